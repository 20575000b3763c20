"""Device-resident entry points on torch tensors (HIP path of libvdbraster).

``Voxelizer.raster_node2`` runs the whole raster3d pipeline on Gaussian arrays that already
live in GPU memory and returns the node2 NanoVDB grid as a uint8 tensor on the same device.
"""
import ctypes

import torch

from ._lib import (VdbrBatchItem, VdbrGaussianGrads, VdbrGaussians, VdbrRasterParams, VdbrRasterStats,
                   VdbrTuneResult, lib)

VALUE_RGBA8 = 0
VALUE_DENSITY = 1
FLAG_RAW_INPUTS = 1
FLAG_SPATIAL_ORDER = 2
FLAG_VOXEL_COORDS = 4
FLAG_WIDE_SEGMENTS = 8


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p()


def _check_f32(name, t, numel):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 GPU tensor")
    if t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")


class Voxelizer:
    """One library context (HIP device, stream, reusable workspace)."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        self._ctx = self._lib.vdbr_context_create(int(device))
        if not self._ctx:
            raise RuntimeError(f"vdbr_context_create({device}) failed (no HIP device?)")
        self.device = torch.device("cuda", int(device))
        self.last_backward_used_pairs = False
        self.last_pair_list_id = 0  # stats["pair_list_id"] of the last raster_node2 call
        self.last_batch_graph_clouds = 0  # clouds of the last batched call run inside its graph
        self.last_batch_timing = {}  # host-clock phases of the last batched call

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.vdbr_context_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, what):
        msg = self._lib.vdbr_last_error(self._ctx)
        raise RuntimeError(f"{what}: {msg.decode() if msg else 'unknown error'}")

    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device)
        self._lib.vdbr_context_set_stream(self._ctx, ctypes.c_void_p(s.cuda_stream))

    def raster_node2(self, means, quats, scales, colors, opacities, voxel_size, *, value_mode=VALUE_RGBA8,
                     cull_scale=0.0, raw=False, sh0=None, prim_batch_size=0, ijk_batch_size=0, max_node_count=0,
                     spatial_order=False, copy_out=True, voxel_coords=False, slab_pairs=0, wide_segments=False):
        """Gaussians (float32 GPU tensors, reference layouts) -> node2 grid.

        With ``raw=True`` the arrays hold opacity logits, unnormalised (w,x,y,z) quaternions and
        log-scales and are preprocessed in place (colors receives 0.28209479*sh0 + 0.5 from
        ``sh0``). With ``voxel_coords=True`` the call also records every value slot's voxel
        (``last_voxels``). ``wide_segments=True`` keeps the emission's candidate segments in 16-byte
        records even where 8-byte ones fit (same grid). Returns ``(grid_uint8_tensor or None, stats_dict)``.
        """
        n = means.numel() // 3
        _check_f32("means", means, 3 * n)
        _check_f32("quats", quats, 4 * n)
        _check_f32("scales", scales, 3 * n)
        _check_f32("colors", colors, 3 * n)
        _check_f32("opacities", opacities, n)
        if raw:
            if sh0 is None:
                raise ValueError("raw inputs need sh0")
            _check_f32("sh0", sh0, 3 * n)
        g = VdbrGaussians(_ptr(means), _ptr(opacities), _ptr(quats), _ptr(scales), _ptr(colors),
                          _ptr(sh0), n)
        p = VdbrRasterParams(float(voxel_size), float(cull_scale), int(value_mode),
                             (FLAG_RAW_INPUTS if raw else 0) | (FLAG_SPATIAL_ORDER if spatial_order else 0)
                             | (FLAG_VOXEL_COORDS if voxel_coords else 0) | (FLAG_WIDE_SEGMENTS if wide_segments else 0),
                             int(prim_batch_size), int(ijk_batch_size), int(max_node_count), int(slab_pairs))
        grid_dev = ctypes.c_void_p()
        size = ctypes.c_uint64()
        st = VdbrRasterStats()
        self._bind_stream()
        rc = self._lib.vdbr_raster_node2(self._ctx, ctypes.byref(g), ctypes.byref(p), ctypes.byref(grid_dev),
                                         ctypes.byref(size), ctypes.byref(st))
        if rc != 0:
            self._err("vdbr_raster_node2")
        self.last_pair_list_id = st.pair_list_id
        out = None
        if copy_out:
            out = torch.empty(size.value, dtype=torch.uint8, device=self.device)
            if size.value and self._lib.vdbr_copy(self._ctx, _ptr(out), grid_dev, size.value, 2) != 0:
                self._err("vdbr_copy")
        return out, st.as_dict()

    def raster_node2_batch(self, clouds, *, workers=4, copy_out=True):
        """Many clouds in one call. ``clouds`` is a sequence of dicts with the tensors of
        ``raster_node2`` (means, quats, scales, colors, opacities, and sh0 when raw) plus
        ``voxel_size`` and optional ``raw``, ``value_mode``, ``cull_scale``. Each result equals a
        ``raster_node2`` call on that cloud alone. After the setup of every cloud, the stages of
        all clouds run as one captured HIP graph over ``workers`` streams;
        ``last_batch_graph_clouds`` counts the clouds that ran inside it. Returns
        ``(grids, stats)``: uint8 GPU tensors (or None without ``copy_out``) and stats dicts, in
        input order."""
        items = (VdbrBatchItem * max(len(clouds), 1))()
        keep = []
        for i, c in enumerate(clouds):
            n = c["means"].numel() // 3
            raw = bool(c.get("raw", False))
            for name, k in (("means", 3), ("quats", 4), ("scales", 3), ("colors", 3), ("opacities", 1)):
                _check_f32(f"clouds[{i}].{name}", c[name], k * n)
            if raw:
                _check_f32(f"clouds[{i}].sh0", c["sh0"], 3 * n)
            items[i].gaussians = VdbrGaussians(_ptr(c["means"]), _ptr(c["opacities"]), _ptr(c["quats"]),
                                               _ptr(c["scales"]), _ptr(c["colors"]), _ptr(c.get("sh0")), n)
            items[i].params = VdbrRasterParams(float(c["voxel_size"]), float(c.get("cull_scale", 0.0)),
                                               int(c.get("value_mode", VALUE_RGBA8)),
                                               FLAG_RAW_INPUTS if raw else 0, 0, 0, 0, 0)
            keep.append(c)
        self._bind_stream()
        if self._lib.vdbr_raster_node2_batch(self._ctx, items, len(clouds), int(workers)) != 0:
            self._err("vdbr_raster_node2_batch")
        self.last_batch_graph_clouds = int(self._lib.vdbr_last_batch_graph_clouds(self._ctx))
        ms = (ctypes.c_float * 5)()
        self._lib.vdbr_last_batch_timing(self._ctx, ms)
        self.last_batch_timing = {"setup_ms": ms[0], "capture_ms": ms[1], "graph_ms": ms[2],
                                  "captured": bool(ms[3]), "rerun_ms": ms[4]}
        grids, stats = [], []
        for i in range(len(clouds)):
            stats.append(items[i].stats.as_dict())
            if not copy_out:
                grids.append(None)
                continue
            size = items[i].grid_size
            out = torch.empty(size, dtype=torch.uint8, device=self.device)
            if size and self._lib.vdbr_copy(self._ctx, _ptr(out), ctypes.c_void_p(items[i].grid_dev), size, 2) != 0:
                self._err("vdbr_copy")
            grids.append(out)
        return grids, stats

    def density_backward(self, means, log_scales, quats, opacity_logits, voxel_size, grid, grad_values,
                         pair_list_id=0):
        """Gradients (means, log_scales, quats, opacity_logits) of sum(grad_values * values) where
        values is the value array of the density grid ``grid`` that ``raster_node2`` built from these
        raw parameters (``raw=True``, ``value_mode=VALUE_DENSITY``).

        With the forward's ``stats["pair_list_id"]`` the adjoint runs over the sorted pair list that
        the forward kept on this context (``vdbr_density_backward_pairs``); when that list is gone it
        enumerates the pairs again. ``last_backward_used_pairs`` tells which path ran."""
        n = means.numel() // 3
        _check_f32("means", means, 3 * n)
        _check_f32("log_scales", log_scales, 3 * n)
        _check_f32("quats", quats, 4 * n)
        _check_f32("opacity_logits", opacity_logits, n)
        if grid.dtype != torch.uint8 or not grid.is_cuda or not grid.is_contiguous():
            raise ValueError("grid must be a contiguous uint8 GPU tensor")
        if grad_values.dtype != torch.float32 or not grad_values.is_cuda or not grad_values.is_contiguous():
            raise ValueError("grad_values must be a contiguous float32 GPU tensor")
        gm, gs, gq, go = (torch.empty_like(t) for t in (means, log_scales, quats, opacity_logits))
        g = VdbrGaussians(_ptr(means), _ptr(opacity_logits), _ptr(quats), _ptr(log_scales), None, None, n)
        out = VdbrGaussianGrads(_ptr(gm), _ptr(go), _ptr(gq), _ptr(gs))
        self._bind_stream()
        if pair_list_id:
            rc = self._lib.vdbr_density_backward_pairs(self._ctx, ctypes.byref(g), float(voxel_size), _ptr(grid),
                                                       grid.numel(), _ptr(grad_values), grad_values.numel(),
                                                       ctypes.byref(out), int(pair_list_id))
            if rc < 0:
                self._err("vdbr_density_backward_pairs")
            self.last_backward_used_pairs = rc == 0
        else:
            if self._lib.vdbr_density_backward(self._ctx, ctypes.byref(g), float(voxel_size), _ptr(grid),
                                               grid.numel(), _ptr(grad_values), grad_values.numel(),
                                               ctypes.byref(out)) != 0:
                self._err("vdbr_density_backward")
            self.last_backward_used_pairs = False
        return gm, gs, gq, go

    def node2_to_onindex(self, grid):
        """NanoVDB OnIndex grid (uint8 GPU tensor, zero padded to cover the header's grid size) from a
        node2 grid held in a uint8 GPU tensor (putil/Convert.h:40-313)."""
        if grid.dtype != torch.uint8 or not grid.is_cuda or not grid.is_contiguous():
            raise ValueError("grid must be a contiguous uint8 GPU tensor")
        self._bind_stream()
        dev = ctypes.c_void_p()
        size = ctypes.c_uint64()
        buf = ctypes.c_uint64()
        if self._lib.vdbr_node2_to_onindex(self._ctx, _ptr(grid), grid.numel(), ctypes.byref(dev), ctypes.byref(size),
                                           ctypes.byref(buf)) != 0:
            self._err("vdbr_node2_to_onindex")
        out = torch.empty(buf.value, dtype=torch.uint8, device=self.device)
        if buf.value and self._lib.vdbr_copy(self._ctx, _ptr(out), dev, buf.value, 2) != 0:
            self._err("vdbr_copy")
        return out, size.value

    def selftest_alpha8(self, count, seed=1):
        """(mismatches, fast_decided) of the fast RGBA8 alpha against the correctly rounded chain."""
        mism = ctypes.c_uint64()
        fast = ctypes.c_uint64()
        self._bind_stream()
        if self._lib.vdbr_selftest_alpha8(self._ctx, int(count), int(seed), ctypes.byref(mism), ctypes.byref(fast)) != 0:
            self._err("vdbr_selftest_alpha8")
        return mism.value, fast.value

    def autotune_sort(self, key_count, key_bits):
        """Times every onesweep tile shape x digit width on key_count hashed keys and keeps the
        fastest for sorts of that size or larger. Returns (choice dict, list of measurement dicts)."""
        cap = 64
        res = (VdbrTuneResult * cap)()
        cnt, bv, bw = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        self._bind_stream()
        if self._lib.vdbr_autotune_sort(self._ctx, int(key_count), int(key_bits), res, cap, ctypes.byref(cnt),
                                        ctypes.byref(bv), ctypes.byref(bw)) != 0:
            self._err("vdbr_autotune_sort")
        rows = [{k: getattr(res[i], k) for k, _ in VdbrTuneResult._fields_} for i in range(min(cnt.value, cap))]
        best = next(r for r in rows if r["variant"] == bv.value and r["digit_bits"] == bw.value)
        return best, rows

    def set_sort_config(self, from_count, variant, digit_bits):
        if self._lib.vdbr_set_sort_config(self._ctx, int(from_count), int(variant), int(digit_bits)) != 0:
            self._err("vdbr_set_sort_config")

    def last_voxels(self):
        """int32 [value_count, 3] voxel coordinates of the last ``raster_node2(voxel_coords=True)``
        call in value-index order (rows 0 and 1, the background slots, hold sentinels)."""
        cnt = ctypes.c_uint64()
        self._lib.vdbr_last_voxels(self._ctx, None, 0, ctypes.byref(cnt))
        out = torch.empty((cnt.value, 3), dtype=torch.int32, device=self.device)
        if cnt.value and self._lib.vdbr_last_voxels(self._ctx, _ptr(out), cnt.value, ctypes.byref(cnt)) != 0:
            self._err("vdbr_last_voxels")
        return out

    def node2_sample(self, grid, values, ijk):
        """Value words (same dtype as ``values``, 32-bit) of a node2 grid at int32 voxel coordinates
        [M, 3]; 0 where the voxel is inactive."""
        if grid.dtype != torch.uint8 or not grid.is_cuda or not grid.is_contiguous():
            raise ValueError("grid must be a contiguous uint8 GPU tensor")
        if values.element_size() != 4 or not values.is_cuda or not values.is_contiguous():
            raise ValueError("values must be a contiguous 32-bit GPU tensor")
        if ijk.dtype != torch.int32 or not ijk.is_cuda or not ijk.is_contiguous() or ijk.numel() % 3:
            raise ValueError("ijk must be a contiguous int32 GPU tensor of shape [M, 3]")
        out = torch.empty(ijk.numel() // 3, dtype=values.dtype, device=self.device)
        self._bind_stream()
        if self._lib.vdbr_node2_sample(self._ctx, _ptr(grid), grid.numel(), _ptr(values), values.numel(), _ptr(ijk),
                                       ijk.numel() // 3, _ptr(out)) != 0:
            self._err("vdbr_node2_sample")
        return out

    def last_pairs(self):
        """Sorted pair list of the last raster call: (ijk int32 [P,3], gid int32 [P]) on the GPU."""
        cnt = ctypes.c_uint64()
        self._lib.vdbr_last_pairs(self._ctx, None, None, 0, ctypes.byref(cnt))
        p = cnt.value
        ijk = torch.empty((p, 3), dtype=torch.int32, device=self.device)
        gid = torch.empty(p, dtype=torch.int32, device=self.device)
        if p and self._lib.vdbr_last_pairs(self._ctx, _ptr(ijk), _ptr(gid), p, ctypes.byref(cnt)) != 0:
            self._err("vdbr_last_pairs")
        return ijk, gid

    def radix_sort_u64(self, keys, vals=None, bit_begin=0, bit_end=64):
        """Stable in-place sort of int64 keys (as unsigned) over [bit_begin, bit_end)."""
        if keys.dtype != torch.int64 or not keys.is_cuda or not keys.is_contiguous():
            raise ValueError("keys must be a contiguous int64 GPU tensor")
        if vals is not None and (vals.dtype != torch.int32 or vals.numel() != keys.numel()):
            raise ValueError("vals must be an int32 tensor with one value per key")
        self._bind_stream()
        if self._lib.vdbr_radix_sort_u64(self._ctx, _ptr(keys), _ptr(vals), keys.numel(), bit_begin, bit_end) != 0:
            self._err("vdbr_radix_sort_u64")

    def radix_sort_u32(self, keys, vals=None, key_bit_count=32):
        if keys.dtype != torch.int32 or not keys.is_cuda or not keys.is_contiguous():
            raise ValueError("keys must be a contiguous int32 GPU tensor")
        self._bind_stream()
        if self._lib.vdbr_radix_sort_u32(self._ctx, _ptr(keys), _ptr(vals), keys.numel(), key_bit_count) != 0:
            self._err("vdbr_radix_sort_u32")

    def voxelbvh_ijkl(self, means, opacities, quats, scales, resolution, transform=None):
        """VoxelBVH Gaussian keys (raster/VoxelBVH.cpp:1225-1435) of preprocessed float32 GPU
        arrays (opacities after the sigmoid, normalised (w,x,y,z) quats, scales after exp).
        Returns (keys int64 [8N] sorted, Gaussian ids int32 [8N], run headers int32 [8N, 2],
        world box float32 [6]); unused key slots hold -1 (all bits set)."""
        n = opacities.numel()
        _check_f32("means", means, 3 * n)
        _check_f32("opacities", opacities, n)
        _check_f32("quats", quats, 4 * n)
        _check_f32("scales", scales, 3 * n)
        keys = torch.empty(8 * n, dtype=torch.int64, device=self.device)
        ids = torch.empty(8 * n, dtype=torch.int32, device=self.device)
        ranges = torch.empty((8 * n, 2), dtype=torch.int32, device=self.device)
        bbox = torch.empty(6, dtype=torch.float32, device=self.device)
        xf = None
        if transform is not None:
            xf = (ctypes.c_float * 16)(*[float(v) for v in torch.as_tensor(transform).reshape(-1).tolist()])
        g = VdbrGaussians(_ptr(means), _ptr(opacities), _ptr(quats), _ptr(scales), None, None, n)
        self._bind_stream()
        if self._lib.vdbr_voxelbvh_ijkl(self._ctx, ctypes.byref(g), int(resolution), xf, _ptr(keys), _ptr(ids),
                                        _ptr(ranges), _ptr(bbox)) != 0:
            self._err("vdbr_voxelbvh_ijkl")
        return keys, ids, ranges, bbox

    def tile_bin(self, means2d, radii, depths, width, height, tile_size=16):
        """Raster2D tile binning (raster/Raster2D.cpp:347-535) of projected Gaussians: float32
        means2d [N, 2] pixels, int32 radii [N], float32 depths [N] on the GPU. Returns int32 tensors
        (keys_low = depth bits, keys_high = tile index, vals = Gaussian index) sorted by (tile,
        depth), and the int32 tile offsets [tiles_h * tiles_w]."""
        n = radii.numel()
        _check_f32("means2d", means2d, 2 * n)
        _check_f32("depths", depths, n)
        if radii.dtype != torch.int32 or not radii.is_cuda or not radii.is_contiguous():
            raise ValueError("radii must be a contiguous int32 GPU tensor")
        tiles = ((width + tile_size - 1) // tile_size) * ((height + tile_size - 1) // tile_size)
        offsets = torch.empty(tiles, dtype=torch.int32, device=self.device)
        lo, hi, va = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        cnt = ctypes.c_uint64()
        self._bind_stream()
        if self._lib.vdbr_tile_bin(self._ctx, _ptr(means2d), _ptr(radii), _ptr(depths), n, int(width), int(height),
                                   int(tile_size), _ptr(offsets), ctypes.byref(lo), ctypes.byref(hi), ctypes.byref(va),
                                   ctypes.byref(cnt)) != 0:
            self._err("vdbr_tile_bin")
        outs = []
        for src in (lo, hi, va):
            t = torch.empty(cnt.value, dtype=torch.int32, device=self.device)
            if cnt.value and self._lib.vdbr_copy(self._ctx, _ptr(t), src, 4 * cnt.value, 2) != 0:
                self._err("vdbr_copy")
            outs.append(t)
        return outs[0], outs[1], outs[2], offsets

    def radix_sort_dual_key(self, key_low, key_high, vals=None, low_bits=32, high_bits=32):
        """Stable in-place sort of int32 tensors by (key_high, key_low) over the given low bits."""
        for name, t in (("key_low", key_low), ("key_high", key_high), ("vals", vals)):
            if t is not None and (t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous()
                                  or t.numel() != key_low.numel()):
                raise ValueError(f"{name} must be a contiguous int32 GPU tensor of the key count")
        self._bind_stream()
        if self._lib.vdbr_radix_sort_dual_key(self._ctx, _ptr(key_low), _ptr(key_high), _ptr(vals), key_low.numel(),
                                              int(low_bits), int(high_bits)) != 0:
            self._err("vdbr_radix_sort_dual_key")

    def inclusive_scan(self, values, op="add"):
        """Inclusive scan: op 'add' (int32 as u32 or int64 as u64) or 'max' (int32 as u32)."""
        if not values.is_cuda or not values.is_contiguous():
            raise ValueError("values must be a contiguous GPU tensor")
        out = torch.empty_like(values)
        if op == "add" and values.dtype == torch.int32:
            code = 0
        elif op == "add" and values.dtype == torch.int64:
            code = 1
        elif op == "max" and values.dtype == torch.int32:
            code = 2
        else:
            raise ValueError("unsupported scan op / dtype")
        self._bind_stream()
        if self._lib.vdbr_inclusive_scan(self._ctx, _ptr(values), _ptr(out), values.numel(), code) != 0:
            self._err("vdbr_inclusive_scan")
        return out

    def trim(self):
        """Frees the context's device workspace (results of earlier calls become invalid)."""
        self._lib.vdbr_context_trim(self._ctx)

    def workspace_bytes(self):
        return int(self._lib.vdbr_workspace_bytes(self._ctx))
