"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
It loads oracle/liboracle.so (built by oracle/Makefile) and exposes numpy-level helpers
that restate the reference raster3d path (see vdb_oracle.cpp for per-function citations).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Stats(ctypes.Structure):
    _fields_ = [
        ("fragment_count", ctypes.c_uint64),
        ("pair_count", ctypes.c_uint64),
        ("window_count", ctypes.c_uint64),
        ("grid_size", ctypes.c_uint64),
        ("upper_count", ctypes.c_uint32),
        ("lower_count", ctypes.c_uint32),
        ("leaf_count", ctypes.c_uint32),
        ("pad", ctypes.c_uint32),
        ("value_count", ctypes.c_uint64),
        ("bbox_min", ctypes.c_int32 * 3),
        ("bbox_max", ctypes.c_int32 * 3),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "bbox_min", "bbox_max")}
        d["bbox_min"] = tuple(self.bbox_min)
        d["bbox_max"] = tuple(self.bbox_max)
        return d


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        u64, u32, i32, f32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
        L.vdbo_preprocess.argtypes = [u64, P, P, P, P, P, P]
        L.vdbo_preprocess.restype = None
        L.vdbo_prim_aabb.argtypes = [u64, P, P, P, P, f32, P, P]
        L.vdbo_prim_aabb.restype = None
        L.vdbo_pairs.argtypes = [u64, P, P, P, P, P, f32, f32, P, P, P, u64, i32]
        L.vdbo_pairs.restype = u64
        L.vdbo_raster_node2.argtypes = [u64, P, P, P, P, P, f32, f32, u32, u32, u32, i32, P, u64, i32,
                                        ctypes.POINTER(Stats)]
        L.vdbo_raster_node2.restype = u64
        L.vdbo_density_f64.argtypes = [u64, P, P, P, P, P, P, f32, P, P]
        L.vdbo_density_f64.restype = u64
        L.vdbo_density_grad_f64.argtypes = [u64, u64, P, P, P, P, P, P, f32, P, P, P, P, P]
        L.vdbo_density_grad_f64.restype = None
        L.vdbo_node2_active_voxels.argtypes = [P, u64, P, P, u64]
        L.vdbo_node2_active_voxels.restype = u64
        L.vdbo_node2_to_onindex.argtypes = [P, u64, P, u64]
        L.vdbo_tile_bin.argtypes = [u32, P, P, P, u32, u32, u32, P, P, P, u64, P]
        L.vdbo_tile_bin.restype = u64
        L.vdbo_voxelbvh_ijkl.argtypes = [u64, P, P, P, P, u32, P, P, P, P, P]
        L.vdbo_voxelbvh_ijkl.restype = i32
        L.vdbo_mat_inverse.argtypes = [P, P]
        L.vdbo_node2_to_onindex.restype = u64
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def preprocess(means, opacity_logits, quats, log_scales, sh0):
    """RasterUtils.cpp:31-136. Returns (means, opacities, quats, scales, colors) as new arrays."""
    means = _f32(means).copy()
    op = _f32(opacity_logits).copy()
    q = _f32(quats).copy()
    s = _f32(log_scales).copy()
    sh0 = _f32(sh0)
    n = means.size // 3
    colors = np.zeros(3 * n, np.float32)
    lib().vdbo_preprocess(n, _p(means), _p(op), _p(q), _p(s), _p(sh0), _p(colors))
    return means, op, q, s, colors


def prim_aabb(means, quats, scales, opacities, voxel_size):
    n = _f32(means).size // 3
    bbox = np.zeros(6 * n, np.int32)
    fp = np.zeros(n, np.uint64)
    lib().vdbo_prim_aabb(n, _p(_f32(means)), _p(_f32(quats)), _p(_f32(scales)), _p(_f32(opacities)),
                         voxel_size, _p(bbox), _p(fp))
    return bbox.reshape(n, 6), fp


def pairs(means, quats, scales, colors, opacities, voxel_size, cull_scale=0.05, nthreads=0, want_rgba=True):
    """Surviving pairs in reference fragment order: (gid[P], ijk[P,3], rgba[P])."""
    means, quats, scales, opacities = map(_f32, (means, quats, scales, opacities))
    colors = _f32(colors) if colors is not None else None
    n = means.size // 3
    L = lib()
    total = L.vdbo_pairs(n, _p(means), _p(quats), _p(scales), _p(colors), _p(opacities), voxel_size, cull_scale,
                         None, None, None, 0, nthreads)
    gid = np.zeros(total, np.uint32)
    ijk = np.zeros(3 * total, np.int32)
    rgba = np.zeros(total, np.uint32) if (want_rgba and colors is not None) else None
    L.vdbo_pairs(n, _p(means), _p(quats), _p(scales), _p(colors), _p(opacities), voxel_size, cull_scale,
                 _p(gid), _p(ijk), _p(rgba), total, nthreads)
    return gid, ijk.reshape(total, 3), rgba


def raster_node2(means, quats, scales, colors, opacities, voxel_size, cull_scale=0.05, prim_batch_size=0,
                 ijk_batch_size=0, max_node_count=0, value_mode=0, nthreads=0):
    """Windowed reference build. Returns (grid bytes as np.uint8, stats dict)."""
    means, quats, scales, opacities = map(_f32, (means, quats, scales, opacities))
    colors = _f32(colors) if colors is not None else np.zeros_like(means)
    n = means.size // 3
    L = lib()
    st = Stats()
    size = L.vdbo_raster_node2(n, _p(means), _p(quats), _p(scales), _p(colors), _p(opacities), voxel_size,
                               cull_scale, prim_batch_size, ijk_batch_size, max_node_count, value_mode, None, 0,
                               nthreads, ctypes.byref(st))
    if size == 0:
        raise RuntimeError("oracle: node capacity exceeded")
    out = np.zeros(size, np.uint8)
    L.vdbo_raster_node2(n, _p(means), _p(quats), _p(scales), _p(colors), _p(opacities), voxel_size, cull_scale,
                        prim_batch_size, ijk_batch_size, max_node_count, value_mode, _p(out), size, nthreads,
                        ctypes.byref(st))
    return out, st.as_dict()


def density_f64(gid, ijk, means, quats, scales, opacities, voxel_size):
    gid = np.ascontiguousarray(gid, np.uint32)
    ijk = np.ascontiguousarray(ijk, np.int32)
    p = gid.size
    vox = np.zeros(3 * max(p, 1), np.int32)
    dens = np.zeros(max(p, 1), np.float64)
    nv = lib().vdbo_density_f64(p, _p(gid), _p(ijk), _p(_f32(means)), _p(_f32(quats)), _p(_f32(scales)),
                                _p(_f32(opacities)), voxel_size, _p(vox), _p(dens))
    return vox[: 3 * nv].reshape(nv, 3), dens[:nv]


def density_grad_f64(gid, ijk, means, log_scales, quats_raw, opacity_logits, voxel_size, residual_per_pair):
    means = _f32(means)
    n = means.size // 3
    gid = np.ascontiguousarray(gid, np.uint32)
    ijk = np.ascontiguousarray(ijk, np.int32)
    res = np.ascontiguousarray(residual_per_pair, np.float64)
    gm = np.zeros(3 * n)
    gs = np.zeros(3 * n)
    gq = np.zeros(4 * n)
    go = np.zeros(n)
    lib().vdbo_density_grad_f64(n, gid.size, _p(gid), _p(ijk), _p(means), _p(_f32(log_scales)),
                                _p(_f32(quats_raw)), _p(_f32(opacity_logits)), voxel_size, _p(res), _p(gm),
                                _p(gs), _p(gq), _p(go))
    return gm.reshape(n, 3), gs.reshape(n, 3), gq.reshape(n, 4), go


def node2_active_voxels(grid):
    grid = np.ascontiguousarray(grid, np.uint8)
    L = lib()
    cnt = L.vdbo_node2_active_voxels(_p(grid), grid.size, None, None, 0)
    ijk = np.zeros(3 * max(cnt, 1), np.int32)
    vidx = np.zeros(max(cnt, 1), np.uint64)
    L.vdbo_node2_active_voxels(_p(grid), grid.size, _p(ijk), _p(vidx), cnt)
    return ijk[: 3 * cnt].reshape(cnt, 3), vidx[:cnt]


def node2_to_onindex(grid):
    """putil/Convert.h:40-313: node2 grid bytes -> NanoVDB OnIndex grid bytes."""
    grid = np.ascontiguousarray(grid, np.uint8)
    L = lib()
    size = L.vdbo_node2_to_onindex(_p(grid), grid.size, None, 0)
    out = np.zeros(size, np.uint8)
    L.vdbo_node2_to_onindex(_p(grid), grid.size, _p(out), size)
    return out


def tile_bin(means2d, radii, depths, width, height, tile_size):
    """Raster2D tile binning (Raster2D.cpp:347-535): sorted (keys_low = depth bits, keys_high =
    tile index, vals = Gaussian index) and the tile offsets."""
    means2d = np.ascontiguousarray(means2d, np.float32)
    radii = np.ascontiguousarray(radii, np.int32)
    depths = np.ascontiguousarray(depths, np.float32)
    n = radii.size
    tiles = ((width + tile_size - 1) // tile_size) * ((height + tile_size - 1) // tile_size)
    L = lib()
    offsets = np.zeros(tiles, np.uint32)
    total = L.vdbo_tile_bin(n, _p(means2d), _p(radii), _p(depths), width, height, tile_size, None, None, None, 0,
                            None)
    lo = np.zeros(total, np.uint32)
    hi = np.zeros(total, np.uint32)
    va = np.zeros(total, np.uint32)
    L.vdbo_tile_bin(n, _p(means2d), _p(radii), _p(depths), width, height, tile_size, _p(lo), _p(hi), _p(va), total,
                    _p(offsets))
    return lo, hi, va, offsets


def voxelbvh_ijkl(means, opacities, quats, scales, resolution, transform=None):
    """VoxelBVH Gaussian keys (VoxelBVH.cpp:1225-1435) on preprocessed arrays: sorted keys [8N]
    (uint64), Gaussian ids [8N], run headers [8N, 2] and the world box [6]."""
    means, opacities, quats, scales = (np.ascontiguousarray(a, np.float32) for a in (means, opacities, quats, scales))
    n = opacities.size
    keys = np.zeros(8 * n, np.uint64)
    ids = np.zeros(8 * n, np.uint32)
    ranges = np.zeros(16 * n, np.uint32)
    bbox = np.zeros(6, np.float32)
    xf = None if transform is None else np.ascontiguousarray(transform, np.float32).ravel()
    rc = lib().vdbo_voxelbvh_ijkl(n, _p(means), _p(opacities), _p(quats), _p(scales), resolution, _p(xf), _p(keys),
                                  _p(ids), _p(ranges), _p(bbox))
    if rc != 0:
        raise ValueError("resolution must be in [1, 4096]")
    return keys, ids, ranges.reshape(-1, 2), bbox


def mat_inverse(m):
    """pnanovdb_camera_mat_inverse (putil/Camera.h:316-375) in float32."""
    m = np.ascontiguousarray(m, np.float32).ravel()
    out = np.zeros(16, np.float32)
    lib().vdbo_mat_inverse(_p(m), _p(out))
    return out.reshape(4, 4)
