"""ctypes binding of libvdbraster.so (the HIP library built from ../csrc).

The library is loaded from this package directory. There is no fallback: when the library
is missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvdbraster.so")
_LIB = None

P = ctypes.c_void_p
u64, u32, i32, f32 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_float


class VdbrGaussians(ctypes.Structure):
    _fields_ = [("means", P), ("opacities", P), ("quats", P), ("scales", P), ("colors", P), ("sh0", P),
                ("count", u64)]


class VdbrGaussianGrads(ctypes.Structure):
    _fields_ = [("means", P), ("opacities", P), ("quats", P), ("scales", P)]


class VdbrTuneResult(ctypes.Structure):
    _fields_ = [("variant", u32), ("threads", u32), ("keys_per_thread", u32), ("digit_bits", u32), ("ms", f32)]


class VdbrRasterParams(ctypes.Structure):
    _fields_ = [("voxel_size", f32), ("cull_scale", f32), ("value_mode", u32), ("flags", u32),
                ("prim_batch_size", u32), ("ijk_batch_size", u32), ("max_node_count", u32), ("slab_pairs", u32)]


class VdbrRasterStats(ctypes.Structure):
    _fields_ = [("gaussian_count", u64), ("fragment_count", u64), ("pair_count", u64), ("voxel_count", u64),
                ("grid_size", u64), ("upper_count", u32), ("lower_count", u32), ("leaf_count", u32),
                ("window_count", u32), ("key_bits", u32), ("emit_attempts", u32), ("ms_setup", f32),
                ("ms_emit", f32), ("ms_sort", f32), ("ms_reduce", f32), ("ms_build", f32), ("ms_total", f32),
                ("sort_passes", u32), ("ms_sort_passes", f32), ("ms_emit_kernel", f32), ("ms_reduce_kernel", f32),
                ("sort_bit_begin", u32), ("slab_count", u32), ("pair_list_id", u64),
                ("ms_row_kernel", f32), ("ms_walk_kernel", f32), ("candidate_voxels", u64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VdbrBatchItem(ctypes.Structure):
    _fields_ = [("gaussians", VdbrGaussians), ("params", VdbrRasterParams), ("grid_dev", P), ("grid_size", u64),
                ("stats", VdbrRasterStats)]


class VdbrBuffer(ctypes.Structure):
    _fields_ = [("device_ptr", P), ("size_bytes", u64)]


class PnanovdbParallelPrimitives(ctypes.Structure):
    """Field order of pnanovdb_parallel_primitives_t (nanovdb_editor/putil/ParallelPrimitives.h:22-88)."""
    _fields_ = [(name, P) for name in (
        "interface_pnanovdb_reflect_data_type", "create_context", "destroy_context", "global_scan",
        "global_scan_uint64", "global_scan_max", "radix_sort", "radix_sort_dual_key", "radix_sort_key64", "compute")]


# exported symbols and their signatures (include/pnanovdb_amd.h)
SIGNATURES = {
    "pnanovdb_get_raster": ([], P),
    "pnanovdb_get_compute": ([], P),
    "pnanovdb_get_parallel_primitives": ([], P),
    "pnanovdb_get_fileformat": ([], P),
    "pnanovdb_get_grid_build": ([], P),
    "vdbr_context_create": ([ctypes.c_int], P),
    "vdbr_context_destroy": ([P], None),
    "vdbr_last_error": ([P], ctypes.c_char_p),
    "vdbr_context_set_stream": ([P, P], ctypes.c_int),
    "vdbr_context_get_stream": ([P], P),
    "vdbr_synchronize": ([P], ctypes.c_int),
    "vdbr_copy": ([P, P, P, u64, ctypes.c_int], ctypes.c_int),
    "vdbr_workspace_bytes": ([P], u64),
    "vdbr_release_workspace": ([P], ctypes.c_int),
    "vdbr_raster_node2": ([P, ctypes.POINTER(VdbrGaussians), ctypes.POINTER(VdbrRasterParams),
                           ctypes.POINTER(P), ctypes.POINTER(u64), ctypes.POINTER(VdbrRasterStats)], ctypes.c_int),
    "vdbr_last_pairs": ([P, P, P, u64, ctypes.POINTER(u64)], ctypes.c_int),
    "vdbr_radix_sort_u64": ([P, P, P, u64, u32, u32], ctypes.c_int),
    "vdbr_radix_sort_u64_noalloc": ([P, P, P, u64, u32, u32], ctypes.c_int),
    "vdbr_density_backward": ([P, ctypes.POINTER(VdbrGaussians), f32, P, u64, P, u64,
                               ctypes.POINTER(VdbrGaussianGrads)], ctypes.c_int),
    "vdbr_density_backward_pairs": ([P, ctypes.POINTER(VdbrGaussians), f32, P, u64, P, u64,
                                     ctypes.POINTER(VdbrGaussianGrads), u64], ctypes.c_int),
    "vdbr_selftest_alpha8": ([P, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)], ctypes.c_int),
    "vdbr_context_trim": ([P], ctypes.c_int),
    "vdbr_last_voxels": ([P, P, u64, ctypes.POINTER(u64)], ctypes.c_int),
    "vdbr_node2_sample": ([P, P, u64, P, u64, P, u64, P], ctypes.c_int),
    "vdbr_voxelbvh_ijkl": ([P, ctypes.POINTER(VdbrGaussians), u32, P, P, P, P, P], ctypes.c_int),
    "vdbr_tile_bin": ([P, P, P, P, u32, u32, u32, u32, P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(P),
                       ctypes.POINTER(u64)], ctypes.c_int),
    "vdbr_radix_sort_dual_key": ([P, P, P, P, u64, u32, u32], ctypes.c_int),
    "vdbr_raster_node2_batch": ([P, ctypes.POINTER(VdbrBatchItem), u32, u32], ctypes.c_int),
    "vdbr_last_batch_graph_clouds": ([P], u32),
    "vdbr_last_batch_timing": ([P, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "vdbr_autotune_sort": ([P, u64, u32, ctypes.POINTER(VdbrTuneResult), u32, ctypes.POINTER(u32),
                            ctypes.POINTER(u32), ctypes.POINTER(u32)], ctypes.c_int),
    "vdbr_set_sort_config": ([P, u64, u32, u32], ctypes.c_int),
    "vdbr_node2_to_onindex": ([P, P, u64, ctypes.POINTER(P), ctypes.POINTER(u64), ctypes.POINTER(u64)],
                              ctypes.c_int),
    "vdbr_radix_sort_u32": ([P, P, P, u64, u32], ctypes.c_int),
    "vdbr_inclusive_scan": ([P, P, P, u64, ctypes.c_int], ctypes.c_int),
}


def lib():
    """Load libvdbraster.so and declare its signatures. Raises when the library is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libvdbraster.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB
