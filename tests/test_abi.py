"""C-ABI checks that need no GPU: the library loads, exports every entry point the public
header declares, its function tables and parameter structs have the header's layout (compiled
with the host C compiler), and the host-only compute-array entry points behave like
Compute.cpp:707-760."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from nanovdb_editor_amd import _lib
from nanovdb_editor_amd.compute import Compute, pnanovdb_Compute, pnanovdb_ComputeArray
from nanovdb_editor_amd.raster import pnanovdb_Raster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        names += re.findall(r"^PNANOVDB_API\s+[^;(]*?\b(\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pnanovdb_get_raster" in names and "vdbr_raster_node2" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signatures cover exactly the declared set
    assert set(_lib.SIGNATURES) == set(declared_functions())


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "pnanovdb_amd.h"
#define S(t) printf(#t " %zu\n", sizeof(t));
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f));
int main(void) {
  S(pnanovdb_compute_array_t) O(pnanovdb_compute_array_t, element_size) O(pnanovdb_compute_array_t, filepath)
  S(pnanovdb_compute_t) O(pnanovdb_compute_t, device_interface) O(pnanovdb_compute_t, create_array)
  O(pnanovdb_compute_t, destroy_array) O(pnanovdb_compute_t, map_array) O(pnanovdb_compute_t, duplicate_array)
  S(pnanovdb_raster_t) O(pnanovdb_raster_t, raster_gaussian_3d) O(pnanovdb_raster_t, raster_to_nanovdb)
  O(pnanovdb_raster_t, raster_to_nanovdb_from_arrays) O(pnanovdb_raster_t, create_gaussian_data_from_arrays)
  O(pnanovdb_raster_t, raster_file)
  S(pnanovdb_fileformat_t) O(pnanovdb_fileformat_t, load_file) O(pnanovdb_fileformat_t, module)
  S(vdbr_batch_item_t) O(vdbr_batch_item_t, stats) S(vdbr_tune_result_t)
  S(pnanovdb_parallel_primitives_t) O(pnanovdb_parallel_primitives_t, radix_sort_key64)
  O(pnanovdb_parallel_primitives_t, compute)
  S(vdbr_gaussians_t) O(vdbr_gaussians_t, count)
  S(vdbr_raster_params_t) O(vdbr_raster_params_t, max_node_count)
  S(vdbr_raster_stats_t) O(vdbr_raster_stats_t, key_bits) O(vdbr_raster_stats_t, ms_total)
  O(vdbr_raster_stats_t, ms_reduce_kernel) O(vdbr_raster_stats_t, sort_bit_begin)
  S(vdbr_buffer_t)
  return 0;
}
"""


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("abi")
    src = d / "probe.c"
    src.write_text(PROBE)
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_struct_layouts_match_ctypes(c_layout):
    from nanovdb_editor_amd._lib import VdbrBuffer, VdbrGaussians, VdbrRasterParams, VdbrRasterStats
    from nanovdb_editor_amd._lib import PnanovdbParallelPrimitives

    def off(t, f):
        return getattr(t, f).offset

    c = c_layout
    assert c["pnanovdb_compute_array_t"] == ctypes.sizeof(pnanovdb_ComputeArray)
    assert c["pnanovdb_compute_array_t.filepath"] == off(pnanovdb_ComputeArray, "filepath")
    assert c["pnanovdb_compute_t"] == ctypes.sizeof(pnanovdb_Compute)
    for f in ("device_interface", "create_array", "destroy_array", "map_array", "duplicate_array"):
        assert c[f"pnanovdb_compute_t.{f}"] == off(pnanovdb_Compute, f), f
    assert c["pnanovdb_raster_t"] == ctypes.sizeof(pnanovdb_Raster)
    for f in ("raster_gaussian_3d", "raster_to_nanovdb", "raster_to_nanovdb_from_arrays",
              "create_gaussian_data_from_arrays", "raster_file"):
        assert c[f"pnanovdb_raster_t.{f}"] == off(pnanovdb_Raster, f), f
    assert c["pnanovdb_parallel_primitives_t"] == ctypes.sizeof(PnanovdbParallelPrimitives)
    assert c["pnanovdb_parallel_primitives_t.radix_sort_key64"] == off(PnanovdbParallelPrimitives, "radix_sort_key64")
    assert c["vdbr_gaussians_t"] == ctypes.sizeof(VdbrGaussians)
    assert c["vdbr_raster_params_t"] == ctypes.sizeof(VdbrRasterParams)
    assert c["vdbr_raster_stats_t"] == ctypes.sizeof(VdbrRasterStats)
    assert c["vdbr_raster_stats_t.key_bits"] == off(VdbrRasterStats, "key_bits")
    assert c["vdbr_raster_stats_t.ms_reduce_kernel"] == off(VdbrRasterStats, "ms_reduce_kernel")
    assert c["vdbr_buffer_t"] == ctypes.sizeof(VdbrBuffer)
    assert c["vdbr_raster_stats_t.sort_bit_begin"] == off(VdbrRasterStats, "sort_bit_begin")
    from nanovdb_editor_amd._lib import VdbrBatchItem, VdbrTuneResult
    from nanovdb_editor_amd.fileformat import pnanovdb_FileFormat
    assert c["pnanovdb_fileformat_t"] == ctypes.sizeof(pnanovdb_FileFormat)
    assert c["pnanovdb_fileformat_t.load_file"] == off(pnanovdb_FileFormat, "load_file")
    assert c["pnanovdb_fileformat_t.module"] == off(pnanovdb_FileFormat, "module")
    assert c["vdbr_batch_item_t"] == ctypes.sizeof(VdbrBatchItem)
    assert c["vdbr_batch_item_t.stats"] == off(VdbrBatchItem, "stats")
    assert c["vdbr_tune_result_t"] == ctypes.sizeof(VdbrTuneResult)


def test_function_tables_populated():
    L = _lib.lib()
    raster = ctypes.cast(L.pnanovdb_get_raster(), ctypes.POINTER(pnanovdb_Raster)).contents
    for f in ("create_context", "destroy_context", "create_gaussian_data", "upload_gaussian_data",
              "destroy_gaussian_data", "raster_gaussian_3d", "raster_to_nanovdb", "raster_to_nanovdb_from_arrays",
              "create_gaussian_data_from_arrays", "raster_file"):
        assert ctypes.cast(getattr(raster, f), ctypes.c_void_p).value, f
    assert not raster.raster_gaussian_2d and not raster.create_gaussian_data_from_desc
    comp = ctypes.cast(L.pnanovdb_get_compute(), ctypes.POINTER(pnanovdb_Compute)).contents
    for f in ("create_array", "destroy_array", "map_array", "unmap_array", "duplicate_array"):
        assert ctypes.cast(getattr(comp, f), ctypes.c_void_p).value, f
    from nanovdb_editor_amd._lib import PnanovdbParallelPrimitives
    pp = ctypes.cast(L.pnanovdb_get_parallel_primitives(), ctypes.POINTER(PnanovdbParallelPrimitives)).contents
    for f, _ in PnanovdbParallelPrimitives._fields_[1:]:
        assert pp.__getattribute__(f), f
    # the getters hand out the same writable static each time
    assert L.pnanovdb_get_raster() == L.pnanovdb_get_raster()


def test_compute_arrays_host_semantics():
    c = Compute()
    data = np.arange(10, dtype=np.float32)
    arr = c.create_array(data)
    assert arr.element_size == 4 and arr.element_count == 10
    assert np.array_equal(c.map_array(arr, np.float32), data)
    c.unmap_array(arr)
    dup = c.duplicate_array(arr)
    assert np.array_equal(c.map_array(dup, np.float32), data)
    c.unmap_array(dup)
    # create_array without data gives zero-filled storage (Compute.cpp:707-733)
    table = c.get_compute().contents
    empty = table.create_array(4, 5, None)
    assert np.all(c.map_array(empty.contents, np.uint32) == 0)
    for a in (arr, dup, empty.contents):
        c.destroy_array(a)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU failure path")
def test_no_device_fails_loudly():
    from nanovdb_editor_amd.device import Voxelizer
    with pytest.raises(RuntimeError):
        Voxelizer(0)
