"""The public header against the reference's function tables (tests/golden/abi_fields.json, read
from nanovdb_editor/putil/*.h by tests/golden/make_abi_fields.py): every table and parameter
struct has the reference's fields in the reference's order, data fields of the same types, and
function-pointer slots either implemented with a function type or left as a NULL `void*` slot.
Host-side behaviour of the tables that needs no GPU: the compute and fileformat tables carry the
library's own handle in `module`, errors of the raster table reach the host's log sink through the
compute table's device interface, and short arrays are rejected."""
import ctypes
import json
import os

import numpy as np
import pytest

from abi_fields import struct_fields
from nanovdb_editor_amd import _lib
from nanovdb_editor_amd.compute import Compute, pnanovdb_Compute, pnanovdb_ComputeArray
from nanovdb_editor_amd.raster import Raster

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = json.load(open(os.path.join(ROOT, "tests", "golden", "abi_fields.json")))


@pytest.fixture(scope="module")
def ours():
    return struct_fields(open(os.path.join(ROOT, "include", "pnanovdb_amd.h")).read(), list(FIXTURE))


@pytest.mark.parametrize("name", sorted(FIXTURE))
def test_table_fields_match_reference(ours, name):
    ref = FIXTURE[name]
    assert name in ours, f"{name} is not defined in include/pnanovdb_amd.h"
    got = ours[name]
    assert got["fields"] == ref["fields"], f"{name}: field order differs from {ref['source']}"
    assert not set(got["fields"]) & set(ref["conditional"])
    for field, rt, gt in zip(ref["fields"], ref["types"], got["types"]):
        if rt == "fn":
            assert gt in ("fn", "void*"), f"{name}.{field}: {gt}"
        else:
            assert gt == rt, f"{name}.{field}: {gt} != {rt}"


def test_module_handle_resolves_the_getters():
    L = _lib.lib()
    comp = ctypes.cast(L.pnanovdb_get_compute(), ctypes.POINTER(pnanovdb_Compute)).contents
    assert comp.module == L._handle                  # dlopen of the same file: the same handle
    libdl = ctypes.CDLL(None)
    libdl.dlsym.restype = ctypes.c_void_p
    libdl.dlsym.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    for name in ("pnanovdb_get_raster", "pnanovdb_get_grid_build", "pnanovdb_get_parallel_primitives",
                 "pnanovdb_get_compute", "pnanovdb_get_fileformat"):
        assert libdl.dlsym(comp.module, name.encode()) == ctypes.cast(getattr(L, name), ctypes.c_void_p).value
    ff = ctypes.cast(L.pnanovdb_get_fileformat(), ctypes.POINTER(ctypes.c_void_p * 4)).contents
    assert ff[3] == L._handle


# --------------------------------------------------------------------- log routing (host side)
LOG = CFUNC = None


class _FakeComputeInterface(ctypes.Structure):
    # leading fields of pnanovdb_compute_interface_t (putil/Compute.h:338-349)
    _fields_ = [("iface", ctypes.c_void_p), ("get_api", ctypes.c_void_p), ("is_feature_supported", ctypes.c_void_p),
                ("get_frame_info", ctypes.c_void_p), ("get_log_print", ctypes.c_void_p)]


@pytest.fixture
def logging_compute():
    """A copy of the library's compute table whose device interface hands out a log sink."""
    messages = []
    log_print_t = ctypes.CFUNCTYPE(None, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_char_p)

    def sink(level, fmt, arg):
        messages.append((level, arg.decode() if fmt == b"%s" else fmt.decode()))

    sink_c = log_print_t(sink)
    get_log_print_t = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p)
    get_log_print = get_log_print_t(lambda ctx: ctypes.cast(sink_c, ctypes.c_void_p).value)
    iface = _FakeComputeInterface(None, None, None, None, ctypes.cast(get_log_print, ctypes.c_void_p).value)
    get_iface_t = ctypes.CFUNCTYPE(ctypes.c_void_p, ctypes.c_void_p)
    get_iface = get_iface_t(lambda q: ctypes.addressof(iface))
    get_ctx = get_iface_t(lambda q: 1234)
    base = Compute().get_compute().contents
    comp = pnanovdb_Compute()
    ctypes.memmove(ctypes.addressof(comp), ctypes.addressof(base), ctypes.sizeof(comp))
    comp.device_interface.get_compute_interface = ctypes.cast(get_iface, ctypes.c_void_p).value
    comp.device_interface.get_compute_context = ctypes.cast(get_ctx, ctypes.c_void_p).value
    keep = (sink_c, get_log_print, iface, get_iface, get_ctx)
    yield comp, messages, keep


def _arrays(compute, n, quat_count=None):
    rng = np.random.default_rng(3)
    a = [rng.standard_normal(3 * n), rng.standard_normal(n), rng.standard_normal(quat_count or 4 * n),
         rng.standard_normal(3 * n), rng.standard_normal(3 * n), np.zeros(0)]
    return [compute.create_array(x.astype(np.float32)) for x in a]


def test_errors_reach_the_host_log_sink(logging_compute):
    comp, messages, _ = logging_compute
    r = Raster()
    table = r._raster.contents
    compute = Compute()
    arrays = _arrays(compute, 10, quat_count=4 * 10 - 3)           # one quaternion short
    arr_p = (ctypes.POINTER(pnanovdb_ComputeArray) * 6)(*[ctypes.pointer(a) for a in arrays])
    data = ctypes.c_void_p()
    ctxp = ctypes.c_void_p()
    queue = ctypes.c_void_p(1)
    ok = table.create_gaussian_data_from_arrays(r._raster, ctypes.pointer(comp), queue, arr_p, 6, ctypes.byref(data),
                                                None, ctypes.byref(ctxp))
    assert not ok and not data
    levels = [m[0] for m in messages]
    assert 1 in levels                                              # the reference's deprecation warning
    assert any(lv == 0 and "create_gaussian_data_from_arrays" in txt for lv, txt in messages)
    # create_gaussian_data with a short opacity array
    messages.clear()
    m, q, s, c, sh0, shn = (ctypes.pointer(x) for x in (arrays[0], arrays[2], arrays[3], arrays[4], arrays[4],
                                                          arrays[5]))
    short = compute.create_array(np.zeros(4, np.float32))
    got = table.create_gaussian_data(ctypes.pointer(comp), queue, None, m, q, s, c, sh0, shn, ctypes.pointer(short),
                                     None, None)
    assert not got
    assert messages and messages[0][0] == 0 and "create_gaussian_data" in messages[0][1]
    for a in arrays + [short]:
        compute.destroy_array(a)


def test_short_arrays_rejected_without_a_host_sink():
    """The library's own compute table has no device interface: failures return FALSE / NULL."""
    r = Raster()
    table = r._raster.contents
    compute = Compute()
    for bad in (4 * 10 - 1, 0):
        arrays = _arrays(compute, 10, quat_count=bad or None)
        if bad == 0:
            arrays[1] = compute.create_array(np.zeros(3, np.float32))   # opacities short
        arr_p = (ctypes.POINTER(pnanovdb_ComputeArray) * 6)(*[ctypes.pointer(a) for a in arrays])
        data = ctypes.c_void_p()
        assert not table.create_gaussian_data_from_arrays(r._raster, compute.get_compute(), None, arr_p, 6,
                                                          ctypes.byref(data), None, None)
        out = ctypes.POINTER(pnanovdb_ComputeArray)()
        assert not table.raster_to_nanovdb_from_arrays(r._raster, compute.get_compute(), None, ctypes.c_float(0.01),
                                                       arr_p, 6, ctypes.byref(out))
        # fewer than five arrays, or a NULL one
        assert not table.raster_to_nanovdb_from_arrays(r._raster, compute.get_compute(), None, ctypes.c_float(0.01),
                                                       arr_p, 4, ctypes.byref(out))
        arr_p[3] = ctypes.POINTER(pnanovdb_ComputeArray)()
        assert not table.create_gaussian_data_from_arrays(r._raster, compute.get_compute(), None, arr_p, 6,
                                                          ctypes.byref(data), None, None)
        for a in arrays:
            compute.destroy_array(a)
