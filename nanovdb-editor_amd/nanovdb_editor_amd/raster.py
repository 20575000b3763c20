"""Raster function table binding (mirror of pymodule/nanovdb_editor/raster.py:14-275).

``pnanovdb_Raster`` has the field order of pnanovdb_raster_t (nanovdb_editor/putil/Raster.h:58-174);
``Raster.raster_to_nanovdb`` and ``Raster.raster_to_nanovdb_from_arrays`` keep the reference
argument order, validation and error behaviour.
"""
from ctypes import CFUNCTYPE, POINTER, Structure, byref, c_char_p, c_float, c_uint32, c_uint64, c_void_p, pointer

from ._lib import lib
from .compute import Compute, pnanovdb_bool_t, pnanovdb_Compute, pnanovdb_ComputeArray
from .exceptions import InvalidArgumentError

_ArrP = POINTER(pnanovdb_ComputeArray)


class pnanovdb_Raster(Structure):
    """Definition equivalent to pnanovdb_raster_t."""

    _fields_ = [
        ("interface_pnanovdb_reflect_data_type", c_void_p),
        ("compute", POINTER(pnanovdb_Compute)),
        ("create_context", CFUNCTYPE(c_void_p, POINTER(pnanovdb_Compute), c_void_p)),
        ("destroy_context", CFUNCTYPE(None, POINTER(pnanovdb_Compute), c_void_p, c_void_p)),
        ("create_gaussian_data", CFUNCTYPE(c_void_p, POINTER(pnanovdb_Compute), c_void_p, c_void_p, _ArrP, _ArrP,
                                           _ArrP, _ArrP, _ArrP, _ArrP, _ArrP, POINTER(_ArrP), c_void_p)),
        ("upload_gaussian_data", CFUNCTYPE(None, POINTER(pnanovdb_Compute), c_void_p, c_void_p, c_void_p)),
        ("destroy_gaussian_data", CFUNCTYPE(None, POINTER(pnanovdb_Compute), c_void_p, c_void_p)),
        ("raster_gaussian_2d", c_void_p),
        ("raster_gaussian_3d", CFUNCTYPE(None, POINTER(pnanovdb_Compute), c_void_p, c_void_p, c_float, c_void_p,
                                         c_void_p, c_uint64, c_void_p)),
        ("raster_to_nanovdb", CFUNCTYPE(_ArrP, POINTER(pnanovdb_Compute), c_void_p, c_float, _ArrP, _ArrP, _ArrP,
                                        _ArrP, _ArrP, _ArrP, _ArrP, POINTER(_ArrP), c_void_p, c_void_p)),
        ("raster_file", CFUNCTYPE(pnanovdb_bool_t, c_void_p, POINTER(pnanovdb_Compute), c_void_p, c_char_p, c_float,
                                  POINTER(_ArrP), POINTER(c_void_p), POINTER(c_void_p), POINTER(_ArrP), c_void_p,
                                  c_void_p, c_void_p)),
        ("raster_to_nanovdb_from_arrays", CFUNCTYPE(pnanovdb_bool_t, c_void_p, POINTER(pnanovdb_Compute), c_void_p,
                                                    c_float, POINTER(_ArrP), c_uint32, POINTER(_ArrP))),
        ("create_gaussian_data_from_arrays", CFUNCTYPE(pnanovdb_bool_t, c_void_p, POINTER(pnanovdb_Compute),
                                                       c_void_p, POINTER(_ArrP), c_uint32, POINTER(c_void_p),
                                                       c_void_p, POINTER(c_void_p))),
        ("create_gaussian_data_from_desc", c_void_p),
    ]


class Raster:
    """Python wrapper for pnanovdb_raster_t."""

    def __init__(self, compute: Compute = None):
        L = lib()
        ptr = L.pnanovdb_get_raster()
        if not ptr:
            raise RuntimeError("Failed to get raster")
        self._raster = POINTER(pnanovdb_Raster).from_buffer(c_void_p(ptr))
        self._compute = compute if compute is not None else Compute()
        self._raster.contents.compute = self._compute.get_compute()

    def raster_to_nanovdb(self, voxel_size, means, quaternions, scales, colors, sh_0, sh_n, opacities,
                          shader_params_arrays=None, profiler_report=None, userdata=None) -> pnanovdb_ComputeArray:
        """Preprocessed arrays -> node2 grid array (raster.py:190-229 call shape)."""
        func = self._raster.contents.raster_to_nanovdb
        arr = func(self._compute.get_compute(), None, voxel_size, pointer(means), pointer(quaternions),
                   pointer(scales), pointer(colors), pointer(sh_0), pointer(sh_n), pointer(opacities),
                   shader_params_arrays, profiler_report or c_void_p(), userdata or c_void_p())
        if not arr:
            raise RuntimeError("Failed to raster points")
        return arr.contents

    def raster_to_nanovdb_from_arrays(self, voxel_size: float, arrays) -> pnanovdb_ComputeArray:
        """Raw arrays [means, opacities, quaternions, scales, sh_0, sh_n] -> grid array.

        Preprocesses the arrays (quaternion normalisation, scale exponentiation, sh_0 colour,
        opacity sigmoid) in place, as raster.py:231-275 documents.
        """
        if len(arrays) != 6:
            raise InvalidArgumentError(
                "Gaussian rasterization requires 6 arrays: [means, opacities, quaternions, scales, sh_0, sh_n]")
        func = self._raster.contents.raster_to_nanovdb_from_arrays
        arrays_c = (_ArrP * 6)(*[pointer(a) for a in arrays])
        out = _ArrP()
        ok = func(self._raster, self._compute.get_compute(), None, c_float(voxel_size), arrays_c, c_uint32(6),
                  byref(out))
        if not ok or not out:
            raise RuntimeError("Failed to rasterize Gaussian arrays to NanoVDB")
        return out.contents

    def raster_file(self, filename: str, voxel_size: float, shader_params_arrays=None) -> pnanovdb_ComputeArray:
        """Gaussian PLY / NPZ file -> OnIndex grid array (RasterUtils.cpp:295-395)."""
        func = self._raster.contents.raster_file
        out = _ArrP()
        ok = func(self._raster, self._compute.get_compute(), None, filename.encode(), c_float(voxel_size), byref(out),
                  None, None, shader_params_arrays, None, None, None)
        if not ok or not out:
            raise RuntimeError(f"Failed to rasterize Gaussian file {filename}")
        return out.contents
