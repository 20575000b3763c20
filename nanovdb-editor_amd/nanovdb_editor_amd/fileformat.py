"""FileFormat function table binding (nanovdb_editor/putil/FileFormat.h:19-35).

``FileFormat.load_file`` returns the requested Gaussian arrays of a binary PLY or NPZ file as
compute arrays created by the library's compute table (fileformat/FileFormat.cpp:754-790);
``FileFormat.load_numpy`` copies them out as numpy arrays.
"""
from ctypes import CFUNCTYPE, POINTER, Structure, c_char_p, c_uint32, c_void_p

import numpy as np

from ._lib import lib
from .compute import Compute, pnanovdb_bool_t, pnanovdb_ComputeArray

_ArrP = POINTER(pnanovdb_ComputeArray)

GAUSSIAN_ARRAYS = ("means", "opacities", "quaternions", "scales", "sh_0", "sh_n")


class pnanovdb_FileFormat(Structure):
    _fields_ = [
        ("interface_pnanovdb_reflect_data_type", c_void_p),
        ("can_load_file", CFUNCTYPE(pnanovdb_bool_t, c_char_p)),
        ("load_file", CFUNCTYPE(pnanovdb_bool_t, c_char_p, c_uint32, POINTER(c_char_p), POINTER(_ArrP))),
        ("module", c_void_p),
    ]


class FileFormat:
    def __init__(self, compute: Compute = None):
        ptr = lib().pnanovdb_get_fileformat()
        if not ptr:
            raise RuntimeError("Failed to get fileformat")
        self._ff = POINTER(pnanovdb_FileFormat).from_buffer(c_void_p(ptr))
        self._compute = compute if compute is not None else Compute()

    def can_load_file(self, filename: str) -> bool:
        return bool(self._ff.contents.can_load_file(filename.encode()))

    def load_file(self, filename: str, names=GAUSSIAN_ARRAYS):
        """Compute arrays (or None where the file lacks one) in the order of ``names``."""
        names_c = (c_char_p * len(names))(*[n.encode() for n in names])
        out = (_ArrP * len(names))()
        if not self._ff.contents.load_file(filename.encode(), len(names), names_c, out):
            raise RuntimeError(f"Failed to load file: {filename}")
        return [out[i].contents if out[i] else None for i in range(len(names))]

    def load_numpy(self, filename: str, names=GAUSSIAN_ARRAYS, dtype=np.float32):
        """The arrays as numpy copies (the compute arrays are freed)."""
        arrays = self.load_file(filename, names)
        res = []
        for a in arrays:
            if a is None:
                res.append(None)
                continue
            res.append(self._compute.map_array(a, dtype).copy())
            self._compute.unmap_array(a)
            self._compute.destroy_array(a)
        return res
