"""Host arrays through the compute function table (mirror of pymodule/nanovdb_editor/compute.py).

``pnanovdb_ComputeArray`` and ``Compute.create_array/destroy_array/map_array`` keep the
reference's layout and call shapes (compute.py:43-53, 273-343); the table comes from this
library's ``pnanovdb_get_compute`` (nanovdb_editor/putil/Compute.h:784-853 layout).
"""
from ctypes import (CFUNCTYPE, POINTER, Structure, c_char_p, c_int32, c_size_t, c_uint32, c_uint64, c_void_p,
                    pointer)

import numpy as np

from ._lib import lib
from .exceptions import InvalidArgumentError, PipelineError

pnanovdb_bool_t = c_int32


class pnanovdb_ComputeArray(Structure):
    """Definition equivalent to pnanovdb_compute_array_t (Compute.h:765-771)."""

    _fields_ = [("data", c_void_p), ("element_size", c_uint64), ("element_count", c_uint64), ("filepath", c_char_p)]


class pnanovdb_ShaderInterface(Structure):
    _fields_ = [("interface_pnanovdb_reflect_data_type", c_void_p), ("create_shader", c_void_p),
                ("map_shader_build", c_void_p), ("destroy_shader", c_void_p)]


class pnanovdb_DeviceInterface(Structure):
    _fields_ = [("interface_pnanovdb_reflect_data_type", c_void_p)] + [
        (name, c_void_p) for name in (
            "create_device_manager", "destroy_device_manager", "enumerate_devices", "create_device",
            "destroy_device", "get_memory_stats", "create_semaphore", "destroy_semaphore",
            "get_semaphore_external_handle", "close_semaphore_external_handle", "get_device_queue",
            "get_compute_queue", "flush", "get_frame_global_completed", "wait_for_frame", "wait_idle",
            "get_compute_interface", "get_compute_context", "create_swapchain", "destroy_swapchain",
            "resize_swapchain", "present_swapchain", "get_swapchain_front_texture", "create_encoder",
            "destroy_encoder", "present_encoder", "get_encoder_front_texture", "map_encoder_data",
            "unmap_encoder_data", "enable_profiler", "disable_profiler", "set_resource_min_lifetime",
            "get_device_index")]


class pnanovdb_Compute(Structure):
    """Definition equivalent to pnanovdb_compute_t (Compute.h:784-853)."""

    _fields_ = [
        ("interface_pnanovdb_reflect_data_type", c_void_p),
        ("module", c_void_p),
        ("compiler", c_void_p),
        ("shader_interface", pnanovdb_ShaderInterface),
        ("device_interface", pnanovdb_DeviceInterface),
        ("load_nanovdb", c_void_p),
        ("save_nanovdb", c_void_p),
        ("create_shader_context", c_void_p),
        ("destroy_shader_context", c_void_p),
        ("init_shader", c_void_p),
        ("destroy_shader", c_void_p),
        ("dispatch_shader", c_void_p),
        ("dispatch_shader_on_array", c_void_p),
        ("dispatch_shader_on_nanovdb_array", c_void_p),
        ("create_array", CFUNCTYPE(POINTER(pnanovdb_ComputeArray), c_size_t, c_uint64, c_void_p)),
        ("destroy_array", CFUNCTYPE(None, POINTER(pnanovdb_ComputeArray))),
        ("map_array", CFUNCTYPE(c_void_p, POINTER(pnanovdb_ComputeArray))),
        ("unmap_array", CFUNCTYPE(None, POINTER(pnanovdb_ComputeArray))),
        ("compute_array_print_range", c_void_p),
        ("nanovdb_from_image_rgba8", c_void_p),
        ("duplicate_array", CFUNCTYPE(POINTER(pnanovdb_ComputeArray), POINTER(pnanovdb_ComputeArray))),
    ]


class Compute:
    """Host array management through the library's compute table."""

    def __init__(self):
        L = lib()
        ptr = L.pnanovdb_get_compute()
        if not ptr:
            raise RuntimeError("Failed to get compute")
        self._compute = POINTER(pnanovdb_Compute).from_buffer(c_void_p(ptr))

    def get_compute(self) -> POINTER(pnanovdb_Compute):
        return self._compute

    def create_array(self, data: np.ndarray) -> pnanovdb_ComputeArray:
        data = np.ascontiguousarray(data)
        array = self._compute.contents.create_array(data.itemsize, data.size, data.ctypes.data_as(c_void_p))
        if not array:
            raise PipelineError("Failed to create compute array")
        return array.contents

    def duplicate_array(self, array: pnanovdb_ComputeArray) -> pnanovdb_ComputeArray:
        dup = self._compute.contents.duplicate_array(pointer(array))
        if not dup:
            raise PipelineError("Failed to duplicate compute array")
        return dup.contents

    def destroy_array(self, array: pnanovdb_ComputeArray) -> None:
        self._compute.contents.destroy_array(pointer(array))

    def map_array(self, array: pnanovdb_ComputeArray, np_dtype) -> np.ndarray:
        ptr = self._compute.contents.map_array(pointer(array))
        if not ptr:
            raise PipelineError("Failed to map compute array")
        dt = np.dtype(np_dtype)
        count = array.element_count * array.element_size // dt.itemsize
        buf = (np.ctypeslib.as_array((c_uint32 * 0).from_address(ptr)) if count == 0 else
               np.ctypeslib.as_array((np.ctypeslib.as_ctypes_type(dt) * count).from_address(ptr)))
        return buf.view(dt)

    def unmap_array(self, array: pnanovdb_ComputeArray) -> None:
        self._compute.contents.unmap_array(pointer(array))

    def array_bytes(self, array: pnanovdb_ComputeArray) -> np.ndarray:
        """Copy of the array contents as uint8."""
        n = int(array.element_count * array.element_size)
        return np.ctypeslib.as_array((np.ctypeslib.as_ctypes_type(np.uint8) * n).from_address(array.data)).copy() \
            if n else np.zeros(0, np.uint8)


__all__ = ["Compute", "pnanovdb_Compute", "pnanovdb_ComputeArray", "pnanovdb_bool_t", "InvalidArgumentError"]
