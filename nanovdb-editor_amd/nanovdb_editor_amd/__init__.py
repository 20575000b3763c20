"""MI355X (gfx950) Gaussian -> NanoVDB voxelizer: Python front end of libvdbraster.so.

Host-array API (reference-compatible): ``Compute``, ``Raster``.
Device API on torch tensors: ``device.Voxelizer``.
"""
from .exceptions import InvalidArgumentError, PipelineError

__all__ = ["InvalidArgumentError", "PipelineError"]
