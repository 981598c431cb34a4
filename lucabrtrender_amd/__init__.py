"""MI355X-native Gaussian-splat tile rasteriser: HIP kernels for gfx950 behind
the C ABI of include/lrt_hip.h, with a Python mirror of the reference's
lrt::render interface. Importing fails loudly when liblrt_hip.so is missing."""
from ._abi import LrtError, lib  # noqa: F401  (loads liblrt_hip.so or raises)
from .raster import (Camera, FrameStats, GpuSplats, Lens, RenderSettings, RenderTargets, SplatEdit,  # noqa: F401
                     SplatInstance, TileRasterizer, projection_for)

__all__ = ["Camera", "FrameStats", "GpuSplats", "Lens", "LrtError", "RenderSettings", "RenderTargets", "SplatEdit",
           "SplatInstance", "TileRasterizer", "projection_for"]
