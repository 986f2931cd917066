"""athenea_amd: an MI355X-native (gfx950) Gaussian-splat tile rasterizer.

The hot path of athenea's splat renderer -- projection, 64-bit (tile, depth)
key duplication, onesweep radix sort, tile ranges, LDS-staged front-to-back
blend, and a hand-written adjoint -- as HIP kernels in libathx.so behind the
C ABI of include/athx.h. This package is the host-side mirror of athenea's
render::TileRasterizer API over that library, plus the camera math and the
seeded synthetic workloads.
"""
from .camera import Camera, Lens, Projection, make_view, projection_for, view_for  # noqa: F401
from .raster import (  # noqa: F401
    FrameStats,
    GaussianCloud,
    RenderSettings,
    RenderTargets,
    SplatCounters,
    SplatInstance,
    TileRasterizer,
    UnderLayer,
)

__all__ = [
    "Camera",
    "Lens",
    "Projection",
    "make_view",
    "projection_for",
    "view_for",
    "FrameStats",
    "GaussianCloud",
    "RenderSettings",
    "RenderTargets",
    "SplatCounters",
    "SplatInstance",
    "TileRasterizer",
    "UnderLayer",
]
