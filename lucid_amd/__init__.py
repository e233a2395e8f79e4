"""MI355X-native exact order-independent-transparency compute rasterizer.

Drop-in for nadult/lucid's LucidRenderer forward path: scene quads + LucidConfig in,
RGBA image and per-pixel fragment counts out, computed by hand-written HIP kernels
for gfx950 behind the C ABI in include/lucid_c.h.
"""

from .scenes import Camera, Scene, make_config  # noqa: F401


def __getattr__(name):
    if name in ("LucidRenderer", "GpuScene"):
        from . import renderer
        return getattr(renderer, name)
    raise AttributeError(name)
