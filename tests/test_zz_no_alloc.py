"""lucid_render_noalloc (the allocation-free frame) against lucid_render through the C ABI.

Both run the same kernels on the same inputs, so fragment counts, the RGBA8 image and the
fp32 image must be identical. Two calls in a row, the second on a smaller frame (the
shared scratch array keeps the first call's contents), then sizes beyond the limits.
"""

import ctypes

import numpy as np
import pytest

from lucid_amd import _lib, scenes

pytestmark = pytest.mark.gpu


def _noalloc(sc, gs, opts):
    import torch
    lib = _lib.load()
    h, w = sc.height, sc.width
    img = torch.zeros((h, w, 4), dtype=torch.uint8, device="cuda")
    img_f = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    cnt = torch.zeros((h, w), dtype=torch.int32, device="cuda")
    o = _lib.FrameOutputs(img.data_ptr(), img_f.data_ptr(), cnt.data_ptr())
    inp = gs.inputs()
    st = torch.cuda.current_stream()
    rc = lib.lucid_render_noalloc(w, h, opts, None, ctypes.byref(inp), ctypes.byref(o),
                                  ctypes.c_void_p(st.cuda_stream))
    _lib.check(rc, "lucid_render_noalloc")
    return img.cpu().numpy(), img_f.cpu().numpy(), cnt.cpu().numpy()


def _allocating(sc, gs, opts):
    from lucid_amd.renderer import LucidRenderer
    r = LucidRenderer((sc.width, sc.height), opts=opts)
    out = r.render(gs, r.alloc_outputs(rgba8=True, rgba32f=True, counts=True))
    return (out["image"].cpu().numpy(), out["image_f"].cpu().numpy(),
            out["frag_counts"].cpu().numpy())


@pytest.mark.parametrize("opts", [0, _lib.OPT_FP32_SAMPLES])
def test_noalloc_matches_lucid_render_twice(opts):
    from lucid_amd.renderer import GpuScene
    # first a clustered frame (21 high bins, up to 209 fragments per pixel), then a smaller,
    # sparser one
    cases = [scenes.config1(num_tris=20_000, width=256, height=256),
             scenes.config0(seed=72, num_tris=2_000, width=160, height=96)]
    for sc in cases:
        print(f"\n[case] noalloc {sc.name} {sc.width}x{sc.height} quads={sc.num_quads} "
              f"opts={opts:#x}", flush=True)
        gs = GpuScene(sc)
        a = _noalloc(sc, gs, opts)
        b = _allocating(sc, gs, opts)
        assert int(a[2].sum()) > 0
        assert np.array_equal(a[2], b[2]), "fragment counts"
        assert np.array_equal(a[0], b[0]), "RGBA8 image"
        assert np.array_equal(a[1], b[1]), "fp32 image"


def test_noalloc_rejects_sizes_beyond_its_limits():
    import torch
    from lucid_amd.renderer import GpuScene
    lib = _lib.load()
    sc = scenes.config0(seed=73, num_tris=100, width=64, height=64)
    gs = GpuScene(sc)
    o = _lib.FrameOutputs(None, None, None)
    inp = gs.inputs()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # 1024 x 1024 is 1024 bins (limit 256)
    assert lib.lucid_render_noalloc(1024, 1024, 0, None, ctypes.byref(inp), ctypes.byref(o),
                                    st) == _lib.LUCID_ERR_INVALID_ARG
    for bad in (_lib.Tuning(3, 256, 2048, 1), _lib.Tuning(1, 256, 2048, 3)):
        assert lib.lucid_render_noalloc(64, 64, 0, ctypes.byref(bad), ctypes.byref(inp),
                                        ctypes.byref(o), st) == _lib.LUCID_ERR_INVALID_ARG
