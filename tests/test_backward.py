"""Backward pass: the GPU adjoint (lucid_backward through the C ABI) against a dense float64
PyTorch compositing reference built from the oracle's exact fragment lists, whose own
autograd is pinned by float64 central finite differences (h = 1e-4).

Bar: gradients within 1e-4 relative (vector norm) of the float64 reference, and on 64
random parameters within 1e-4 relative (plus a 1e-9 absolute floor for the finite
differences' truncation error) of the finite differences.
"""

import numpy as np
import pytest

from lucid_amd import scenes

from dense_ref import DenseCompositor

H = 1e-4


def _fd_check_params(rng, grad, k=64):
    """k random parameters among those with a gradient (visible triangles)."""
    nz = np.flatnonzero(np.abs(grad) > 1e-6 * np.abs(grad).max())
    return rng.choice(nz, size=min(k, len(nz)), replace=False)


def _loss64(torch, comp, colors, alpha, target):
    img = comp.image(colors, alpha)
    return ((img - target) ** 2).mean()


def _setup(oracle_mod, sc, seed=5):
    import torch
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES, want_fragments=True)
    comp = DenseCompositor(sc, res["fragments"])
    rng = np.random.default_rng(seed)
    colors = sc.colors_rgb.astype(np.float64)
    alpha = sc.tri_alpha.astype(np.float64)
    # target: the image of N(0, 0.1)-perturbed parameters
    with torch.no_grad():
        tc = torch.from_numpy(np.clip(colors + rng.normal(0, 0.1, colors.shape), 0, 1))
        ta = torch.from_numpy(np.clip(alpha + rng.normal(0, 0.1, alpha.shape), 0.01, 0.99))
        target = comp.image(tc, ta)
    return res, comp, colors, alpha, target


def _ref_grads(oracle_mod, sc):
    import torch
    res, comp, colors, alpha, target = _setup(oracle_mod, sc)
    c64 = torch.from_numpy(colors).requires_grad_(True)
    a64 = torch.from_numpy(alpha).requires_grad_(True)
    loss = _loss64(torch, comp, c64, a64, target)
    loss.backward()
    return res, comp, colors, alpha, target, c64.grad.numpy(), a64.grad.numpy()


def test_dense_reference_matches_oracle_image(oracle_mod):
    import torch
    sc = scenes.config0(num_tris=600, width=64, height=64, seed=21)
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES, want_fragments=True)
    comp = DenseCompositor(sc, res["fragments"])
    img = comp.image(torch.from_numpy(sc.colors_rgb.astype(np.float64)),
                     torch.from_numpy(sc.tri_alpha.astype(np.float64))).numpy()
    ref = res["image_f"][..., :3]
    assert np.abs(img - ref).max() < 1e-5


def test_dense_reference_gradient_matches_finite_differences(oracle_mod):
    """Pins the float64 reference's autograd with central differences (h = 1e-4)."""
    import torch
    sc = scenes.config0(num_tris=300, width=48, height=48, seed=22)
    res, comp, colors, alpha, target, gc, ga = _ref_grads(oracle_mod, sc)
    rng = np.random.default_rng(3)
    for name, base, grad in (("colors", colors, gc), ("alpha", alpha, ga)):
        flat = base.reshape(-1)
        for i in _fd_check_params(rng, grad.reshape(-1), k=32):
            vals = []
            for sgn in (1.0, -1.0):
                p = flat.copy()
                p[i] += sgn * H
                cc = torch.from_numpy(p.reshape(colors.shape) if name == "colors" else colors)
                aa = torch.from_numpy(p if name == "alpha" else alpha)
                vals.append(float(_loss64(torch, comp, cc, aa, target)))
            fd = (vals[0] - vals[1]) / (2 * H)
            g = grad.reshape(-1)[i]
            assert abs(g - fd) <= 1e-4 * abs(fd) + 1e-9, (name, i, g, fd)


def _gpu_grads(sc, target, record=False):
    import torch
    from lucid_amd import _lib
    from lucid_amd.renderer import GpuScene, LucidRenderer
    from lucid_amd.autograd import render_image
    print(f"\n[case] backward {sc.name} {sc.width}x{sc.height} quads={sc.num_quads} "
          f"record={record}", flush=True)
    r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES)
    r.record_backward = record
    gs = GpuScene(sc)
    col = torch.tensor(sc.colors_rgb, device="cuda", requires_grad=True)
    alpha = torch.tensor(sc.tri_alpha, device="cuda", requires_grad=True)
    img = render_image(r, gs, col, alpha)
    t = torch.tensor(target, dtype=torch.float32, device="cuda")
    loss = ((img[..., :3] - t) ** 2).mean()
    loss.backward()
    return img.detach().cpu().numpy(), col.grad.cpu().numpy(), alpha.grad.cpu().numpy()


def _check_grads(g, ref, what):
    n = np.linalg.norm(ref)
    assert n > 0, what
    rel = np.linalg.norm(g - ref) / n
    assert rel < 1e-4, f"{what}: relative error {rel:.3g}"


@pytest.mark.gpu
@pytest.mark.parametrize("ntris,size,seed", [(400, 64, 31), (3000, 128, 32)])
def test_backward_matches_dense_reference(oracle_mod, ntris, size, seed):
    sc = scenes.config0(num_tris=ntris, width=size, height=size, seed=seed)
    res, comp, colors, alpha, target, gc, ga = _ref_grads(oracle_mod, sc)
    img, col_g, alpha_g = _gpu_grads(sc, target.numpy())
    assert np.abs(img[..., :3] - res["image_f"][..., :3]).max() <= 1.0 / 255.0
    _check_grads(col_g, gc, "dL/dcolors")
    _check_grads(alpha_g[0::2], ga[0::2], "dL/dalpha")
    assert np.all(alpha_g[1::2] == 0.0), "culled second triangles get no gradient"


@pytest.mark.gpu
def test_backward_finite_differences_64_params(oracle_mod):
    """GPU gradients on 64 random parameters against float64 central differences."""
    import torch
    sc = scenes.config0(num_tris=2000, width=96, height=96, seed=33)
    res, comp, colors, alpha, target = _setup(oracle_mod, sc)
    img, col_g, alpha_g = _gpu_grads(sc, target.numpy())
    rng = np.random.default_rng(4)
    grads = np.concatenate([col_g.reshape(-1), alpha_g])
    params = np.concatenate([colors.reshape(-1), alpha])
    nc = colors.size
    for i in _fd_check_params(rng, grads, k=64):
        vals = []
        for sgn in (1.0, -1.0):
            p = params.copy()
            p[i] += sgn * H
            cc = torch.from_numpy(p[:nc].reshape(colors.shape))
            aa = torch.from_numpy(p[nc:])
            vals.append(float(_loss64(torch, comp, cc, aa, target)))
        fd = (vals[0] - vals[1]) / (2 * H)
        assert abs(grads[i] - fd) <= 1e-4 * abs(fd) + 1e-9, (i, grads[i], fd)


@pytest.mark.gpu
def test_backward_exact_fallback_and_deep_pixels(oracle_mod):
    """Ordering violations (tilted planes) send pixels through the exact per-pixel adjoint."""
    sc = scenes.tilted_planes(40, 32, 24, seed=11)
    res, comp, colors, alpha, target, gc, ga = _ref_grads(oracle_mod, sc)
    img, col_g, alpha_g = _gpu_grads(sc, target.numpy())
    _check_grads(col_g, gc, "dL/dcolors")
    _check_grads(alpha_g, ga, "dL/dalpha")


@pytest.mark.gpu
def test_backward_is_deterministic(oracle_mod):
    sc = scenes.config0(num_tris=1500, width=96, height=96, seed=34)
    res, comp, colors, alpha, target = _setup(oracle_mod, sc)
    a = _gpu_grads(sc, target.numpy())
    b = _gpu_grads(sc, target.numpy())
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.gpu
def test_backward_reusing_the_frame_matches_full_recompute(oracle_mod):
    """lucid_backward_ex with LUCID_BACKWARD_REUSE_FRAME (the autograd path) gives the same
    gradients, bit for bit, as the full recompute; it falls back to the full recompute when
    another frame was rendered in between. Covers ordering violations (tilted planes: exact
    per-pixel adjoints) and a block too long to pre-sort (block stack: its pixels are queued
    again from the reused lists)."""
    import torch
    from lucid_amd import _lib
    from lucid_amd.renderer import GpuScene, LucidRenderer
    for sc in (scenes.config0(num_tris=1500, width=96, height=96, seed=35),
               scenes.tilted_planes(40, 32, 24, seed=12),
               scenes.block_stack(32, 32, 8300, center_px=(12, 12), size_px=4.0, seed=14)):
        print(f"\n[case] backward reuse {sc.name} {sc.width}x{sc.height}", flush=True)
        r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES)
        gs = GpuScene(sc)
        other = GpuScene(scenes.config0(num_tris=500, width=sc.width, height=sc.height, seed=36))
        out = r.alloc_outputs(rgba8=False, rgba32f=True, counts=False)
        gen = torch.Generator(device="cuda").manual_seed(7)
        grad = torch.randn((sc.height, sc.width, 4), device="cuda", generator=gen) * 1e-2
        r.render(gs, out)
        full = r.backward(gs, grad)                   # rebuilds everything
        r.render(gs, out)
        reused = r.backward(gs, grad, reuse_frame=True)
        again = r.backward(gs, grad, reuse_frame=True)  # the lists stay valid
        r.render(other, out)
        fallback = r.backward(gs, grad, reuse_frame=True)  # other inputs: full recompute
        torch.cuda.synchronize()
        for got in (reused, again, fallback):
            assert torch.equal(got[0], full[0]) and torch.equal(got[1], full[1])


@pytest.mark.gpu
def test_backward_from_a_recording_forward_matches_full_recompute():
    """A frame rendered with LUCID_RENDER_RECORD_BACKWARD (what the autograd binding renders)
    has the plain frame's fragment counts and colours (within fp32 rounding: pixels the
    recording pass leaves to the exact sort), and the backward that starts from its record
    gives the full recompute's gradients bit for bit, also when run twice on one record.
    Covers ordering violations (tilted planes) and a block too long to pre-sort."""
    import torch
    from lucid_amd import _lib
    from lucid_amd.renderer import GpuScene, LucidRenderer
    for sc in (scenes.config0(num_tris=1500, width=96, height=96, seed=35),
               scenes.tilted_planes(40, 32, 24, seed=12),
               scenes.block_stack(32, 32, 8300, center_px=(12, 12), size_px=4.0, seed=14)):
        print(f"\n[case] recording forward {sc.name} {sc.width}x{sc.height}", flush=True)
        r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES)
        gs = GpuScene(sc)
        plain = r.alloc_outputs(rgba8=False, rgba32f=True, counts=True)
        rec = r.alloc_outputs(rgba8=False, rgba32f=True, counts=True)
        gen = torch.Generator(device="cuda").manual_seed(8)
        grad = torch.randn((sc.height, sc.width, 4), device="cuda", generator=gen) * 1e-2
        r.render(gs, plain)
        full = r.backward(gs, grad)
        r.render(gs, rec, record_backward=True)
        assert r.frame_recorded(), "the full backward sized the buffers: the record must fit"
        assert torch.equal(rec["frag_counts"], plain["frag_counts"])
        assert (rec["image_f"] - plain["image_f"]).abs().max().item() <= 1e-6
        first = r.backward(gs, grad, reuse_frame=True)
        again = r.backward(gs, grad, reuse_frame=True)
        torch.cuda.synchronize()
        for got in (first, again):
            assert torch.equal(got[0], full[0]) and torch.equal(got[1], full[1])


@pytest.mark.gpu
def test_recording_forward_beyond_the_backward_buffers_renders_without_record():
    """On a fresh renderer the backward buffers have their initial sizes (65536 recorded
    rows); a 60k-triangle frame needs more, so it renders without the record (same image),
    the buffers grow, the backward walks the frame itself (same gradients), and the next
    recording frame records."""
    import torch
    from lucid_amd import _lib
    from lucid_amd.renderer import GpuScene, LucidRenderer
    sc = scenes.config0(num_tris=60000, width=256, height=256, seed=37)
    print(f"\n[case] recording forward beyond the buffers {sc.name}", flush=True)
    gs = GpuScene(sc)
    gen = torch.Generator(device="cuda").manual_seed(9)
    grad = torch.randn((sc.height, sc.width, 4), device="cuda", generator=gen) * 1e-2
    ref = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES)
    plain = ref.render(gs, ref.alloc_outputs(rgba8=False, rgba32f=True, counts=True))
    full = ref.backward(gs, grad)
    r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES)
    out = r.render(gs, r.alloc_outputs(rgba8=False, rgba32f=True, counts=True),
                   record_backward=True)
    assert not r.frame_recorded()
    assert torch.equal(out["frag_counts"], plain["frag_counts"])
    assert torch.equal(out["image_f"], plain["image_f"])
    got = r.backward(gs, grad, reuse_frame=True)
    assert torch.equal(got[0], full[0]) and torch.equal(got[1], full[1])
    r.render(gs, out, record_backward=True)
    assert r.frame_recorded()
    got = r.backward(gs, grad, reuse_frame=True)
    torch.cuda.synchronize()
    assert torch.equal(got[0], full[0]) and torch.equal(got[1], full[1])


@pytest.mark.gpu
def test_autograd_with_recording_forward_matches_default(oracle_mod):
    """renderer.record_backward (the forward stores each pixel's blend order) gives the same
    image and gradients, bit for bit, as the default (the backward recomputes them)."""
    sc = scenes.config0(num_tris=3000, width=128, height=128, seed=32)
    target = np.random.default_rng(3).uniform(0, 1, (sc.height, sc.width, 3))
    a = _gpu_grads(sc, target)
    b = _gpu_grads(sc, target, record=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.gpu
def test_c4_full_size_crop_reference_and_finite_differences(oracle_mod):
    """The differentiable workload at its full size: 200k soup triangles at 512x512 (~41M
    fragments), per-vertex RGB and per-triangle opacity logits (alpha = sigmoid) as fp32 leaf
    tensors, MSE over a 64x64 crop against the image of N(0, 0.1)-perturbed parameters.
    The GPU gradient (all parameters, through the autograd binding) against the dense
    float64 compositor built from the oracle's exact fragment lists inside the crop (1e-4
    relative, vector norm), and 64 random parameters against float64 central finite
    differences (h = 1e-4) of that compositor's loss (1e-4 relative + 1e-9 absolute)."""
    import dataclasses

    import torch
    from lucid_amd import _lib
    from lucid_amd.autograd import render_image_logits
    from lucid_amd.renderer import GpuScene, LucidRenderer
    sc = scenes.config4()
    x0, y0, cw = 224, 224, 64
    print(f"\n[case] C4 full size {sc.name} crop ({x0},{y0})+{cw}", flush=True)
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES, want_records=False,
                            want_fragments=True, fragment_rect=(x0, y0, x0 + cw, y0 + cw))
    fr = res["fragments"].copy()
    fr["x"] -= x0
    fr["y"] -= y0
    crop_scene = dataclasses.replace(sc, width=cw, height=cw)
    comp = DenseCompositor(crop_scene, fr)
    rng = np.random.default_rng(1004)
    colors = sc.colors_rgb.astype(np.float32)
    a = sc.tri_alpha.astype(np.float64)
    logits = np.log(a / (1.0 - a)).astype(np.float32)
    with torch.no_grad():
        tc = torch.from_numpy(np.clip(colors.astype(np.float64) +
                                      rng.normal(0, 0.1, colors.shape), 0, 1))
        tl = torch.from_numpy(logits.astype(np.float64) + rng.normal(0, 0.1, logits.shape))
        target = comp.image(tc, torch.sigmoid(tl))

    def loss64(c, l):
        return ((comp.image(c, torch.sigmoid(l)) - target) ** 2).mean()

    c64 = torch.from_numpy(colors.astype(np.float64)).requires_grad_(True)
    l64 = torch.from_numpy(logits.astype(np.float64)).requires_grad_(True)
    loss64(c64, l64).backward()
    ref_gc, ref_gl = c64.grad.numpy(), l64.grad.numpy()

    r = LucidRenderer((sc.width, sc.height), opts=_lib.OPT_FP32_SAMPLES, max_quads=sc.num_quads)
    gs = GpuScene(sc)
    col = torch.tensor(colors, device="cuda", requires_grad=True)
    lg = torch.tensor(logits, device="cuda", requires_grad=True)
    img = render_image_logits(r, gs, col, lg)
    t = torch.tensor(target.numpy(), dtype=torch.float32, device="cuda")
    loss = ((img[y0:y0 + cw, x0:x0 + cw, :3] - t) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    crop = img.detach()[y0:y0 + cw, x0:x0 + cw, :3].cpu().numpy()
    assert np.abs(crop - res["image_f"][y0:y0 + cw, x0:x0 + cw, :3]).max() <= 1.0 / 255.0
    gc, gl = col.grad.cpu().numpy(), lg.grad.cpu().numpy()
    _check_grads(gc, ref_gc, "dL/dcolors")
    _check_grads(gl, ref_gl, "dL/dlogits")
    assert np.all(gl[1::2] == 0.0), "culled second triangles get no gradient"
    print(f"[c4] crop fragments {len(fr)}, triangles with a gradient "
          f"{int((gl != 0).sum())}", flush=True)

    grads = np.concatenate([gc.reshape(-1), gl])
    params = np.concatenate([colors.reshape(-1).astype(np.float64), logits.astype(np.float64)])
    nc = colors.size
    for i in _fd_check_params(np.random.default_rng(4), grads, k=64):
        vals = []
        for sgn in (1.0, -1.0):
            p = params.copy()
            p[i] += sgn * H
            with torch.no_grad():
                vals.append(float(loss64(torch.from_numpy(p[:nc].reshape(colors.shape)),
                                         torch.from_numpy(p[nc:]))))
        fd = (vals[0] - vals[1]) / (2 * H)
        assert abs(grads[i] - fd) <= 1e-4 * abs(fd) + 1e-9, (i, grads[i], fd)
