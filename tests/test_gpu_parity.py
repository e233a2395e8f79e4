"""GPU parity: the HIP path (through the C ABI) against the C oracle.

Bar: bin counts, per-bin lists (as multisets) and per-pixel fragment counts bit-exact; blended image within 1/255 per channel (exact per-pixel order, ties by
triangle id). Scenes cover the reference's edge cases: empty frame, culled/between-sample
quads, low bins promoted to high (> 256 tris in an 8x8 block), high bins (>= 1024 tris),
blocks too large for LDS, per-pixel ordering violations (exact fallback) and pixels deeper
than the fallback's LDS capacity (global spill).
"""

import numpy as np
import pytest

from lucid_amd import scenes
from lucid_amd import _lib

pytestmark = pytest.mark.gpu

INFO_U32 = 1152


def _render_both(sc, oracle_mod, opts=0, tuning=None):
    from lucid_amd.renderer import GpuScene, LucidRenderer
    print(f"\n[case] {sc.name} {sc.width}x{sc.height} quads={sc.num_quads} opts={opts:#x} "
          f"tuning={tuning}", flush=True)
    r = LucidRenderer((sc.width, sc.height), opts=opts)
    if tuning:
        r.set_tuning(**tuning)
    out = r.render(GpuScene(sc), r.alloc_outputs(rgba8=True, rgba32f=True, counts=True))
    gpu = {
        "image": out["image"].cpu().numpy(),
        "image_f": out["image_f"].cpu().numpy(),
        "frag_counts": out["frag_counts"].cpu().numpy().astype(np.uint32),
        "info": r.read_info(),
        "verify": r.verify_info(),
    }
    gpu["bin_quads"], gpu["bin_tris"] = r.read_bin_lists()
    gpu["stats"] = r.get_stats()
    oopts = oracle_mod.OPT_FP32_SAMPLES if opts & _lib.OPT_FP32_SAMPLES else 0
    if opts & _lib.OPT_ADDITIVE_BLENDING:
        oopts |= oracle_mod.OPT_ADDITIVE
    ref = oracle_mod.render(sc, opts=oopts)
    return r, gpu, ref


def _check(sc, gpu, ref, oracle_mod, max_diff=1, max_diff_f=1.0 / 255.0):
    nb = ref["bin_count"]
    ginfo, rinfo = gpu["info"], ref["info"]
    # LucidInfo counters restated by the oracle
    assert ginfo[0] == rinfo[0], "num_input_quads"
    assert ginfo[1] == rinfo[1] and ginfo[2] == rinfo[2], "num_visible_quads (small, large)"
    assert np.array_equal(ginfo[32:36], rinfo[32:36]), "num_rejected_quads"
    assert np.array_equal(ginfo[5:10], rinfo[5:10]), "bin_level_counts (incl. promotions)"
    # per-bin counts and offsets
    for k in range(6):
        g = ginfo[INFO_U32 + k * nb: INFO_U32 + (k + 1) * nb]
        o = rinfo[INFO_U32 + k * nb: INFO_U32 + (k + 1) * nb]
        assert np.array_equal(g, o), f"g_counts array {k}"
    assert gpu["verify"] == 0, "verifyInfo invariants"
    # per-bin lists as multisets
    qc, qo = ginfo[INFO_U32: INFO_U32 + nb], ginfo[INFO_U32 + nb: INFO_U32 + 2 * nb]
    tc, to = ginfo[INFO_U32 + 3 * nb: INFO_U32 + 4 * nb], ginfo[INFO_U32 + 4 * nb: INFO_U32 + 5 * nb]
    for b in range(nb):
        gq = np.sort(gpu["bin_quads"][qo[b]: qo[b] + qc[b]])
        rq = np.sort(ref["bin_quads"][qo[b]: qo[b] + qc[b]])
        assert np.array_equal(gq, rq), f"bin {b} quad list"
        gt = np.sort(gpu["bin_tris"][to[b]: to[b] + tc[b]])
        rt = np.sort(ref["bin_tris"][to[b]: to[b] + tc[b]])
        assert np.array_equal(gt, rt), f"bin {b} tri list"
    # per-pixel fragment counts, bit-exact
    mism = np.argwhere(gpu["frag_counts"] != ref["frag_counts"])
    assert mism.size == 0, f"{len(mism)} pixels with different fragment counts, first {mism[:5]}"
    assert int(ginfo[60]) == ref["num_fragments"], "stats[0] (fragments)"
    # colour within 1/255 (tolerance in u8 steps)
    d = np.abs(gpu["image"].astype(np.int32) - ref["image"].astype(np.int32))
    assert d.max() <= max_diff, f"max colour diff {d.max()}/255 at {np.argwhere(d == d.max())[:3]}"
    if gpu.get("image_f") is not None:
        df = np.abs(gpu["image_f"] - ref["image_f"]).max()
        assert df <= max_diff_f, f"float image diff {df}"
    return d


def test_c0_small(oracle_mod):
    sc = scenes.config0(num_tris=2000, width=128, height=128)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)


def test_c0_full_f32(oracle_mod):
    sc = scenes.config0()
    r, gpu, ref = _render_both(sc, oracle_mod)
    d = _check(sc, gpu, ref, oracle_mod)
    print(f"[c0] fragments {ref['num_fragments']} exact-colour pixels "
          f"{(d.max(-1) == 0).mean():.6f}", flush=True)


def test_c0_full_rgba8(oracle_mod):
    sc = scenes.config0(color_format="rgba8")
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)


def test_c0_fp32_samples(oracle_mod):
    sc = scenes.config0(num_tris=3000, width=256, height=192)
    r, gpu, ref = _render_both(sc, oracle_mod, opts=_lib.OPT_FP32_SAMPLES)
    _check(sc, gpu, ref, oracle_mod)


def test_additive(oracle_mod):
    sc = scenes.config0(num_tris=1500, width=128, height=96)
    r, gpu, ref = _render_both(sc, oracle_mod, opts=_lib.OPT_ADDITIVE_BLENDING)
    _check(sc, gpu, ref, oracle_mod)


def test_alpha_threshold(oracle_mod):
    """LUCID_OPT_ALPHA_THRESHOLD (shading.glsl:62,282-284): blocks whose pixels all reach
    transmittance < 1/128 stop shading. Bins, lists and fragment counts stay bit-exact with
    the oracle's exact frame; colours stay within the skipped remainder (< 1/128 per
    channel) plus the 1/255 quantisation: at most 3/255 in RGBA8 and 1/128 + 1e-6 in fp32.
    ~100 fragments per pixel at alpha U[0.2, 0.8], so nearly every block stops early."""
    sc = scenes.config0(num_tris=20000, width=128, height=128, seed=41)
    for opts in (_lib.OPT_ALPHA_THRESHOLD, _lib.OPT_ALPHA_THRESHOLD | _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod, max_diff=3, max_diff_f=1.0 / 128.0 + 1e-6)


def test_empty_frame(oracle_mod):
    sc = scenes.config0(num_tris=0, width=64, height=64)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert (gpu["frag_counts"] == 0).all()


def test_planes_closed_form(oracle_mod):
    layers = [(3.0, (1.0, 0.0, 0.0), 0.5), (7.0, (0.0, 1.0, 0.0), 0.25),
              (5.0, (0.0, 0.0, 1.0), 0.75)]
    sc = scenes.planes_scene(96, 64, layers)
    r, gpu, ref = _render_both(sc, oracle_mod, opts=_lib.OPT_FP32_SAMPLES)
    _check(sc, gpu, ref, oracle_mod)
    # nearest first: red (3), blue (5), green (7)
    c = np.zeros(3)
    t = 1.0
    for _, col, a in sorted(layers):
        c += np.array(col) * a * t
        t *= 1 - a
    got = gpu["image_f"][20, 70, :3]
    assert np.allclose(got, c, atol=2e-3), (got, c)
    assert np.median(gpu["frag_counts"]) == 3


def test_promotion_low_to_high(oracle_mod):
    sc = scenes.block_stack(64, 64, 300)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert gpu["stats"]["promoted_bins"] >= 1


def test_high_bin(oracle_mod):
    sc = scenes.block_stack(96, 64, 700, center_px=(40, 20), size_px=10.0)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert gpu["stats"]["high_bins"] >= 1


def test_block_over_lds_capacity(oracle_mod):
    sc = scenes.block_stack(64, 64, 1300, center_px=(20, 20), size_px=5.0, seed=11)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)


def test_long_lists_second_sort_pass(oracle_mod):
    """Blocks of 2048-8192 entries: high_finalize's first pass queues them and its long-list
    pass (a fixed grid of 256 work-groups) pre-sorts them. 289 such blocks (136 x 136 px),
    so work-groups take more than one queued block."""
    rng = np.random.default_rng(17)
    depths = rng.permutation(np.linspace(2.0, 200.0, 2100))
    layers = [(float(d), tuple(rng.uniform(0, 1, 3)), float(rng.uniform(0.005, 0.05)))
              for d in depths]
    sc = scenes.planes_scene(136, 136, layers)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert int(ref["frag_counts"].min()) >= 2100


def test_block_over_sort_capacity(oracle_mod):
    # more than 8192 triangles in one 8x8 block: every pixel takes the exact per-pixel path,
    # and pixels deeper than the fallback's LDS capacity spill their keys to global memory
    sc = scenes.block_stack(32, 32, 8300, center_px=(12, 12), size_px=4.0, seed=13)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert gpu["stats"]["oversized_blocks"] >= 1


@pytest.mark.parametrize("window", [1, 4])
def test_ordering_violations_fallback(oracle_mod, window):
    """Steep intersecting planes: the block order is not the pixels' depth order. The wide
    window (4 slots) puts this scene's pixels in order; the narrow one leaves pixels to the
    exact per-pixel sort. Both match the oracle."""
    sc = scenes.tilted_planes(128, 96, 24)
    r, gpu, ref = _render_both(sc, oracle_mod, tuning={"window": window})
    _check(sc, gpu, ref, oracle_mod)
    if window == 1:
        assert gpu["stats"]["fallback_pixels"] > 0


def test_deep_pixels_spill(oracle_mod):
    sc = scenes.tilted_planes(48, 40, 600, seed=5)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert int(ref["frag_counts"].max()) > 512


def test_determinism(oracle_mod):
    from lucid_amd.renderer import GpuScene, LucidRenderer
    sc = scenes.tilted_planes(64, 64, 40, seed=9)
    r = LucidRenderer((sc.width, sc.height))
    g = GpuScene(sc)
    a = r.render(g, r.alloc_outputs(rgba32f=True))
    fa = a["image_f"].cpu().numpy().copy()
    b = r.render(g, r.alloc_outputs(rgba32f=True))
    assert np.array_equal(fa, b["image_f"].cpu().numpy())


def test_c2_icospheres_reduced(oracle_mod):
    """Mesh quads from the quad generator (two triangles per quad, shared vertices)."""
    sc = scenes.config2(grid=(4, 4, 4), width=640, height=360)
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)


def test_c2_full_size(oracle_mod):
    """The full mesh workload: 4096 icospheres (5.2M triangles in 2.6M generated quads) at
    2560x1440, alpha 0.5."""
    sc = scenes.config2()
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)


def _lit_scene(seed=43):
    """config0 with the shading features the benchmark scenes leave off: per-vertex normals
    (INST_HAS_VERTEX_NORMALS, packed 10:10:10), instance colours (INST_HAS_COLOR) and a sun
    with a tinted ambient, so finalShading takes its power path."""
    sc = scenes.config0(num_tris=3000, width=128, height=96, seed=seed)
    rng = np.random.default_rng(seed)
    n = rng.normal(size=(sc.num_verts, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    q = np.floor(512.0 + n * 511.0).astype(np.uint32) & 0x3FF
    sc.normals = (q[:, 0] | (q[:, 1] << 10) | (q[:, 2] << 20)).astype(np.uint32)
    sc.instances["flags"] |= scenes.INST_HAS_VERTEX_NORMALS | scenes.INST_HAS_COLOR
    c = rng.integers(64, 256, size=(len(sc.instances), 4)).astype(np.uint32)
    c[:, 3] = 255
    sc.instance_colors = (c[:, 0] | (c[:, 1] << 8) | (c[:, 2] << 16) | (c[:, 3] << 24)).astype(
        np.uint32)
    sun = np.array([0.3, -0.5, -0.8])
    sun /= np.linalg.norm(sun)
    sc.config = scenes.make_config(scenes.Camera(sc.width, sc.height), len(sc.instances),
                                   ambient=(0.3, 0.35, 0.4), ambient_power=0.6,
                                   sun_color=(1.0, 0.9, 0.7), sun_dir=tuple(sun),
                                   sun_power=1.2)
    sc.name = f"lit_{sc.name}"
    return sc


def test_normals_instance_colours_and_sun_lighting(oracle_mod):
    """Per-vertex normals, instance colours and sun lighting (shading.glsl:107-184,
    funcs.glsl:261-271) against the oracle: bins and counts bit-exact, colours within 1/255,
    in RGBA8 and fp32 samples."""
    sc = _lit_scene()
    for opts in (0, _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod)


def _flat_colour_scene(seed=44):
    """Frames that take the block rasterisers' flat colour planes (RGBA8 samples, unit
    lighting, fp32 vertex colours with a per-triangle alpha), with every input those planes
    fold in: instances with vertex and instance colours, instance colours only (a constant
    plane), vertex colours only, translucent instance colours, and triangles of alpha 0."""
    sc = scenes.config0(num_tris=5000, width=128, height=96, seed=seed)
    rng = np.random.default_rng(seed)
    assert len(sc.instances) == 5
    flags = [scenes.INST_HAS_VERTEX_COLORS | scenes.INST_HAS_COLOR, scenes.INST_HAS_COLOR,
             scenes.INST_HAS_VERTEX_COLORS, scenes.INST_HAS_VERTEX_COLORS | scenes.INST_HAS_COLOR,
             scenes.INST_HAS_VERTEX_COLORS]
    sc.instances["flags"] = np.array(flags, np.uint32)
    c = rng.integers(0, 256, size=(5, 4)).astype(np.uint32)
    c[:, 3] = [255, 160, 255, 90, 255]
    sc.instance_colors = (c[:, 0] | (c[:, 1] << 8) | (c[:, 2] << 16) | (c[:, 3] << 24)).astype(
        np.uint32)
    ta = sc.tri_alpha.copy()
    ta[0::2][rng.random(sc.num_quads) < 0.1] = 0.0
    sc.tri_alpha = ta
    sc.name = f"flat_{sc.name}"
    return sc


def test_flat_colour_planes(oracle_mod):
    """The flat colour-plane records (colour = K + L(px, py) / inverse depth per channel,
    instance colour folded in) against the oracle's barycentric shading: bins and counts
    bit-exact, colours within 1/255; the same scene with fp32 samples (general records)
    as well."""
    sc = _flat_colour_scene()
    for opts in (0, _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod)


def _near_plane_scene(seed):
    """config0 with a fifth of the triangles moved so that one vertex lies between the
    near plane and the camera or behind it (view z in [-0.06, 2]): setup's z-near clipped
    bounding box (quad_setup.glsl:77-128) and the scanline setup of crossing triangles."""
    sc = scenes.config0(num_tris=1500, width=96, height=96, seed=seed)
    rng = np.random.default_rng(seed)
    t = sc.num_quads
    pos = sc.positions.reshape(t, 3, 3).copy()
    k = rng.integers(0, 3, t)
    z = rng.uniform(-0.06, 2.0, t)
    for i in np.flatnonzero(rng.random(t) < 0.2):
        pos[i, k[i], 2] = z[i]
    sc.positions = pos.reshape(-1, 3).astype(np.float32)
    sc.name = f"near_plane_{seed}"
    return sc


@pytest.mark.parametrize("seed", [51, 52])
def test_triangles_crossing_the_near_plane(oracle_mod, seed):
    """Parity with the oracle's restatement of the reference's near-plane handling (bins,
    lists and counts bit-exact, colours within 1/255). The oracle's clipping itself is not
    checked against an independent model (its coverage pins use projected edges)."""
    sc = _near_plane_scene(seed)
    for opts in (0, _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod)


def test_unsorted_baseline(oracle_mod):
    """LUCID_OPT_UNSORTED (the unsorted baseline): bins, lists and fragment counts are the
    exact frame's. With additive blending (order-independent up to fp32 rounding) the
    colours match the oracle's exact frame; with "over" blending on two overlapping planes
    every pixel is the composite of one of the two orders (no sort, no mixing)."""
    sc = scenes.config0(num_tris=1500, width=128, height=96)
    r, gpu, ref = _render_both(sc, oracle_mod,
                               opts=_lib.OPT_UNSORTED | _lib.OPT_ADDITIVE_BLENDING)
    _check(sc, gpu, ref, oracle_mod)
    layers = [(9.0, (1.0, 0.0, 0.0), 0.5), (3.0, (0.0, 0.0, 1.0), 0.5)]
    sc = scenes.planes_scene(96, 64, layers)
    r, gpu, ref = _render_both(sc, oracle_mod, opts=_lib.OPT_UNSORTED | _lib.OPT_FP32_SAMPLES)
    assert np.array_equal(gpu["frag_counts"], ref["frag_counts"])
    orders = []
    for seq in (layers, layers[::-1]):
        c, t = np.zeros(3), 1.0
        for _, col, a in seq:
            c += np.array(col) * a * t
            t *= 1 - a
        orders.append(c)
    img = gpu["image_f"][8:56, 8:88, :3].reshape(-1, 3)
    near = [np.abs(img - o).max(axis=1) < 1e-5 for o in orders]
    assert (near[0] | near[1]).all()


def test_c1_full(oracle_mod):
    """The headline workload at full size: 1M soup triangles at 1920x1080 in 64 Gaussian blobs
    (~68M fragments, ~645 high bins, depth-range splitting and pre-sorted long blocks,
    raster_high.glsl:287-369). Bin counts, per-bin lists and per-pixel fragment counts
    bit-exact; colour within 1/255."""
    sc = scenes.config1()
    r, gpu, ref = _render_both(sc, oracle_mod)
    _check(sc, gpu, ref, oracle_mod)
    assert gpu["stats"]["high_bins"] >= 500
    print(f"[c1] fragments {ref['num_fragments']} max per pixel "
          f"{int(ref['frag_counts'].max())}", flush=True)


def test_frames_in_flight_share_buffers(oracle_mod):
    """Two different scenes queued back to back without waiting in between, on a side stream
    and then split across two streams: the second frame is ordered behind the first (they
    share the renderer's buffers and instance table), both images match the oracle, and an
    overflow of the first frame is reported by lucid_finish although the second frame's
    status replaced its own."""
    import torch
    from lucid_amd.renderer import GpuScene, LucidRenderer
    sa = scenes.tilted_planes(48, 40, 600, seed=5)  # deep pixels: overflows a fresh renderer
    sb = scenes.config0(num_tris=800, width=48, height=40, seed=7)
    ga, gb = GpuScene(sa), GpuScene(sb)
    ref_a, ref_b = oracle_mod.render(sa), oracle_mod.render(sb)
    # does the first scene overflow a fresh renderer on its own?
    solo = LucidRenderer((48, 40))
    solo.render_async(ga, solo.alloc_outputs())
    a_overflows = solo.finish() == _lib.LUCID_ERR_CAPACITY
    print(f"\n[case] in-flight pair; first scene overflows a fresh renderer: {a_overflows}",
          flush=True)
    side = torch.cuda.Stream()
    r = LucidRenderer((48, 40))
    oa, ob = r.alloc_outputs(rgba32f=True), r.alloc_outputs(rgba32f=True)
    r.render_async(ga, oa, stream=side)
    r.render_async(gb, ob, stream=side)
    rc = r.finish()
    assert rc == (_lib.LUCID_ERR_CAPACITY if a_overflows else _lib.LUCID_OK), rc
    # Same instance table and quad count, nothing visible (every quad behind the camera):
    # nothing is re-uploaded or reallocated, so the second frame is queued while the first
    # may still run, and the first one's overflow reaches the host through the carry word.
    import dataclasses
    pos = sa.positions.copy()
    pos[:, 2] = np.abs(pos[:, 2]) + 1.0
    sc_hidden = dataclasses.replace(sa, positions=pos, name="hidden")
    gh = GpuScene(sc_hidden)
    r2 = LucidRenderer((48, 40))
    o2a, o2b = r2.alloc_outputs(), r2.alloc_outputs()
    r2.render_async(ga, o2a, stream=side)
    r2.render_async(gh, o2b, stream=side)
    rc = r2.finish()
    assert rc == (_lib.LUCID_ERR_CAPACITY if a_overflows else _lib.LUCID_OK), rc
    for _ in range(4):
        r2.render_async(ga, o2a, stream=side)
        r2.render_async(gh, o2b, stream=side)
        if r2.finish() == _lib.LUCID_OK:
            break
    torch.cuda.synchronize()
    assert np.array_equal(o2a["frag_counts"].cpu().numpy().astype(np.uint32), ref_a["frag_counts"])
    assert int(o2b["frag_counts"].sum()) == 0
    for streams in ((side, side), (side, torch.cuda.current_stream())):
        for _ in range(4):
            r.render_async(ga, oa, stream=streams[0])
            r.render_async(gb, ob, stream=streams[1])
            rc = r.finish()
            if rc == _lib.LUCID_OK:
                break
            assert rc == _lib.LUCID_ERR_CAPACITY
        assert rc == _lib.LUCID_OK
        torch.cuda.synchronize()
        for o, ref in ((oa, ref_a), (ob, ref_b)):
            cnt = o["frag_counts"].cpu().numpy().astype(np.uint32)
            assert np.array_equal(cnt, ref["frag_counts"])
            d = np.abs(o["image"].cpu().numpy().astype(np.int32) - ref["image"].astype(np.int32))
            assert d.max() <= 1


def _camera_pose():
    rot = scenes.look_at((0.0, 0.0, 0.0), (0.3, -0.2, -1.0), up=(0.2, 1.0, 0.1))
    return (37.5, -12.25, 54.0), rot


@pytest.mark.parametrize("cull", [False, True])
def test_general_camera_and_backface_culling(oracle_mod, cull):
    """A translated and rotated camera (frustum origin ~67 units from the world origin,
    world-space corner rays) with and without back-face culling (quad_setup.glsl:155-170):
    the frustum-origin subtraction in setup, the cull and the rotated rays against the
    oracle; bins, lists and counts bit-exact, colours within 1/255."""
    base = scenes.config0(num_tris=4000, width=160, height=128, seed=81, vertex_sigma=0.08)
    sc = scenes.posed(base, *_camera_pose(), backface_culling=cull,
                      name=f"posed_cull{int(cull)}")
    for opts in (0, _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod)
        if cull:
            assert gpu["stats"]["rejected"]["backface"] > 1000


def test_setup_records_match_oracle(oracle_mod):
    """lucid_read_setup returns the per-triangle setup records in the reference's storage
    order (bary planes, scan parameters, depth plane + flags, quad_setup.glsl:274-340) and
    the per-quad AABB words: bit-exact against the oracle's for every visible triangle, on
    soup quads, mesh quads (both triangles of a quad), a posed camera with culling and
    triangles crossing the near plane."""
    from lucid_amd.renderer import GpuScene, LucidRenderer
    cases = [scenes.config0(num_tris=3000, width=128, height=96, seed=82),
             scenes.config2(grid=(3, 3, 3), width=320, height=192),
             scenes.posed(scenes.config0(num_tris=2000, width=128, height=96, seed=83),
                          *_camera_pose(), backface_culling=True, name="posed_cull"),
             _near_plane_scene(84)]
    for sc in cases:
        print(f"\n[case] setup records {sc.name} quads={sc.num_quads}", flush=True)
        r = LucidRenderer((sc.width, sc.height))
        r.render(GpuScene(sc))
        recs, aabb = r.read_setup(sc.num_quads)
        ref = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
        assert np.array_equal(aabb, ref["quad_aabb"]), "quad AABB words"
        vis = aabb != 0xFFFFFFFF
        live = np.zeros(2 * sc.num_quads, bool)
        live[0::2] = vis & ((aabb >> 30) & 1 == 0)
        live[1::2] = vis & ((aabb >> 31) & 1 == 0)
        assert live.sum() > 0
        bad = np.argwhere((recs[live] != ref["tri_records"][live]).any(axis=1))
        assert len(bad) == 0, f"{len(bad)} triangles differ, first {np.flatnonzero(live)[bad[:3, 0]]}"
