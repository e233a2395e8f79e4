"""Pins the C oracle (CPU, no GPU needed).

The reference ships no tests, fixtures or golden vectors, so each part of the oracle is pinned
by something the mathematics fixes, chosen so that a plausible slip (dropped term, wrong sign
or index, off-by-one span) fails:
  * encoders against closed forms (funcs.glsl:28-59,100-129);
  * conversion semantics (truncate, saturate, NaN -> 0);
  * per-pixel coverage against a float64 pixel-centre edge test;
  * bin counts against per-quad brute force and the verifyInfo invariants
    (src/lucid_renderer.cpp:650-680);
  * exact blending against a float64 per-pixel sort and against closed-form layer stacks;
  * the reference's 3-deep window (REF3) agrees with the exact sort where no disorder exists
    and disagrees where it does;
  * promotion of low bins with more than 256 triangles in an 8x8 block.
"""

import hashlib
import json
import os

import numpy as np
import pytest

from lucid_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
INFO = 1152


# ----------------------------------------------------------------------------- encoders


def test_encode_aabb28_closed_form(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x0, y0, x1, y1 = (int(v) for v in rng.integers(0, 128, 4))
        enc = L.oracle_encode_aabb28(x0, y0, x1, y1)
        assert enc == x0 | (y0 << 7) | (x1 << 14) | (y1 << 21)
        assert (enc & 0x7f, (enc >> 7) & 0x7f, (enc >> 14) & 0x7f, (enc >> 21) & 0x7f) == \
            (x0, y0, x1, y1)
    assert L.oracle_encode_aabb28(200, 0, 0, 0) == 200 & 0x7f  # 7-bit fields


def test_encode_rgba8_truncates(oracle_mod):
    L = oracle_mod.lib()
    rng = np.random.default_rng(2)
    c = rng.uniform(0, 1, (4000, 4)).astype(np.float32)
    for v in c[:500]:
        enc = L.oracle_encode_rgba8(*[float(x) for x in v])
        q = np.floor(v * np.float32(255.0)).astype(np.uint32)  # float32 product, then trunc
        assert enc == int(q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24))
    # exact boundaries: k/255 * 255 may round below k in float32
    assert L.oracle_encode_rgba8(1.0, 0.0, 0.0, 1.0) == 0xff0000ff
    assert L.oracle_encode_rgba8(-0.1, 0.0, 0.0, 0.0) == 0  # negative saturates to 0


def test_encode_normal_uint(oracle_mod):
    L = oracle_mod.lib()
    for n in [(0, 0, 1), (1, 0, 0), (0, -1, 0), (0.6, 0.8, 0.0)]:
        enc = L.oracle_encode_normal_uint(*map(float, n))
        f = [np.float32(512.0) + np.float32(v) * np.float32(511.0) for v in n]
        expect = [int(x) & 0x3ff for x in f]
        assert enc == expect[0] | (expect[1] << 10) | (expect[2] << 20)


def test_conversions_saturate(oracle_mod):
    L = oracle_mod.lib()
    assert L.oracle_f2i(float("nan")) == 0
    assert L.oracle_f2i(float("inf")) == 2**31 - 1
    assert L.oracle_f2i(float("-inf")) == -2**31
    assert L.oracle_f2i(-3.7) == -3 and L.oracle_f2i(3.7) == 3
    assert L.oracle_f2u(-5.0) == 0 and L.oracle_f2u(float("nan")) == 0
    assert L.oracle_f2u(1e20) == 2**32 - 1 and L.oracle_f2u(7.9) == 7


# ----------------------------------------------------------------------------- coverage


def _pixel_centre_coverage(scene):
    """float64 brute force: pixel centres (x + 0.5, y + 0.5) inside each visible triangle,
    plus the distance of the centre to the nearest edge (in pixels)."""
    cam = scenes.scene_camera(scene)
    vp = cam.projection() @ cam.view_matrix()  # float64, row-major
    P = scene.positions.astype(np.float64)
    q = scene.quad_indices
    h, w = scene.height, scene.width
    ys, xs = np.mgrid[0:h, 0:w]
    cx, cy = xs + 0.5, ys + 0.5
    count = np.zeros((h, w), np.int64)
    margin = np.full((h, w), np.inf)
    tris = []
    for quad in q:
        tris.append((quad[0], quad[1], quad[2]))
        if quad[2] != quad[3]:
            tris.append((quad[0], quad[2], quad[3]))
    for t in tris:
        pts = []
        for v in t:
            p = np.append(P[v], 1.0)
            clip = vp @ p
            ndc = clip[:2] / clip[3]
            pts.append(((ndc[0] + 1) * w / 2, (ndc[1] + 1) * h / 2))
        (x0, y0), (x1, y1), (x2, y2) = pts
        area = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
        if area == 0:
            continue
        s = np.sign(area)
        e = []
        for (ax, ay), (bx, by) in [((x0, y0), (x1, y1)), ((x1, y1), (x2, y2)), ((x2, y2), (x0, y0))]:
            ln = np.hypot(bx - ax, by - ay)
            e.append(s * ((bx - ax) * (cy - ay) - (by - ay) * (cx - ax)) / ln)
        e = np.stack(e)
        inside = (e >= 0).all(axis=0)
        count += inside
        margin = np.minimum(margin, np.abs(e).min(axis=0))
    return count, margin


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_coverage_matches_pixel_centre_test(oracle_mod, seed):
    sc = scenes.config0(seed=seed, num_tris=300, width=96, height=80, vertex_sigma=0.15)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    count, margin = _pixel_centre_coverage(sc)
    away = margin > 1e-3  # pixel centres not on an edge
    assert away.mean() > 0.95
    diff = res["frag_counts"].astype(np.int64) - count
    assert np.abs(diff[away]).max() == 0, f"{(diff[away] != 0).sum()} pixels differ"
    # the grid total agrees up to centres lying exactly on edges
    assert abs(int(res["frag_counts"].sum()) - int(count.sum())) <= (~away).sum() * 2


# ----------------------------------------------------------------------------- bins


def test_bin_counts_brute_force_and_invariants(oracle_mod):
    sc = scenes.config0(seed=9, num_tris=2000, width=256, height=160, vertex_sigma=0.2)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    info, nb = res["info"], res["bin_count"]
    cnt = info[INFO:]
    qc, qo, qt = cnt[0:nb], cnt[nb:2 * nb], cnt[2 * nb:3 * nb]
    tc, to, tt = cnt[3 * nb:4 * nb], cnt[4 * nb:5 * nb], cnt[5 * nb:6 * nb]
    # verifyInfo invariants
    assert (qo[1:] == qo[:-1] + qc[:-1]).all() and (qt == qo + qc).all()
    assert (to[1:] == to[:-1] + tc[:-1]).all() and (tt == to + tc).all()
    # small quads: each bin of the AABB rectangle once
    bcx = (sc.width + 31) // 32
    expect_q = np.zeros(nb, np.int64)
    for a in res["quad_aabb"]:
        if a == 0xFFFFFFFF:
            continue
        x0, y0, x1, y1 = a & 0x7F, (a >> 7) & 0x7F, (a >> 14) & 0x7F, (a >> 21) & 0x7F
        if (x1 - x0 + 1) * (y1 - y0 + 1) <= 4:
            for by in range(y0, y1 + 1):
                expect_q[by * bcx + x0: by * bcx + x1 + 1] += 1
    assert np.array_equal(expect_q, qc.astype(np.int64))
    # large tris: every bin holding a covered pixel of the tri lists it (conservative test)
    ql, tl = oracle_mod.bin_lists(res)
    count, margin = _pixel_centre_coverage(sc)
    covered_bins = set()
    h, w = sc.height, sc.width
    for b in range(nb):
        x0, y0 = (b % bcx) * 32, (b // bcx) * 32
        if (count[y0:y0 + 32, x0:x0 + 32] > 0).any():
            covered_bins.add(b)
    nonempty = {b for b in range(nb) if qc[b] + tc[b] > 0}
    assert covered_bins <= nonempty
    # levels: low < 1024 tris (quads count twice), else high (bin_categorizer.glsl:345-359)
    ntri = tc.astype(np.int64) + 2 * qc.astype(np.int64)
    assert info[5 + 2] == ((ntri > 0) & (ntri < 1024)).sum()
    assert info[5 + 4] == (ntri >= 1024).sum()


# ----------------------------------------------------------------------------- blending


def test_planes_closed_form(oracle_mod):
    layers = [(3.0, (1.0, 0.0, 0.0), 0.5), (9.0, (0.0, 1.0, 0.0), 0.25),
              (5.0, (0.0, 0.0, 1.0), 0.75), (7.0, (1.0, 1.0, 1.0), 0.4)]
    sc = scenes.planes_scene(64, 48, layers)
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES)
    c, t = np.zeros(3), 1.0
    for _, col, a in sorted(layers):
        c += np.array(col) * a * t
        t *= 1 - a
    assert (res["frag_counts"] == len(layers)).all()
    img = res["image_f"][..., :3].reshape(-1, 3)
    assert np.abs(img - c).max() < 2e-6 * 10  # sRGB round trip in float32
    # a wrong blend order would give a different colour
    c2, t2 = np.zeros(3), 1.0
    for _, col, a in sorted(layers, reverse=True):
        c2 += np.array(col) * a * t2
        t2 *= 1 - a
    assert np.abs(c2 - c).max() > 0.1


def _srgb_to_linear(x):
    return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def _linear_to_srgb(y):
    return np.where(y <= 0.0031308, 12.92 * y, 1.055 * np.abs(y) ** (1.0 / 2.4) - 0.055)


def test_lit_planes_closed_form(oracle_mod):
    """The oracle's shading with instance colours and sun lighting (shading.glsl:107-184,
    funcs.glsl:261-271) against a float64 closed form on two screen-aligned planes: flat
    normal of each quad (quad_setup.glsl, cross(v0 - v2, v1 - v0), through the 10:10:10
    packing), light = max(0, dot(-sun, n) * 0.7 + 0.3), per-channel factor ambient *
    ambient_power + sun_colour * sun_power * light, colour = sRGB(linear(c * inst) * factor),
    then "over" front to back. A dropped term, a wrong normal sign or a swapped power would
    move the colour far beyond the tolerance."""
    layers = [(4.0, (0.8, 0.5, 0.2), 0.5), (6.0, (0.3, 0.6, 0.9), 0.6)]
    sc = scenes.planes_scene(64, 48, layers)
    inst_rgba = np.array([200, 150, 255, 255], np.uint32)
    sc.instances["flags"] |= scenes.INST_HAS_COLOR
    sc.instance_colors = np.full(len(sc.instances), inst_rgba[0] | (inst_rgba[1] << 8) |
                                 (inst_rgba[2] << 16) | (inst_rgba[3] << 24), np.uint32)
    sun = np.array([0.4, -0.3, -0.866])
    sun /= np.linalg.norm(sun)
    amb, amb_pow, sun_col, sun_pow = np.array([0.3, 0.35, 0.4]), 0.6, np.array([1.0, 0.9, 0.7]), 1.2
    sc.config = scenes.make_config(scenes.Camera(sc.width, sc.height), len(sc.instances),
                                   ambient=tuple(amb), ambient_power=amb_pow,
                                   sun_color=tuple(sun_col), sun_dir=tuple(sun),
                                   sun_power=sun_pow)
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES)
    inst = inst_rgba[:3].astype(np.float64) / 255.0
    c, t = np.zeros(3), 1.0
    for i, (_, col, a) in sorted(enumerate(layers), key=lambda x: x[1][0]):
        p = sc.positions[sc.quad_indices[i, :3]].astype(np.float32)
        n = np.cross(p[0] - p[2], p[1] - p[0]).astype(np.float64)
        n /= np.linalg.norm(n)
        q = np.floor(512.0 + n * 511.0).astype(np.int64) & 0x3FF
        n = (q - 512) * np.float64(np.float32(1.0 / 511.0))
        light = max(0.0, float(np.dot(-sun, n)) * 0.7 + 0.3)
        factor = amb * amb_pow + (sun_col * sun_pow) * light
        lit = np.clip(_linear_to_srgb(_srgb_to_linear(np.array(col) * inst) * factor), 0, 1)
        c += lit * a * t
        t *= 1 - a
    img = res["image_f"][16:32, 24:40, :3].reshape(-1, 3)
    assert np.abs(img - c).max() < 5e-5, (img[0], c)
    assert light != 0.3  # the sun term is live (not clamped away)


def _brute_force_image(scene):
    """float64 per-pixel sort + front-to-back blend with unquantised colours."""
    cam = scenes.scene_camera(scene)
    h, w = scene.height, scene.width
    ys, xs = np.mgrid[0:h, 0:w]
    cx, cy = xs + 0.5, ys + 0.5
    P = cam.world_to_view(scene.positions.astype(np.float64))
    f, a = cam.focal, cam.aspect
    frags = [[] for _ in range(h * w)]
    for qi, quad in enumerate(scene.quad_indices):
        for s, tri in enumerate([(quad[0], quad[1], quad[2]), (quad[0], quad[2], quad[3])]):
            if tri[1] == tri[2]:
                continue
            V = P[list(tri)]
            sx = ((V[:, 0] / -V[:, 2]) * f / a + 1) * w / 2
            sy = ((V[:, 1] * -f / -V[:, 2]) + 1) * h / 2
            area = (sx[1] - sx[0]) * (sy[2] - sy[0]) - (sx[2] - sx[0]) * (sy[1] - sy[0])
            if area == 0:
                continue
            l0 = ((sx[1] - cx) * (sy[2] - cy) - (sx[2] - cx) * (sy[1] - cy)) / area
            l1 = ((sx[2] - cx) * (sy[0] - cy) - (sx[0] - cx) * (sy[2] - cy)) / area
            l2 = 1 - l0 - l1
            inside = (l0 >= 0) & (l1 >= 0) & (l2 >= 0)
            # perspective-correct weights and depth
            iz = l0 / -V[0, 2] + l1 / -V[1, 2] + l2 / -V[2, 2]
            rgb = scene.colors_rgb[list(tri)].astype(np.float64)
            wgt = np.stack([l0 / -V[0, 2], l1 / -V[1, 2], l2 / -V[2, 2]]) / iz
            col = np.einsum("kyx,kc->yxc", wgt, rgb)
            alpha = scene.tri_alpha[2 * qi + s]
            for y, x in zip(*np.nonzero(inside)):
                frags[y * w + x].append((-iz[y, x], 2 * qi + s, col[y, x], alpha))
    img = np.zeros((h, w, 3))
    for p, lst in enumerate(frags):
        c, t = np.zeros(3), 1.0
        for _, _, col, al in sorted(lst, key=lambda e: (e[0], e[1])):
            c += np.clip(col, 0, 1) * al * t
            t *= 1 - al
        img[p // w, p % w] = c
    return img


def test_exact_blend_matches_float64_sort(oracle_mod):
    sc = scenes.config0(seed=21, num_tris=120, width=64, height=48, vertex_sigma=0.25)
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES)
    ref = _brute_force_image(sc)
    count, margin = _pixel_centre_coverage(sc)
    ok = (margin > 1e-3) & (res["frag_counts"] == count)
    assert ok.mean() > 0.95
    d = np.abs(res["image_f"][..., :3] - ref)[ok]
    assert d.max() < 2e-3, d.max()


def test_ref3_matches_exact_without_disorder(oracle_mod):
    sc = scenes.config0(num_tris=3000, width=128, height=128)
    ex = oracle_mod.render(sc)
    r3 = oracle_mod.render(sc, mode=oracle_mod.MODE_REF3)
    assert np.array_equal(ex["frag_counts"], r3["frag_counts"])
    assert np.abs(ex["image"].astype(int) - r3["image"].astype(int)).max() <= 1


def test_ref3_window_artifacts_on_tilted_planes(oracle_mod):
    sc = scenes.tilted_planes(128, 96, 24)
    ex = oracle_mod.render(sc)
    r3 = oracle_mod.render(sc, mode=oracle_mod.MODE_REF3)
    assert np.array_equal(ex["frag_counts"], r3["frag_counts"])
    d = np.abs(ex["image"].astype(int) - r3["image"].astype(int)).max(-1)
    assert (d > 1).sum() > 0  # the 3-deep window mis-orders some pixels


def test_promotion_rule(oracle_mod):
    sc = scenes.block_stack(64, 64, 300)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    info, nb = res["info"], res["bin_count"]
    ntri = info[INFO + 3 * nb:INFO + 4 * nb] + 2 * info[INFO:INFO + nb]
    promoted = [b for b in range(nb) if 0 < ntri[b] < 1024 and res["bin_levels"][b] == 4]
    assert promoted, "a low bin with > 256 tris in one 8x8 block must be promoted"
    assert info[5 + 4] == 1  # appended to the high list (raster_low.glsl:229-237)


# ----------------------------------------------------------------------------- golden digests


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_golden_c0(oracle_mod):
    """The seeded C0 scene and the oracle's outputs on it (tests/golden/c0_digests.json)."""
    with open(os.path.join(HERE, "golden", "c0_digests.json")) as f:
        golden = json.load(f)
    sc = scenes.config0()
    assert sc.digest() == golden["scene"]
    res = oracle_mod.render(sc)
    assert _digest(res["frag_counts"]) == golden["frag_counts"]
    assert _digest(res["info"][:INFO + 6 * res["bin_count"]]) == golden["info_counts"]
    assert int(res["num_fragments"]) == golden["num_fragments"]
    assert _digest(res["image"]) == golden["image"]
