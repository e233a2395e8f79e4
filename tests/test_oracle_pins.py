"""Pins of the oracle functions that the first set of pins left open (CPU, no GPU needed):

  * a general camera: the frustum origin subtracted in setup (quad_setup.glsl:155-170,
    274-340) and world-space corner rays, under a translated and rotated camera, against
    float64 pixel-centre coverage and a float64 per-pixel sort of the image;
  * the back-face cull (quad_setup.glsl:155-170): rejected count and coverage against the
    float64 sign of det[v0 - eye, v1 - eye, v2 - eye];
  * triangles crossing the near plane or reaching behind the camera: coverage against a
    float64 model of what the homogeneous edge setup computes, the pixel-centre ray from the
    eye hitting the triangle in front of the eye (the ray lies in the cone spanned by the
    three vertex directions);
  * large-triangle bin lists (bin_counter.glsl:82-134): every listed bin overlaps the
    float64 screen triangle dilated by one pixel (an upper bound to pair with the existing
    lower bound "every bin with a covered pixel lists the triangle");
  * the between-samples reject (quad_setup.glsl:227-231) against its closed form on quads
    of known sub-pixel extent.

Each check is built so that a plausible slip fails it: a dropped origin term moves every
triangle by the camera offset, a flipped cull sign rejects the complementary set, an
unclipped near-plane triangle covers the mirrored cone behind the eye, and an over-counting
bin rule lists bins the triangle never reaches.
"""

import numpy as np
import pytest

from lucid_amd import scenes

from test_oracle import _brute_force_image, _pixel_centre_coverage

INFO = 1152


def _pose(seed):
    """A camera 30-60 units from the origin, rotated about all three axes."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-60.0, 60.0, 3)
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    ang = rng.uniform(0.5, 2.5)
    k = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    rot = np.eye(3) + np.sin(ang) * k + (1 - np.cos(ang)) * (k @ k)
    return pos, rot


def _tri_list(scene):
    """(triangle id, vertex ids) of every triangle not culled by its indices."""
    out = []
    for qi, q in enumerate(scene.quad_indices):
        for s, t in enumerate(((q[0], q[1], q[2]), (q[0], q[2], q[3]))):
            if len({int(v) for v in t}) == 3:
                out.append((2 * qi + s, tuple(int(v) for v in t)))
    return out


# ----------------------------------------------------------------------------- camera


@pytest.mark.parametrize("seed", [61, 62])
def test_general_camera_coverage(oracle_mod, seed):
    base = scenes.config0(seed=seed, num_tris=300, width=96, height=80, vertex_sigma=0.15)
    sc = scenes.posed(base, *_pose(seed))
    assert np.abs(sc.camera.eye).max() > 10.0
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    count, margin = _pixel_centre_coverage(sc)
    away = margin > 2e-3
    assert away.mean() > 0.95
    diff = res["frag_counts"].astype(np.int64) - count
    assert np.abs(diff[away]).max() == 0, f"{(diff[away] != 0).sum()} pixels differ"
    # the same frame as seen by the default camera, up to centres on edges
    ref = oracle_mod.render(base, mode=oracle_mod.MODE_COUNTS)
    assert np.array_equal(res["frag_counts"][away], ref["frag_counts"][away])


def test_general_camera_blending(oracle_mod):
    base = scenes.config0(seed=63, num_tris=120, width=64, height=48, vertex_sigma=0.25)
    sc = scenes.posed(base, *_pose(63))
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES)
    ref = _brute_force_image(sc)
    count, margin = _pixel_centre_coverage(sc)
    ok = (margin > 2e-3) & (res["frag_counts"] == count)
    assert ok.mean() > 0.95
    d = np.abs(res["image_f"][..., :3] - ref)[ok]
    assert d.max() < 2e-3, d.max()


# ----------------------------------------------------------------------------- back faces


def _facing(scene, tri):
    """float64 det[v0 - eye, v1 - eye, v2 - eye] and its scale (product of the lengths)."""
    cam = scenes.scene_camera(scene)
    p = scene.positions[list(tri)].astype(np.float64) - cam.eye
    return float(np.linalg.det(p)), float(np.prod(np.linalg.norm(p, axis=1)))


@pytest.mark.parametrize("posed", [False, True])
def test_backface_cull_closed_form(oracle_mod, posed):
    """The reference keeps a triangle when dot(p0, cross(p2, p1 - p2)) > 0 with p relative
    to the frustum origin, i.e. det[p0, p1, p2] < 0. Soup quads (a, b, c, c) are rejected as
    back faces exactly when their first triangle is. Coverage with the cull on equals the
    pixel-centre coverage of the front-facing triangles alone."""
    base = scenes.config0(seed=64, num_tris=400, width=96, height=80, vertex_sigma=0.15)
    sc = scenes.posed(base, *_pose(64), backface_culling=True) if posed else \
        scenes.posed(base, (0.0, 0.0, 0.0), np.eye(3), backface_culling=True)
    # triangles seen nearly edge-on (|det| within float32 rounding of 0) are left out
    tris = [(i, t) for i, t in _tri_list(sc) if abs(_facing(sc, t)[0]) > 1e-5 * _facing(sc, t)[1]]
    assert len(tris) > 380
    sc = dataclasses_replace_quads(sc, [t for _, t in tris])
    tris = _tri_list(sc)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    dets = [_facing(sc, t) for _, t in tris]
    back = [d >= 0.0 for d, _ in dets]
    assert 0.3 < np.mean(back) < 0.7
    assert int(res["info"][33]) == sum(back), "num_rejected_quads[BACKFACE]"
    front = dataclasses_replace_quads(sc, [t for (i, t), b in zip(tris, back) if not b])
    count, margin = _pixel_centre_coverage(front)
    away = margin > 2e-3
    diff = res["frag_counts"].astype(np.int64) - count
    assert np.abs(diff[away]).max() == 0
    # without the cull every triangle counts
    import dataclasses
    cam = sc.camera
    sc_all = dataclasses.replace(sc, config=scenes.make_config(cam, len(sc.instances)))
    all_count = oracle_mod.render(sc_all, mode=oracle_mod.MODE_COUNTS)["frag_counts"]
    assert int(all_count.sum()) > int(res["frag_counts"].sum())


def dataclasses_replace_quads(scene, tris):
    """The scene with only the given triangles, as degenerate quads (a, b, c, c)."""
    import dataclasses
    q = np.array([[a, b, c, c] for a, b, c in tris], np.uint32).reshape(-1, 4)
    inst = scenes.split_instances(len(q), scenes.INST_HAS_VERTEX_COLORS)
    return dataclasses.replace(scene, quad_indices=q, instances=inst)


# ----------------------------------------------------------------------------- near plane


def _cone_coverage(scene):
    """float64: the pixel-centre ray from the eye hits the triangle in front of the eye,
    i.e. its direction is a non-negative combination of the three vertex directions; plus
    the distance (pixels, on the image plane) of the centre to the nearest edge plane."""
    cam = scenes.scene_camera(scene)
    d = cam.pixel_rays()  # (H, W, 3), view z = -1
    h, w = scene.height, scene.width
    px_per_unit = h * cam.focal / 2.0
    count = np.zeros((h, w), np.int64)
    margin = np.full((h, w), np.inf)
    for _, tri in _tri_list(scene):
        V = cam.world_to_view(scene.positions[list(tri)].astype(np.float64))
        M = V.T  # columns: vertex directions
        if abs(np.linalg.det(M)) < 1e-12:
            continue
        x = np.einsum("ij,hwj->hwi", np.linalg.inv(M), d)
        count += (x >= 0).all(axis=-1)
        for a, b in ((0, 1), (1, 2), (2, 0)):
            n = np.cross(V[a], V[b])
            nn = np.hypot(n[0], n[1])
            if nn > 0:
                margin = np.minimum(margin, np.abs(d @ n) / nn * px_per_unit)
    return count, margin


def _near_plane_soup(seed, width=96, height=80, n=300):
    """Soup triangles with one vertex moved close to the camera (view z in (-0.06, 0))
    or behind it (view z in (0, 3))."""
    sc = scenes.config0(seed=seed, num_tris=n, width=width, height=height, vertex_sigma=0.2)
    rng = np.random.default_rng(seed)
    pos = sc.positions.reshape(n, 3, 3).copy()
    k = rng.integers(0, 3, n)
    for i in np.flatnonzero(rng.random(n) < 0.5):
        pos[i, k[i], 2] = rng.choice([rng.uniform(-0.06, -0.001), rng.uniform(0.01, 3.0)])
    import dataclasses
    return dataclasses.replace(sc, positions=pos.reshape(-1, 3).astype(np.float32),
                               name=f"near_soup_{seed}")


@pytest.mark.parametrize("seed,posed", [(71, False), (72, False), (73, True)])
def test_near_plane_coverage_cone_model(oracle_mod, seed, posed):
    sc = _near_plane_soup(seed)
    if posed:
        sc = scenes.posed(sc, *_pose(seed))
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    count, margin = _cone_coverage(sc)
    away = margin > 2e-3
    assert away.mean() > 0.9
    diff = res["frag_counts"].astype(np.int64) - count
    bad = np.argwhere((diff != 0) & away)
    assert len(bad) == 0, f"{len(bad)} pixels differ, first {bad[:5]} {diff[tuple(bad[0])]}"
    # the crossing triangles do reach the frame
    cam = scenes.scene_camera(sc)
    vz = cam.world_to_view(sc.positions.astype(np.float64))[:, 2].reshape(-1, 3)
    crossing = ((vz > -1.0 / 16.0).any(axis=1)) & ((vz < -1.0 / 16.0).any(axis=1))
    assert crossing.sum() > 50
    assert int(res["info"][1]) + int(res["info"][2]) > 0


# ----------------------------------------------------------------------------- large tris


def _screen_tri(scene, tri):
    cam = scenes.scene_camera(scene)
    vp = cam.projection() @ cam.view_matrix()
    h, w = scene.height, scene.width
    out = []
    for v in tri:
        c = vp @ np.append(scene.positions[v].astype(np.float64), 1.0)
        out.append(((c[0] / c[3] + 1) * w / 2, (c[1] / c[3] + 1) * h / 2))
    return np.array(out)


def _clip_poly(poly, x0, y0, x1, y1):
    """Sutherland-Hodgman clip of a convex polygon to an axis-aligned rectangle."""
    def clip(pts, inside, cut):
        out = []
        for i in range(len(pts)):
            a, b = pts[i - 1], pts[i]
            ia, ib = inside(a), inside(b)
            if ib:
                if not ia:
                    out.append(cut(a, b))
                out.append(b)
            elif ia:
                out.append(cut(a, b))
        return out

    def cx(a, b, x):
        t = (x - a[0]) / (b[0] - a[0])
        return (x, a[1] + t * (b[1] - a[1]))

    def cy(a, b, y):
        t = (y - a[1]) / (b[1] - a[1])
        return (a[0] + t * (b[0] - a[0]), y)

    pts = [tuple(p) for p in poly]
    pts = clip(pts, lambda p: p[0] >= x0, lambda a, b: cx(a, b, x0)) if pts else pts
    pts = clip(pts, lambda p: p[0] <= x1, lambda a, b: cx(a, b, x1)) if pts else pts
    pts = clip(pts, lambda p: p[1] >= y0, lambda a, b: cy(a, b, y0)) if pts else pts
    pts = clip(pts, lambda p: p[1] <= y1, lambda a, b: cy(a, b, y1)) if pts else pts
    return pts


def test_large_triangle_bins_upper_bound(oracle_mod):
    """Upper bound of the large-triangle bin rule (bin_counter.glsl:82-134 with
    loadScanlineParamsBin, scanline.glsl:28-52). In each bin row a triangle lists a
    contiguous run of bins; every bin of the run overlaps the triangle's float64 screen
    projection dilated by one pixel, except possibly the leftmost, whose right neighbour then
    does: the left-bound intercept is lowered by 31.011 + 0.99 and rounded down
    (int(xmin + 1) >> 5), one bin more than the overlap test needs whenever the intercept is
    not bin-aligned (a reference quirk the restatement keeps). With the lower bound of
    test_oracle (every bin holding a covered pixel lists the triangle) this brackets the rule
    from both sides; a rule listing the bounding box fails here."""
    sc = scenes.config0(seed=9, num_tris=600, width=256, height=160, vertex_sigma=0.35)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    info, nb = res["info"], res["bin_count"]
    cnt = info[INFO:]
    tc, to = cnt[3 * nb:4 * nb], cnt[4 * nb:5 * nb]
    bcx = (sc.width + 31) // 32
    assert int(tc.sum()) > 200, "the scene must have large triangles"
    verts = {i: t for i, t in _tri_list(sc)}
    rows = {}  # (tri, bin row) -> listed bin columns
    for b in range(nb):
        for t in res["bin_tris"][to[b]:to[b] + tc[b]]:
            rows.setdefault((int(t), b // bcx), []).append(b % bcx)

    def overlaps(poly, bx, by):
        x0, y0 = bx * 32, by * 32
        return bool(_clip_poly(poly, x0 - 1.0, y0 - 1.0, x0 + 33.0, y0 + 33.0))

    left_extra = 0
    for (t, by), bxs in rows.items():
        bxs = sorted(bxs)
        assert bxs == list(range(bxs[0], bxs[-1] + 1)), "a row's bins are contiguous"
        poly = _screen_tri(sc, verts[t])
        for bx in bxs[1:]:
            assert overlaps(poly, bx, by), (t, by, bx)
        if not overlaps(poly, bxs[0], by):
            left_extra += 1
            assert len(bxs) > 1 or overlaps(poly, bxs[0] + 1, by), (t, by, bxs)
    listed = int(tc.sum())
    # the quirk is real but bounded: at most one extra bin per row
    assert 0 < left_extra <= len(rows) < listed
    # a rule listing every bin of the bounding box would list many more
    aabb_bins = 0
    for t in {t for t, _ in rows}:
        p = _screen_tri(sc, verts[t])
        lo = np.clip(np.floor(p.min(0) / 32), 0, None)
        hi = np.minimum(np.floor(p.max(0) / 32), [bcx - 1, (sc.height + 31) // 32 - 1])
        aabb_bins += int(np.prod(hi - lo + 1))
    assert aabb_bins > listed + left_extra


# ----------------------------------------------------------------------------- between samples


def test_between_samples_reject_closed_form(oracle_mod):
    """A quad is rejected when ceil(min - 0.5001) == floor(max + 0.5001) on either screen
    axis, i.e. (away from the 1e-4 tolerance) when its bounding box contains no pixel-centre
    column or no pixel-centre row. Quads of known extent on both sides of the rule."""
    w, h = 64, 48
    cam = scenes.Camera(w, h)
    rng = np.random.default_rng(81)
    boxes, expect = [], []
    for i in range(200):
        cxp, cyp = rng.integers(4, w - 4), rng.integers(4, h - 4)
        # x extent: inside one pixel column gap, or straddling a centre
        kind = rng.integers(0, 4)
        fx = rng.uniform(0.05, 0.4)
        ex = rng.uniform(0.02, 0.08)
        if kind == 0:    # x between centres, y straddles -> rejected
            xs, ys = (cxp + 0.55 + fx * 0.5, cxp + 0.55 + fx * 0.5 + ex), (cyp + 0.3, cyp + 0.8)
        elif kind == 1:  # y between centres -> rejected
            xs, ys = (cxp + 0.3, cxp + 0.8), (cyp + 0.55 + fx * 0.5, cyp + 0.55 + fx * 0.5 + ex)
        elif kind == 2:  # both straddle -> kept
            xs, ys = (cxp + 0.45, cxp + 0.56), (cyp + 0.47, cyp + 0.53)
        else:            # wide box across a centre in x and y -> kept
            xs, ys = (cxp - 0.9, cxp + 1.7), (cyp + 0.2, cyp + 0.9)
        boxes.append((xs, ys))
        rej = lambda a, b: np.ceil(a - 0.5001) == np.floor(b + 0.5001)
        expect.append(bool(rej(*xs) or rej(*ys)))
    n = len(boxes)
    ndc = np.zeros((n, 3, 2))
    for i, ((xa, xb), (ya, yb)) in enumerate(boxes):
        px = np.array([[xa, ya], [xb, ya], [xa, yb]])
        ndc[i] = np.stack([px[:, 0] * 2 / w - 1, px[:, 1] * 2 / h - 1], -1)
    depth = np.full((n, 3), 5.0)
    rgb = np.full((n, 3, 3), 0.5, np.float32)
    alpha = np.full(n, 0.5, np.float32)
    sc = scenes.soup_scene(w, h, ndc, depth, rgb, alpha, camera=cam)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    assert int(res["info"][35]) == sum(expect), (int(res["info"][35]), sum(expect))
    kept = [i for i in range(n) if not expect[i]]
    assert int(res["info"][1]) + int(res["info"][2]) == len(kept)
    rejected = np.array([res["quad_aabb"][i] == 0xFFFFFFFF for i in range(n)])
    assert np.array_equal(rejected, np.array(expect))
