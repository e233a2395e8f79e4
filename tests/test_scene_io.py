"""OBJ scenes (lucid_amd/scene_io.py): the parse of src/wavefront_obj.cpp against hand-worked
answers, the conversion's meshes, instances, colours and normals, the oracle's frame of an OBJ
scene against a float64 closed form, and (GPU) the HIP path against the oracle on an OBJ
scene."""

import numpy as np
import pytest

from lucid_amd import scene_io, scenes

MTL_TEXT = """# two materials; the repeated 'red' block is ignored
newmtl red
Kd 1 0 0
d 0.5
newmtl blue
Kd 0 0 1
map_Kd -bm 1 tex.png
newmtl red
Kd 0 1 0
d 0.25
"""

OBJ_TEXT = """# corners of a unit square
mtllib scene.mtl
mtllib scene.mtl
v 0 0 0
v 1 0 0
v 1 1 0 1.0
v 0 1 0
vt 0 0
vt 1 0.25
vn 0 0 1
usemtl red
f 1 2 3 4
f -4/1 -3/2 -2/2
usemtl missing
f 1//1 2//1 3//1
f 1/1/1 3/2/1 4/1/1
usemtl blue
"""


def _parse(text=OBJ_TEXT):
    calls = []

    def load_mtl(name, out):
        calls.append(name)
        scene_io.parse_mtl(MTL_TEXT, out)

    return scene_io.parse_obj(text, load_mtl), calls


def test_parse_obj_known_answer():
    """Worked by hand from src/wavefront_obj.cpp:67-261: each distinct (v, vt, vn) index
    triple is a new vertex in first-use order; polygons are fanned from their first corner;
    negative indices count back from the attributes read so far; `vt` v is flipped;
    attributes a triple leaves out read as zero; one `mtllib` load per distinct name; an
    unknown `usemtl` name maps to an appended "default" material; empty groups stay."""
    m, calls = _parse()
    assert calls == ["scene.mtl"]
    assert m.tris.tolist() == [[0, 1, 2], [0, 2, 3], [4, 5, 6], [7, 8, 9], [10, 11, 12]]
    assert len(m.positions) == 13
    corners = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]
    want_pos = corners + corners[:3] + corners[:3] + [corners[0], corners[2], corners[3]]
    assert m.positions.tolist() == want_pos
    uv = np.zeros((13, 2))
    uv[4], uv[5], uv[6] = (0, 1), (1, 0.75), (1, 0.75)
    uv[10], uv[11], uv[12] = (0, 1), (1, 0.75), (0, 1)
    assert np.array_equal(m.tex_coords, uv)
    n = np.zeros((13, 3))
    n[7:13] = (0, 0, 1)
    assert np.array_equal(m.normals, n)
    assert [mt.name for mt in m.materials] == ["red", "blue", "default"]
    red, blue, default = m.materials
    assert red.diffuse == (1.0, 0.0, 0.0) and red.dissolve == 0.5
    assert blue.diffuse == (0.0, 0.0, 1.0) and blue.maps == [("kd", "tex.png")]
    assert default.diffuse == (1.0, 1.0, 1.0) and default.dissolve == 1.0
    assert m.groups == [(0, 0, 3), (2, 3, 2), (1, 5, 0)]


def test_parse_obj_without_materials_and_limits():
    """No `usemtl`: one group of every triangle on material 0, after a "default" material
    is appended (src/wavefront_obj.cpp:213-215). A face of more than 64 corners is an
    error; a `vn`-less, `vt`-less file has neither array."""
    text = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nf 1 2 4 3\n"
    m = scene_io.parse_obj(text)
    assert m.normals is None and m.tex_coords is None
    assert m.tris.tolist() == [[0, 1, 2], [0, 2, 3]]
    assert [mt.name for mt in m.materials] == ["default"] and m.groups == [(0, 0, 2)]
    with pytest.raises(scene_io.ObjError):
        scene_io.parse_obj("v 0 0 0\nf " + " ".join(["1"] * 65) + "\n")
    with pytest.raises(scene_io.ObjError):
        scene_io.parse_obj("usemtl\n")


def test_encode_normals_and_material_colors():
    """encodeNormalUint (src/scene.cpp:338-343) on hand-worked values: 512 + n * 511
    truncated, x in the low bits; colours R lowest, alpha from the opacity."""
    q = scene_io.encode_normals(np.array([[0, 0, 1], [0, 0, 0], [-1, 0.5, 0]], np.float32))
    assert q.tolist() == [512 | 512 << 10 | 1023 << 20, 512 | 512 << 10 | 512 << 20,
                          1 | 767 << 10 | 512 << 20]
    assert scene_io.material_color((1, 1, 1), 1.0) == 0xFFFFFFFF
    assert scene_io.material_color((1, 0, 0), 0.5) == 0x800000FF
    assert scene_io.material_color((0.2, 2.0, -1.0), 0.25) == (51 | 255 << 8 | 0 << 16 | 64 << 24)


def test_obj_scene_meshes_instances_and_view():
    """Conversion of the hand-worked file: empty groups are dropped, each mesh's quads
    cover exactly its triangles, instance runs point at their mesh's quads, instance colours
    and flags follow the materials (a white opaque material gets no colour), every vertex
    lands inside the view after the fit, and scene_opacity scales alpha and clears
    INST_IS_OPAQUE."""
    m, _ = _parse()
    sc, draws = scene_io.obj_scene(m, 64, 48)
    assert [(d.material_id, d.num_tris) for d in draws] == [(0, 3), (2, 2)]
    for d, (mat, first, num) in zip(draws, [g for g in m.groups if g[2]]):
        q = sc.quad_indices[d.quad_offset:d.quad_offset + d.num_quads].astype(np.int64)
        got = [tuple(t) for t in q[:, :3]] + [(a, c, e) for a, _, c, e in q if c != e]
        norm = lambda t: tuple(np.roll(t, -int(np.argmin(t))))
        assert sorted(map(norm, got)) == sorted(map(norm, m.tris[first:first + num].tolist()))
    inst = sc.instances
    assert inst["index_offset"].tolist() == [4 * d.quad_offset for d in draws]
    assert inst["num_quads"].tolist() == [d.num_quads for d in draws]
    red_flags = scenes.INST_HAS_VERTEX_NORMALS | scenes.INST_HAS_COLOR
    white_flags = scenes.INST_HAS_VERTEX_NORMALS | scenes.INST_IS_OPAQUE
    assert inst["flags"].tolist() == [red_flags, white_flags]
    assert sc.instance_colors.tolist() == [0x800000FF, 0xFFFFFFFF]
    assert sc.normals.tolist() == scene_io.encode_normals(m.normals).tolist()
    cam = scenes.Camera(64, 48)
    vp = cam.view_proj().T.astype(np.float64)
    clip = np.c_[sc.positions, np.ones(len(sc.positions))] @ vp.T
    ndc = clip[:, :3] / clip[:, 3:]
    assert (clip[:, 3] > cam.near).all() and (np.abs(ndc[:, :2]) <= 1.0).all()
    sc2, _ = scene_io.obj_scene(m, 64, 48, scene_opacity=0.5)
    assert sc2.instances["flags"].tolist() == [red_flags, red_flags]
    assert sc2.instance_colors.tolist() == [0x400000FF, 0x80FFFFFF]


def test_obj_scene_flip_yz():
    """flip_yz swaps y and z of positions and normals and negates the normals
    (src/scene_convert.cpp:249-255)."""
    m, _ = _parse()
    sc, _ = scene_io.obj_scene(m, 64, 48, flip_yz=True, fit_view=False)
    sc0, _ = scene_io.obj_scene(m, 64, 48, fit_view=False)
    assert np.array_equal(sc.positions, sc0.positions[:, [0, 2, 1]])
    assert sc.normals.tolist() == scene_io.encode_normals(-m.normals[:, [0, 2, 1]]).tolist()


PLANES_OBJ = """mtllib planes.mtl
v -50 -50 -40
v 50 -50 -40
v 50 50 -40
v -50 50 -40
v -50 -50 -60
v 50 -50 -60
v 50 50 -60
v -50 50 -60
vn 0.3 0.2 0.93
usemtl back
f 5//1 6//1 7//1 8//1
usemtl front
f 1 2 3 4
"""

PLANES_MTL = """newmtl front
Kd 0.8 0.5 0.2
d 0.5
newmtl back
Kd 0.3 0.6 0.9
d 0.6
"""


def _srgb_to_linear(x):
    return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def _linear_to_srgb(y):
    return np.where(y <= 0.0031308, y * 12.92, 1.055 * y ** (1.0 / 2.4) - 0.055)


def test_obj_planes_closed_form(oracle_mod, tmp_path):
    """An OBJ file with two screen-covering squares (largest extent 100, so the rescale is
    exact) rendered by the oracle against a float64 closed form: colour = material colour
    as RGBA8 (alpha = dissolve), the back square lit by its `vn` normal through the
    10:10:10 packing, the front square's vertices left without a normal (decoded as zero,
    light 0.3, as the reference's scene-wide normal flag gives them), then "over" front to
    back. A wrong material, a dropped alpha, a missed normal or a flipped order moves the
    colour far beyond the tolerance."""
    (tmp_path / "planes.obj").write_text(PLANES_OBJ)
    (tmp_path / "planes.mtl").write_text(PLANES_MTL)
    sun = np.array([0.4, -0.3, -0.866])
    sun /= np.linalg.norm(sun)
    amb, amb_pow, sun_col, sun_pow = np.array([0.3, 0.35, 0.4]), 0.6, np.array([1.0, 0.9, 0.7]), 1.2
    sc, draws = scene_io.load_obj_scene(str(tmp_path / "planes.obj"), 64, 48, fit_view=False,
                                        ambient=tuple(amb), ambient_power=amb_pow,
                                        sun_color=tuple(sun_col), sun_dir=tuple(sun),
                                        sun_power=sun_pow)
    assert sc.name == "planes" and len(draws) == 2
    res = oracle_mod.render(sc, opts=oracle_mod.OPT_FP32_SAMPLES)
    c, t = np.zeros(3), 1.0
    for n, kd, d in [((0.0, 0.0, 0.0), (0.8, 0.5, 0.2), 0.5), ((0.3, 0.2, 0.93), (0.3, 0.6, 0.9), 0.6)]:
        q = scene_io.encode_normals(np.array([n], np.float32))[0]
        nd = np.array([((q >> s) & 0x3FF) - 512.0 for s in (0, 10, 20)]) * np.float64(np.float32(1.0 / 511.0))
        light = max(0.0, float(np.dot(-sun, nd)) * 0.7 + 0.3)
        factor = amb * amb_pow + (sun_col * sun_pow) * light
        rgba = np.floor(np.clip(np.array([*kd, d]), 0, 1) * 255.0 + 0.5) / 255.0
        lit = np.clip(_linear_to_srgb(_srgb_to_linear(rgba[:3]) * factor), 0, 1)
        c += lit * rgba[3] * t
        t *= 1 - rgba[3]
    img = res["image_f"][16:32, 24:40, :3].reshape(-1, 3)
    assert np.abs(img - c).max() < 5e-5, (img[0], c)
    assert (res["frag_counts"][16:32, 24:40] == 2).all()


def sphere_grid_obj(n=3, level=2, seed=5):
    """OBJ and MTL text: n x n icospheres with per-vertex normals (`f v//vn`, some faces
    with negative indices), one material each with varying colour and dissolve, and a
    square backdrop written as one polygon with a material the MTL lacks."""
    sp, st = scenes.icosphere(level)
    rng = np.random.default_rng(seed)
    obj, mtl = ["mtllib grid.mtl"], []
    base = 0
    for k in range(n * n):
        i, j = k % n, k // n
        centre = np.array([(i - (n - 1) / 2) * 2.2, (j - (n - 1) / 2) * 2.2, -rng.uniform(0, 3)])
        r = rng.uniform(0.8, 1.3)
        for p in sp:
            obj.append("v %.6f %.6f %.6f" % tuple(centre + p * r))
        for p in sp:
            obj.append("vn %.6f %.6f %.6f" % tuple(p / np.linalg.norm(p)))
        obj.append(f"usemtl m{k}")
        nv = len(sp)
        for f, (a, b, c) in enumerate(st.astype(np.int64).tolist()):
            if f % 3 == 0:  # relative to the attributes read so far
                a, b, c = a - nv, b - nv, c - nv
                obj.append(f"f {a}//{a} {b}//{b} {c}//{c}")
            else:
                a, b, c = a + 1 + base, b + 1 + base, c + 1 + base
                obj.append(f"f {a}//{a} {b}//{b} {c}//{c}")
        base += nv
        kd = rng.uniform(0.1, 1.0, 3)
        mtl += [f"newmtl m{k}", "Kd %.3f %.3f %.3f" % tuple(kd), "d %.3f" % rng.uniform(0.3, 0.9)]
    s = 1.5 * n
    for x, y in ((-s, -s), (s, -s), (s, s), (-s, s)):
        obj.append(f"v {x} {y} -6")
    obj.append("usemtl backdrop")
    obj.append("f -4 -3 -2 -1")
    return "\n".join(obj) + "\n", "\n".join(mtl) + "\n"


@pytest.mark.gpu
def test_obj_scene_gpu_parity(oracle_mod, tmp_path):
    """An OBJ scene (spheres with per-vertex normals, per-material instance colours and
    alpha, a backdrop on the default material, sun lighting) through the HIP path against
    the oracle: bins, lists and fragment counts bit-exact, colours within 1/255, RGBA8 and
    fp32 samples."""
    from test_gpu_parity import _check, _render_both
    from lucid_amd import _lib
    obj, mtl = sphere_grid_obj()
    (tmp_path / "grid.obj").write_text(obj)
    (tmp_path / "grid.mtl").write_text(mtl)
    sc, _ = scene_io.load_obj_scene(str(tmp_path / "grid.obj"), 160, 120, scene_opacity=0.9,
                                    sun_dir=(0.3, -0.5, -0.81), sun_power=1.0)
    for opts in (0, _lib.OPT_FP32_SAMPLES):
        r, gpu, ref = _render_both(sc, oracle_mod, opts=opts)
        _check(sc, gpu, ref, oracle_mod)


def test_row_parallel_edges_cover_pixel_centres(oracle_mod, tmp_path):
    """The OBJ sphere grid seen by an unrolled camera has edges exactly parallel to the pixel
    rows (mirror-symmetric vertices, so the edge's x coefficient is +-0). The oracle's
    fragment counts equal a float64 pixel-centre test away from edges; without the tilt of
    such edges their rows past the edge fill with the other two edges' spans."""
    from test_oracle import _pixel_centre_coverage
    obj, mtl = sphere_grid_obj()
    (tmp_path / "grid.obj").write_text(obj)
    (tmp_path / "grid.mtl").write_text(mtl)
    sc, _ = scene_io.load_obj_scene(str(tmp_path / "grid.obj"), 160, 120)
    res = oracle_mod.render(sc, mode=oracle_mod.MODE_COUNTS)
    count, margin = _pixel_centre_coverage(sc)
    away = margin > 1e-3
    assert away.mean() > 0.9
    diff = res["frag_counts"].astype(np.int64) - count
    assert (diff[away] == 0).all(), f"{int((diff[away] != 0).sum())} pixels differ"


def test_merge_vertices():
    """Scene::mergeVertices (src/scene.cpp:169-236): an icosphere written as a triangle soup
    (every triangle with its own three vertices, positions jittered below 10^-6) merges
    back to the sphere's vertex count, in first-appearance order, at the mean position,
    with averaged normals and the sphere's triangles; a mesh with texture coordinates is
    refused, a mesh with nothing to merge comes back unchanged."""
    sp, st = scenes.icosphere(2)
    sp = (np.round(sp.astype(np.float64) * 1e5) / 1e5).astype(np.float32)  # quantum centres
    rng = np.random.default_rng(3)
    soup = sp[st.reshape(-1)].astype(np.float32)
    soup = soup + rng.uniform(-4e-7, 4e-7, soup.shape).astype(np.float32)
    nrm = rng.normal(size=soup.shape).astype(np.float32)
    m = scene_io.ObjMesh(soup, nrm, None, np.arange(len(soup)).reshape(-1, 3),
                         [scene_io.ObjMaterial("default")], [(0, 0, len(st))])
    mm = scene_io.merge_vertices(m)
    assert len(mm.positions) == len(sp)
    first = {}
    for i, v in enumerate(st.reshape(-1).tolist()):
        first.setdefault(v, i)
    order = sorted(first, key=first.get)  # sphere vertex ids by first appearance
    assert mm.tris.tolist() == [[order.index(v) for v in t] for t in st.tolist()]
    for new_id, old_id in enumerate(order[:20]):
        members = np.flatnonzero(st.reshape(-1) == old_id)
        assert np.allclose(mm.positions[new_id], soup[members].mean(0), atol=1e-7)
        assert np.allclose(mm.normals[new_id], nrm[members].mean(0), atol=1e-6)
    assert scene_io.merge_vertices(mm) is mm
    with pytest.raises(scene_io.ObjError):
        scene_io.merge_vertices(_parse()[0])


def test_input_scenes_xml(tmp_path):
    """The input-scenes list format (src/scene_convert.cpp:34-63): root_path joined to each
    path, the name from the file stem unless given, defaults for missing parameters."""
    (tmp_path / "list.xml").write_text(
        '<scenes root_path="/data/scenes/">\n'
        '  <scene path="dragon2.obj" quad_squareness="1.5" merge_verts="true" />\n'
        '  <scene path="trees/tree01.obj" name="tree" flip_yz="true" />\n'
        '</scenes>\n')
    a, b = scene_io.load_input_scenes(str(tmp_path / "list.xml"))
    assert (a.name, a.path, a.quad_squareness, a.merge_verts, a.flip_yz) == \
        ("dragon2", "/data/scenes/dragon2.obj", 1.5, True, False)
    assert (b.name, b.path, b.quad_squareness, b.merge_verts, b.flip_yz) == \
        ("tree", "/data/scenes/trees/tree01.obj", 1.0, False, True)
    (tmp_path / "grid.obj").write_text(sphere_grid_obj(n=2, level=1)[0])
    (tmp_path / "grid.mtl").write_text(sphere_grid_obj(n=2, level=1)[1])
    (tmp_path / "local.xml").write_text('<scenes><scene path="grid.obj" merge_verts="1" /></scenes>')
    (e,) = scene_io.load_input_scenes(str(tmp_path / "local.xml"), root_path=str(tmp_path))
    sc, draws = e.load(64, 48)
    assert sc.name == "grid" and len(draws) == 5 and sc.num_quads > 0
