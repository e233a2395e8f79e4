"""Wavefront OBJ scenes for the renderer.

Parsing follows WavefrontObject::load and WavefrontMaterial::load
(src/wavefront_obj.cpp:10-261): one vertex per distinct (position, uv, normal) index triple,
1-based and negative (relative) indices, fan triangulation of polygons, `usemtl` groups,
`mtllib` files with `newmtl` / `Kd` / `d`. The conversion keeps the parts of convertScene
(src/scene_convert.cpp:241-445) that the exact-OIT path reads: the optional y/z flip, one mesh
per non-empty material group, quads per mesh (Scene::generateQuads, src/scene.cpp:237-247),
10:10:10 normals (encodeNormalUint, src/scene.cpp:338-343) and the rescale to a largest extent
of 100 (src/scene_convert.cpp:226-237). Instances and instance colours are built as
uploadInstances (src/lucid_renderer.cpp:352-429) builds them from Scene::draws
(src/scene.cpp:483-512), with the scene-opacity factor of LucidApp::drawScene
(src/lucid_app.cpp:646-650).

Vertex merging (Scene::mergeVertices, off by default in the reference's InputScene) is
`merge_vertices`, opt-in as there. Left out: decoding the MTL texture maps (their names
are parsed and kept; no image decoder is built, so OBJ scenes render untextured, while the
renderer samples textures a caller provides, `scenes.Texture`) and the vertex-cache
triangle reordering (src/tri_optimizer.cpp), which only permutes triangles within a mesh.
"""

from __future__ import annotations

import dataclasses
import math
import os
import re
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import scenes

MAX_FACE_INDICES = 64  # src/wavefront_obj.cpp:148

_INT_PREFIX = re.compile(r"\s*([+-]?\d+)")


def _atoi(text: str) -> int:
    """C atoi: the leading integer of `text`, 0 when there is none."""
    m = _INT_PREFIX.match(text)
    return int(m.group(1)) if m else 0


def _floats(tokens: List[str], n: int) -> List[float]:
    out = [0.0] * n
    for i in range(min(n, len(tokens))):
        out[i] = float(tokens[i])
    return out


@dataclasses.dataclass
class ObjMaterial:
    """WavefrontMaterial (src/wavefront_obj.h:13-19)."""

    name: str
    dissolve: float = 1.0
    diffuse: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    maps: List[Tuple[str, str]] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class ObjMesh:
    """WavefrontObject (src/wavefront_obj.h:22-37): per-vertex arrays, triangles and
    material groups (material_id, first_tri, num_tris)."""

    positions: np.ndarray  # (V, 3) float32
    normals: Optional[np.ndarray]  # (V, 3) float32, None without `vn` lines
    tex_coords: Optional[np.ndarray]  # (V, 2) float32, None without `vt` lines
    tris: np.ndarray  # (T, 3) int64
    materials: List[ObjMaterial]
    groups: List[Tuple[int, int, int]]


class ObjError(ValueError):
    pass


def parse_mtl(text: str, out: List[ObjMaterial]) -> int:
    """Appends the materials of one .mtl file to `out` and returns how many were added
    (src/wavefront_obj.cpp:10-65). A repeated `newmtl` name is ignored with the lines
    that follow it."""
    count = 0
    cur: Optional[ObjMaterial] = None
    for line in text.split("\n"):
        if not line or line[0] == "#":
            continue
        tok = line.split()
        if not tok:
            continue
        elem, args = tok[0], tok[1:]
        if elem == "newmtl":
            if not args:
                raise ObjError(f"newmtl without a name: {line!r}")
            if any(m.name == args[0] for m in out):
                cur = None
                continue
            cur = ObjMaterial(args[0])
            out.append(cur)
            count += 1
        if cur is None:
            continue
        if elem.startswith("map_"):
            if not args:
                raise ObjError(f"map without a file: {line!r}")
            cur.maps.append((elem[4:].lower(), args[-1]))
        elif elem == "d":
            cur.dissolve = _floats(args, 1)[0]
        elif elem == "Kd":
            cur.diffuse = tuple(_floats(args, 3))
    return count


def parse_obj(text: str, load_mtl: Optional[Callable[[str, List[ObjMaterial]], None]] = None) -> ObjMesh:
    """OBJ text -> ObjMesh (src/wavefront_obj.cpp:67-261). `load_mtl(name, materials)` is
    called once per distinct `mtllib` name after the parse; without it material libraries
    are skipped."""
    src = ([], [], [])  # positions, uvs, normals as read
    counts = [0, 0, 0]
    vertex_map = {}
    tris = []
    use_mtls = []
    libs = []

    def index(elem: str) -> int:
        idx = [_atoi(elem), 0, 0]
        s = elem.find("/")
        if s != -1:
            rest = elem[s + 1:]
            s2 = rest.find("/")
            if s2 == -1:
                idx[1] = _atoi(rest)
            elif s2 == 0:
                idx[2] = _atoi(rest[1:])
            else:
                idx[1] = _atoi(rest)
                idx[2] = _atoi(rest[s2 + 1:])
        key = tuple(i + counts[j] if i < 0 else i - 1 for j, i in enumerate(idx))
        v = vertex_map.get(key)
        if v is None:
            v = vertex_map[key] = len(vertex_map)
        return v

    for line in text.split("\n"):
        if not line or line[0] == "#":
            continue
        tok = line.split()
        if not tok:
            continue
        elem, args = tok[0], tok[1:]
        if elem == "v":
            src[0].append(_floats(args, 3))
            counts[0] += 1
        elif elem == "vt":
            uv = _floats(args, 2)
            uv[1] = 1.0 - uv[1]
            src[1].append(uv)
            counts[1] += 1
        elif elem == "vn":
            src[2].append(_floats(args, 3))
            counts[2] += 1
        elif elem == "f":
            if len(args) > MAX_FACE_INDICES:
                raise ObjError(f"too many face indices ({MAX_FACE_INDICES} is max): {line!r}")
            ids = [index(a) for a in args]
            for i in range(1, len(ids) - 1):
                tris.append((ids[0], ids[i], ids[i + 1]))
        elif elem == "usemtl":
            if not args:
                raise ObjError(f"usemtl without a name: {line!r}")
            use_mtls.append((len(tris), args[0]))
        elif elem == "mtllib":
            if args and args[0] not in libs:
                libs.append(args[0])

    materials: List[ObjMaterial] = []
    if load_mtl is not None:
        for lib in libs:
            load_mtl(lib, materials)

    nv = len(vertex_map)
    keys = np.array(list(vertex_map.keys()), np.int64).reshape(-1, 3)
    slots = np.array(list(vertex_map.values()), np.int64)

    def gather(j: int, width: int) -> np.ndarray:
        table = np.array(src[j], np.float32).reshape(-1, width)
        out = np.zeros((nv, width), np.float32)
        k = keys[:, j]
        ok = (k >= 0) & (k < counts[j])
        out[slots[ok]] = table[k[ok]]
        return out

    positions = gather(0, 3)
    tex_coords = gather(1, 2) if src[1] else None
    normals = gather(2, 3) if src[2] else None
    tri_arr = np.array(tris, np.int64).reshape(-1, 3)

    groups = []
    if not use_mtls:
        materials.append(ObjMaterial("default"))
        groups.append((0, 0, len(tri_arr)))
    else:
        default_idx = -1
        for i, (first, name) in enumerate(use_mtls):
            end = use_mtls[i + 1][0] if i + 1 < len(use_mtls) else len(tri_arr)
            mat = next((j for j, m in enumerate(materials) if m.name == name), -1)
            if mat == -1:
                if default_idx == -1:
                    default_idx = next((j for j, m in enumerate(materials) if m.name == "default"), -1)
                    if default_idx == -1:
                        default_idx = len(materials)
                        materials.append(ObjMaterial("default"))
                mat = default_idx
            groups.append((mat, first, end - first))
    return ObjMesh(positions, normals, tex_coords, tri_arr, materials, groups)


def load_obj(path: str) -> ObjMesh:
    """Reads an OBJ file and the .mtl libraries it names (resolved beside it)."""
    base = os.path.dirname(os.path.abspath(path))

    def load_mtl(name: str, out: List[ObjMaterial]) -> None:
        with open(os.path.join(base, name), "r", encoding="utf-8", errors="replace") as f:
            parse_mtl(f.read(), out)

    with open(path, "r", encoding="utf-8", errors="replace") as f:
        return parse_obj(f.read(), load_mtl)


def merge_vertices(mesh: ObjMesh, decimal_places: int = 5) -> ObjMesh:
    """Scene::mergeVertices (src/scene.cpp:169-236): vertices whose positions round to the
    same multiple of 10^-decimal_places become one, at the mean of their positions (and
    normals, when there are any), in order of first appearance; triangles are re-indexed.
    Sums run in float32 in vertex order, then are scaled by float32(1 / count). The
    reference merges only meshes without texture coordinates (src/scene_convert.cpp:359)."""
    if mesh.tex_coords is not None:
        raise ObjError("merge_vertices needs a mesh without texture coordinates")
    pos = mesh.positions.astype(np.float32)
    scale = np.float32(10.0 ** decimal_places)
    q = np.round(pos * scale)  # round half to even; the reference's rounding is libfwk's
    _, first, inverse = np.unique(q, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    order = np.argsort(first, kind="stable")  # unique rows -> first-appearance order
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    new_idx = rank[inverse]
    n = len(order)
    if n == len(pos):
        return mesh
    count = np.bincount(new_idx, minlength=n).astype(np.float32)
    inv = (np.float32(1.0) / count)[:, None]

    def mean(a):
        acc = np.zeros((n, 3), np.float32)
        np.add.at(acc, new_idx, a.astype(np.float32))
        return (acc * inv).astype(np.float32)

    normals = None if mesh.normals is None else mean(mesh.normals)
    return dataclasses.replace(mesh, positions=mean(pos), normals=normals,
                               tris=new_idx[mesh.tris].astype(np.int64))


def encode_normals(normals: np.ndarray) -> np.ndarray:
    """encodeNormalUint (src/scene.cpp:338-343): each component to uint(512 + n * 511)
    in double precision (truncated), 10 bits each, x lowest."""
    q = np.trunc(512.0 + normals.astype(np.float64) * 511.0).astype(np.int64) & 0x3FF
    return (q[:, 0] | (q[:, 1] << 10) | (q[:, 2] << 20)).astype(np.uint32)


def material_color(diffuse, opacity: float) -> int:
    """IColor(FColor(diffuse, opacity)) as uploadInstances packs it
    (src/lucid_renderer.cpp:364-365): RGBA8, R lowest. The float-to-byte rounding is
    libfwk's (not in the reference tree), restated as clamp to [0, 1], times 255, round to
    nearest."""
    c = np.clip(np.array([*diffuse, opacity], np.float64), 0.0, 1.0)
    b = np.floor(c * 255.0 + 0.5).astype(np.uint32)
    return int(b[0] | (b[1] << 8) | (b[2] << 16) | (b[3] << 24))


@dataclasses.dataclass
class MeshDraw:
    """One mesh of the converted scene: its material and its quads' range."""

    material_id: int
    quad_offset: int
    num_quads: int
    num_tris: int


def obj_scene(
    mesh: ObjMesh,
    width: int,
    height: int,
    *,
    squareness: float = 1.0,
    flip_yz: bool = False,
    scene_opacity: float = 1.0,
    fit_view: bool = True,
    name: str = "obj",
    **lighting,
) -> Tuple[scenes.Scene, List[MeshDraw]]:
    """ObjMesh -> Scene with one instance run per mesh (split at 1024 quads).

    `squareness` is InputScene::quad_squareness (default 1, src/scene_convert.h:14);
    `scene_opacity` multiplies every material's opacity as the app's render config does.
    With `fit_view` the rescaled scene is moved in front of the camera (at the origin,
    looking down -z) so that the front face of its bounding box fills the view; without it
    the positions are used as they are after the rescale. `lighting` goes to scenes.make_config."""
    from .quads import generate_quads

    pos = mesh.positions.astype(np.float32).copy()
    nrm = None if mesh.normals is None else mesh.normals.astype(np.float32).copy()
    if flip_yz:
        pos[:, [1, 2]] = pos[:, [2, 1]]
        if nrm is not None:
            nrm[:, [1, 2]] = nrm[:, [2, 1]]
            nrm = -nrm

    draws: List[MeshDraw] = []
    quad_parts = []
    offset = 0
    for mat, first, num in mesh.groups:
        if num == 0:
            continue
        tris = mesh.tris[first:first + num]
        q = generate_quads(pos, tris.astype(np.uint32), squareness)
        quad_parts.append(q)
        draws.append(MeshDraw(mat, offset, len(q), num))
        offset += len(q)
    quads = np.concatenate(quad_parts) if quad_parts else np.zeros((0, 4), np.uint32)

    if len(pos):
        lo, hi = pos.min(axis=0), pos.max(axis=0)
        extent = float((hi - lo).max())
        if extent > 0:
            pos = (pos * np.float32(100.0 / extent)).astype(np.float32)
    cam = scenes.Camera(width, height)
    if fit_view and len(pos):
        # Centre the box on the view axis with its front face at the distance where that
        # face's x and y extents fill 95% of the view; everything behind it projects smaller.
        lo, hi = pos.min(axis=0).astype(np.float64), pos.max(axis=0).astype(np.float64)
        centre, half_ext = (lo + hi) * 0.5, (hi - lo) * 0.5
        t = math.tan(math.radians(cam.fov_y_deg) * 0.5)
        front = max(half_ext[0] / (t * cam.aspect), half_ext[1] / t) / 0.95
        front = max(front, 2.0 * cam.near)
        pos = (pos - centre + np.array([0.0, 0.0, -front - half_ext[2]])).astype(np.float32)

    insts, colors = [], []
    has_normals = nrm is not None
    for d in draws:
        m = mesh.materials[d.material_id]
        opacity = m.dissolve * scene_opacity
        flags = scenes.INST_HAS_VERTEX_NORMALS if has_normals else 0
        if m.dissolve == 1.0 and scene_opacity >= 1.0:
            flags |= scenes.INST_IS_OPAQUE
        color = material_color(m.diffuse, opacity)
        if color != 0xFFFFFFFF:
            flags |= scenes.INST_HAS_COLOR
        part = scenes.split_instances(d.num_quads, flags, quad_offset=d.quad_offset)
        insts.append(part)
        colors.append(np.full(len(part), color, np.uint32))
    inst = np.concatenate(insts) if insts else np.zeros(0, scenes.INSTANCE_DTYPE)
    inst_colors = np.concatenate(colors) if colors else np.zeros(0, np.uint32)

    sc = scenes.Scene(width, height, pos, quads.astype(np.uint32), inst,
                      scenes.make_config(cam, len(inst), **lighting), name=name)
    if has_normals:
        sc.normals = encode_normals(nrm)
    if inst_colors.size and (inst["flags"] & scenes.INST_HAS_COLOR).any():
        sc.instance_colors = inst_colors
    return sc, draws


def load_obj_scene(path: str, width: int, height: int, merge_verts: bool = False,
                   **kwargs) -> Tuple[scenes.Scene, List[MeshDraw]]:
    """load_obj (then merge_vertices when asked and the file has no texture coordinates,
    as convertScene does) then obj_scene; the scene is named after the file."""
    kwargs.setdefault("name", os.path.splitext(os.path.basename(path))[0])
    mesh = load_obj(path)
    if merge_verts and mesh.tex_coords is None:
        mesh = merge_vertices(mesh)
    return obj_scene(mesh, width, height, **kwargs)


@dataclasses.dataclass
class InputScene:
    """One <scene> of an input-scenes XML file (InputScene, src/scene_convert.h:8-19,
    src/scene_convert.cpp:34-63)."""

    name: str
    path: str
    quad_squareness: float = 1.0
    merge_verts: bool = False
    flip_uv: bool = False
    flip_yz: bool = False
    pbr: bool = False
    env_map: str = ""

    def load(self, width: int, height: int, **kwargs) -> Tuple[scenes.Scene, List[MeshDraw]]:
        """The scene with this entry's conversion parameters (flip_uv, pbr and env_map have
        no effect here: textures are not sampled)."""
        kwargs.setdefault("name", self.name)
        return load_obj_scene(self.path, width, height, merge_verts=self.merge_verts,
                              squareness=self.quad_squareness, flip_yz=self.flip_yz, **kwargs)


def load_input_scenes(xml_path: str, root_path: Optional[str] = None) -> List[InputScene]:
    """Parses an input-scenes XML file (the reference's input_scenes.xml format): a <scenes>
    root with an optional root_path, one <scene path=... /> per scene with optional name,
    quad_squareness, merge_verts, flip_uv, flip_yz, pbr and env_map. `root_path` overrides
    the file's (its scenes may live elsewhere on this machine)."""
    import xml.etree.ElementTree as ET

    root = ET.parse(xml_path).getroot()
    if root.tag != "scenes":
        raise ObjError(f"{xml_path}: root element is <{root.tag}>, not <scenes>")
    base = root_path if root_path is not None else root.get("root_path", "")

    def flag(node, key):
        v = node.get(key, "false").strip().lower()
        if v not in ("true", "false", "1", "0"):
            raise ObjError(f"{xml_path}: {key}={v!r} is not a bool")
        return v in ("true", "1")

    out = []
    for node in root.findall("scene"):
        path = node.get("path")
        if not path:
            raise ObjError(f"{xml_path}: <scene> without a path")
        name = node.get("name") or os.path.splitext(os.path.basename(path))[0]
        out.append(InputScene(name, os.path.join(base, path),
                              float(node.get("quad_squareness", "1.0")), flag(node, "merge_verts"),
                              flag(node, "flip_uv"), flag(node, "flip_yz"), flag(node, "pbr"),
                              node.get("env_map", "")))
    return out


def main(argv=None) -> int:
    """python -m lucid_amd.scene_io scene.obj [--size 1920x1080] [--frames 20]: renders an
    OBJ scene on the GPU and prints one JSON line (quads, fragments, ms per frame, stage
    times); --ppm writes the frame. With --input-scenes list.xml the first argument names
    a scene of that list, converted with its parameters."""
    import argparse
    import json
    import sys

    p = argparse.ArgumentParser(prog="python -m lucid_amd.scene_io")
    p.add_argument("obj", help="an OBJ file, or a scene name with --input-scenes")
    p.add_argument("--input-scenes", default="", help="input-scenes XML file")
    p.add_argument("--root", default=None, help="overrides the XML file's root_path")
    p.add_argument("--size", default="1920x1080")
    p.add_argument("--frames", type=int, default=20)
    p.add_argument("--opacity", type=float, default=0.5, help="scene opacity factor")
    p.add_argument("--squareness", type=float, default=1.0)
    p.add_argument("--flip-yz", action="store_true")
    p.add_argument("--autotune", action="store_true")
    p.add_argument("--ppm", default="")
    a = p.parse_args(argv)
    w, h = (int(x) for x in a.size.lower().split("x"))

    import torch

    from . import _lib
    from .renderer import GpuScene, LucidRenderer

    if a.input_scenes:
        entries = {e.name: e for e in load_input_scenes(a.input_scenes, a.root)}
        if a.obj not in entries:
            print(f"no scene {a.obj!r} in {a.input_scenes} (has {sorted(entries)})",
                  file=sys.stderr)
            return 2
        sc, draws = entries[a.obj].load(w, h, scene_opacity=a.opacity)
    else:
        sc, draws = load_obj_scene(a.obj, w, h, scene_opacity=a.opacity,
                                   squareness=a.squareness, flip_yz=a.flip_yz)
    print(f"[scene_io] {sc.name}: {len(draws)} meshes, {sc.num_quads} quads, "
          f"{sc.num_verts} vertices", file=sys.stderr, flush=True)
    r = LucidRenderer((w, h), opts=_lib.OPT_TIMERS, max_quads=sc.num_quads)
    g = GpuScene(sc)
    out = r.alloc_outputs(rgba8=True, counts=False)
    if a.autotune:
        r.autotune(g, out, iters=3)
    r.render(g, out)
    frags = int(r.read_info()[60])
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(a.frames):
        r.render_async(g, out)
    e1.record(stream)
    torch.cuda.synchronize()
    _lib.check(r.finish(), "frames")
    ms = e0.elapsed_time(e1) / a.frames
    r.render_async(g, out)
    stages = r.stage_times()
    _lib.check(r.finish(), "stage timing")
    print(json.dumps({"scene": sc.name, "size": [w, h], "quads": sc.num_quads,
                      "fragments": frags, "ms_per_frame": ms,
                      "fragments_per_s": frags / (ms * 1e-3), "stages_ms": stages}), flush=True)
    if a.ppm:
        img = out["image"].cpu().numpy()[..., :3]
        with open(a.ppm, "wb") as f:
            f.write(b"P6\n%d %d\n255\n" % (w, h))
            f.write(img.astype(np.uint8).tobytes())
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
