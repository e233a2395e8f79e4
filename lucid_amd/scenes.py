"""Seeded synthetic scenes and LucidConfig bytes for the exact-OIT rasterizer.

The reference builds ``LucidConfig`` from libfwk's Camera/Frustum
(src/shading.cpp:46-74, src/lucid_renderer.cpp:431-451); libfwk is not
available, so the camera model is restated here and the same config bytes feed
both the HIP path and the oracle:

* a camera is a pose (world position, and a rotation whose columns are the
  view x, y and z axes in world space; it looks down its -z axis, +y is up)
  and a symmetric perspective projection; the view matrix is the inverse pose
  and ``view_proj_matrix`` = projection x view (src/lucid_renderer.cpp:440-441);
* the frustum's corner rays start at the camera position (``ws_origins`` and
  ``ws_origin0``) and point along the rotated view-space corner directions,
  unnormalised with view z = -1, so a plane's ``inv_ray_pos`` equals 1 / view
  depth; ``ws_dirx``/``ws_diry`` step one pixel (src/shading.cpp:46-61, with
  ``inverseOrZero(viewMatrix())`` applied as mulPoint/mulNormal);
* the default camera sits at the origin looking down -Z (world == view space),
  90 degree vertical FOV, depth range [1/16, 1024] (src/lucid_app.cpp:95),
  GL-style clip z in [-w, w] and a flipped y row so that image row 0 is NDC
  y = -1 (top of the image);
* identity lighting (ambient 1, sun power 0) so finalShading reduces to an
  sRGB round trip (shared/funcs.glsl:261-271).

Soup triangles become degenerate quads ``(a, b, c, c)`` exactly as the
reference's quad generator emits unpaired triangles
(src/quad_generator.cpp:192); their second triangle is culled by index
(quad_setup.glsl:144). Draw calls are split into instances of at most 1024
quads (src/lucid_renderer.cpp:352-429).
"""

from __future__ import annotations

import dataclasses
import hashlib
import math
from typing import Optional

import numpy as np

# definitions.glsl:84-94
INST_HAS_VERTEX_COLORS = 0x001
INST_HAS_VERTEX_TEX_COORDS = 0x002
INST_HAS_VERTEX_NORMALS = 0x004
INST_IS_OPAQUE = 0x008
INST_TEX_OPAQUE = 0x010
INST_HAS_UV_RECT = 0x020
INST_HAS_ALBEDO_TEXTURE = 0x040
INST_HAS_COLOR = 0x200

MAX_INSTANCE_QUADS = 1024  # src/lucid_renderer.h:29

INSTANCE_DTYPE = np.dtype(
    [("index_offset", "<i4"), ("vertex_offset", "<i4"), ("num_quads", "<i4"), ("flags", "<u4")]
)

CONFIG_BYTES = 352
NEAR, FAR = 1.0 / 16.0, 1024.0


@dataclasses.dataclass
class Camera:
    width: int
    height: int
    fov_y_deg: float = 90.0
    near: float = NEAR
    far: float = FAR
    # pose: world position, and world-from-view rotation (columns: view x, y, z axes)
    position: tuple = (0.0, 0.0, 0.0)
    rotation: Optional[np.ndarray] = None

    @property
    def aspect(self) -> float:
        return self.width / self.height

    @property
    def focal(self) -> float:
        return 1.0 / math.tan(math.radians(self.fov_y_deg) * 0.5)

    @property
    def rot(self) -> np.ndarray:
        """World-from-view rotation (float64 3x3)."""
        return np.eye(3) if self.rotation is None else np.asarray(self.rotation, np.float64)

    @property
    def eye(self) -> np.ndarray:
        return np.asarray(self.position, np.float64)

    def projection(self) -> np.ndarray:
        """Row-major float64 4x4 perspective projection with flipped y (view -> clip)."""
        f, a, n, fa = self.focal, self.aspect, self.near, self.far
        m = np.zeros((4, 4))
        m[0, 0] = f / a
        m[1, 1] = -f
        m[2, 2] = (fa + n) / (n - fa)
        m[2, 3] = 2.0 * fa * n / (n - fa)
        m[3, 2] = -1.0
        return m

    def view_matrix(self) -> np.ndarray:
        """Row-major float64 4x4 world -> view transform (the inverse pose)."""
        m = np.eye(4)
        rt = self.rot.T
        m[:3, :3] = rt
        m[:3, 3] = -rt @ self.eye
        return m

    def view_proj(self) -> np.ndarray:
        """Column-major 4x4 float32 (cols[c][r]) of projection x view, as LucidConfig holds
        view_proj_matrix."""
        return (self.projection() @ self.view_matrix()).T.astype(np.float32)

    def world_to_view(self, pts: np.ndarray) -> np.ndarray:
        return (np.asarray(pts, np.float64) - self.eye) @ self.rot

    def view_to_world(self, pts: np.ndarray) -> np.ndarray:
        return np.asarray(pts, np.float64) @ self.rot.T + self.eye

    def ndc_to_view(self, ndc_xy: np.ndarray, depth: np.ndarray) -> np.ndarray:
        """View-space points at view depth `depth` (positive) that project to NDC `ndc_xy`."""
        f, a = self.focal, self.aspect
        d = np.asarray(depth, np.float64)[..., None]
        x = ndc_xy[..., 0:1] * (a / f) * d
        y = -ndc_xy[..., 1:2] / f * d
        return np.concatenate([x, y, -d], axis=-1)

    def view_corner_dirs(self) -> np.ndarray:
        """View-space ray directions at screen corners (x, y) = (0,0), (0,H), (W,H), (W,0)."""
        f, a = np.float32(self.focal), np.float32(self.aspect)
        af = np.float32(a / f)
        inv_f = np.float32(1.0) / f
        return np.array(
            [[-af, inv_f, -1.0], [-af, -inv_f, -1.0], [af, -inv_f, -1.0], [af, inv_f, -1.0]],
            np.float32,
        )

    def corner_dirs(self) -> np.ndarray:
        """World-space corner ray directions (float32): the view-space ones rotated by the
        pose (mulNormal of the inverse view matrix, src/shading.cpp:53-56)."""
        d = self.view_corner_dirs()
        if self.rotation is None:
            return d
        return (d.astype(np.float64) @ self.rot.T).astype(np.float32)

    def pixel_rays(self) -> np.ndarray:
        """float64 view-space directions (view z = -1) through every pixel centre, (H, W, 3)."""
        h, w = self.height, self.width
        ys, xs = np.mgrid[0:h, 0:w]
        ndc_x = (xs + 0.5) * (2.0 / w) - 1.0
        ndc_y = (ys + 0.5) * (2.0 / h) - 1.0
        return self.ndc_to_view(np.stack([ndc_x, ndc_y], -1), np.ones((h, w)))


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """World-from-view rotation of a camera at `eye` looking at `target`."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    return np.stack([right, np.cross(back, right), back], axis=1)


def make_config(
    camera: Camera,
    num_instances: int,
    *,
    ambient=(1.0, 1.0, 1.0),
    ambient_power: float = 1.0,
    sun_color=(1.0, 1.0, 1.0),
    sun_dir=(0.0, 0.0, -1.0),
    sun_power: float = 0.0,
    background=(0.0, 0.0, 0.0, 1.0),
    backface_culling: bool = False,
    packet_size: int = 1,
) -> bytes:
    """LucidConfig in std140 layout (shared/structures.glsl:101-111)."""
    buf = np.zeros(CONFIG_BYTES // 4, np.float32)
    u32 = buf.view(np.uint32)
    i32 = buf.view(np.int32)
    dirs = camera.corner_dirs()
    eye = camera.eye.astype(np.float32)
    # Frustum: ws_origins[4] at 0, ws_dirs[4] at 16, origin0 at 32, dir0 at 36, dirx 40, diry 44
    for i in range(4):
        buf[4 * i : 4 * i + 4] = (*eye, 0.0)
        buf[16 + 4 * i : 16 + 4 * i + 3] = dirs[i]
    buf[32:36] = (*eye, 1.0)
    buf[36:39] = dirs[0]
    dirx = (dirs[3] - dirs[0]) * (np.float32(1.0) / np.float32(camera.width))
    diry = (dirs[1] - dirs[0]) * (np.float32(1.0) / np.float32(camera.height))
    buf[40:43] = dirx.astype(np.float32)
    buf[44:47] = diry.astype(np.float32)
    buf[48:64] = camera.view_proj().reshape(-1)
    # Lighting at word 64: ambient_color, sun_color, sun_dir, sun_power, ambient_power
    buf[64:68] = (*ambient, 1.0)
    buf[68:72] = (*sun_color, 1.0)
    buf[72:76] = (*sun_dir, 0.0)
    buf[76] = sun_power
    buf[77] = ambient_power
    buf[80:84] = background
    u32[84] = 1 if backface_culling else 0
    i32[85] = num_instances
    i32[86] = packet_size
    return buf.tobytes()


def unpack_rgba8(img: np.ndarray) -> np.ndarray:
    """(..., ) uint32 RGBA8 -> (..., 4) uint8 channels."""
    img = np.asarray(img, np.uint32)
    return np.stack([(img >> (8 * k)) & 0xFF for k in range(4)], axis=-1).astype(np.uint8)


def pack_channels(ch: np.ndarray) -> np.ndarray:
    """(..., 4) uint8 channels -> (...) uint32 RGBA8."""
    ch = np.asarray(ch, np.uint32)
    return ch[..., 0] | (ch[..., 1] << 8) | (ch[..., 2] << 16) | (ch[..., 3] << 24)


def box_mip(level: np.ndarray) -> np.ndarray:
    """Next mip level of an RGBA8 image: the rounded mean of each 2 x 2 texel block (a size
    of 1 stays 1; an odd size drops its last row or column, as max(1, size >> 1) does)."""
    h, w = level.shape
    nh, nw = max(1, h >> 1), max(1, w >> 1)
    ch = unpack_rgba8(level).astype(np.float64)
    ys = np.minimum(np.arange(2 * nh).reshape(nh, 2), h - 1)
    xs = np.minimum(np.arange(2 * nw).reshape(nw, 2), w - 1)
    acc = np.zeros((nh, nw, 4))
    for dy in range(2):
        for dx in range(2):
            acc += ch[ys[:, dy]][:, xs[:, dx]]
    return pack_channels(np.floor(acc / 4.0 + 0.5).astype(np.uint8))


@dataclasses.dataclass
class Texture:
    """An RGBA8 atlas with its mip chain, as lucid_texture takes it: levels[l] is a
    (max(1, H >> l), max(1, W >> l)) uint32 image, row 0 = v 0."""

    levels: list

    @property
    def width(self) -> int:
        return int(self.levels[0].shape[1])

    @property
    def height(self) -> int:
        return int(self.levels[0].shape[0])

    @property
    def num_levels(self) -> int:
        return len(self.levels)

    def packed(self) -> np.ndarray:
        return np.concatenate([np.ascontiguousarray(lv, np.uint32).reshape(-1)
                               for lv in self.levels])

    @staticmethod
    def with_mips(level0: np.ndarray, num_levels: Optional[int] = None) -> "Texture":
        """Full (or num_levels-deep) chain built with box_mip."""
        lv = [np.ascontiguousarray(level0, np.uint32)]
        top = max(level0.shape)
        full = top.bit_length()
        for _ in range((num_levels or full) - 1):
            lv.append(box_mip(lv[-1]))
        return Texture(lv)


@dataclasses.dataclass
class Scene:
    """Host-side scene arrays in the layouts the C ABI takes."""

    width: int
    height: int
    positions: np.ndarray  # (V, 3) float32
    quad_indices: np.ndarray  # (Q, 4) uint32
    instances: np.ndarray  # INSTANCE_DTYPE
    config: bytes
    colors_rgba8: Optional[np.ndarray] = None  # (V,) uint32
    colors_rgb: Optional[np.ndarray] = None  # (V, 3) float32
    tri_alpha: Optional[np.ndarray] = None  # (2Q,) float32, by triangle id
    normals: Optional[np.ndarray] = None  # (V,) uint32 packed 10:10:10
    instance_colors: Optional[np.ndarray] = None  # (N,) uint32
    name: str = ""
    camera: Optional[Camera] = None  # the camera `config` was made from (None: the default)
    # albedo textures: per-vertex coordinates, per-instance uv rectangles (min.xy, size.xy)
    # and the (opaque, transparent) atlases
    tex_coords: Optional[np.ndarray] = None  # (V, 2) float32
    instance_uv_rects: Optional[np.ndarray] = None  # (N, 4) float32
    textures: Optional[tuple] = None  # (Texture | None, Texture | None)

    @property
    def num_quads(self) -> int:
        return int(self.quad_indices.shape[0])

    @property
    def num_verts(self) -> int:
        return int(self.positions.shape[0])

    def input_bytes(self) -> int:
        b = self.quad_indices.nbytes + self.positions.nbytes + self.instances.nbytes
        for arr in (self.colors_rgba8, self.colors_rgb, self.tri_alpha, self.normals,
                    self.instance_colors, self.tex_coords, self.instance_uv_rects):
            if arr is not None:
                b += arr.nbytes
        return int(b)

    def digest(self) -> str:
        h = hashlib.sha256()
        for arr in (self.positions, self.quad_indices, self.instances, self.colors_rgba8,
                    self.colors_rgb, self.tri_alpha, self.normals, self.instance_colors,
                    self.tex_coords, self.instance_uv_rects):
            if arr is not None:
                h.update(np.ascontiguousarray(arr).tobytes())
        for t in (self.textures or ()):
            if t is not None:
                h.update(t.packed().tobytes())
        h.update(self.config)
        return h.hexdigest()


def scene_camera(scene: "Scene") -> Camera:
    return scene.camera if scene.camera is not None else Camera(scene.width, scene.height)


def posed(scene: Scene, position, rotation, *, backface_culling: bool = False,
          name: Optional[str] = None) -> Scene:
    """The scene moved rigidly so that a camera at `position` with world-from-view
    `rotation` sees what the default camera saw (positions were view-space points of the
    default camera), with that camera's LucidConfig. Only float rounding tells the two
    frames apart."""
    cam = Camera(scene.width, scene.height, position=tuple(float(v) for v in position),
                 rotation=np.asarray(rotation, np.float64))
    pos = cam.view_to_world(scene.positions.astype(np.float64)).astype(np.float32)
    cfg = make_config(cam, len(scene.instances), backface_culling=backface_culling)
    return dataclasses.replace(scene, positions=pos, config=cfg, camera=cam,
                               name=name or f"posed_{scene.name}")


def split_instances(num_quads: int, flags: int, quad_offset: int = 0,
                    vertex_offset: int = 0) -> np.ndarray:
    n = (num_quads + MAX_INSTANCE_QUADS - 1) // MAX_INSTANCE_QUADS
    inst = np.zeros(n, INSTANCE_DTYPE)
    starts = np.arange(n, dtype=np.int64) * MAX_INSTANCE_QUADS
    inst["index_offset"] = (starts + quad_offset) * 4
    inst["vertex_offset"] = vertex_offset
    inst["num_quads"] = np.minimum(MAX_INSTANCE_QUADS, num_quads - starts)
    inst["flags"] = flags
    return inst


def pack_rgba8(rgb: np.ndarray, alpha: np.ndarray) -> np.ndarray:
    c = np.concatenate([rgb, alpha[..., None]], axis=-1)
    q = np.clip(np.floor(c.astype(np.float64) * 255.0 + 0.5), 0, 255).astype(np.uint32)
    return q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (q[..., 3] << 24)


def soup_scene(width: int, height: int, ndc_verts: np.ndarray, depth: np.ndarray,
               rgb: np.ndarray, alpha: np.ndarray, *, color_format: str = "f32",
               camera: Optional[Camera] = None, name: str = "", **config_kw) -> Scene:
    """Triangle soup; each triangle owns its 3 vertices and becomes quad (a,b,c,c).

    ndc_verts: (T,3,2) NDC positions; depth: (T,) or (T,3) view depths;
    rgb: (T,3,3) per-vertex colour in [0,1]; alpha: (T,) per-triangle alpha.
    """
    cam = camera or Camera(width, height)
    t = ndc_verts.shape[0]
    depth = np.asarray(depth, np.float64)
    if depth.ndim == 1:
        depth = np.repeat(depth[:, None], 3, axis=1)
    pos = cam.ndc_to_view(ndc_verts.astype(np.float64), depth).astype(np.float32)
    positions = pos.reshape(-1, 3)
    base = np.arange(t, dtype=np.uint32) * 3
    quads = np.stack([base, base + 1, base + 2, base + 2], axis=1).astype(np.uint32)
    flags = INST_HAS_VERTEX_COLORS
    inst = split_instances(t, flags)
    sc = Scene(width, height, positions, quads, inst,
               make_config(cam, len(inst), **config_kw), name=name)
    rgb = rgb.astype(np.float32).reshape(-1, 3)
    if color_format == "rgba8":
        sc.colors_rgba8 = pack_rgba8(rgb, np.repeat(alpha.astype(np.float32), 3))
    else:
        sc.colors_rgb = rgb
        ta = np.zeros(2 * t, np.float32)
        ta[0::2] = alpha
        sc.tri_alpha = ta
    return sc


# ----------------------------------------------------------------------------- configs


def config0(seed: int = 1000, num_tris: int = 10_000, width: int = 512, height: int = 512,
            color_format: str = "f32", vertex_sigma: float = 0.0615) -> Scene:
    """10k soup triangles: centroids uniform in NDC, Gaussian vertex offsets in NDC (sigma
    calibrated to a mean depth complexity of ~8 fragments/pixel at 10k tris / 512x512),
    view depth U[1,100] per triangle, RGB U[0,1] per vertex, alpha U[0.2,0.8]."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (num_tris, 1, 2))
    e = rng.normal(0.0, vertex_sigma, (num_tris, 3, 2))
    depth = rng.uniform(1.0, 100.0, num_tris)
    rgb = rng.uniform(0.0, 1.0, (num_tris, 3, 3))
    alpha = rng.uniform(0.2, 0.8, num_tris)
    return soup_scene(width, height, c + e, depth, rgb, alpha, color_format=color_format,
                      name=f"c0_{num_tris}_{width}x{height}")


def icosphere(level: int = 3):
    """Unit icosphere: the icosahedron with each triangle split into 4, `level` times
    (20 * 4^level triangles; level 3 = 1280). Returns (positions (V, 3), tris (T, 3))."""
    p = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p),
             (0, 1, -p), (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
            (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
            (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]
        nxt = []
        for a, b, c in tris:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = nxt
    return np.array(verts, np.float32), np.array(tris, np.uint32)


def config2(seed: int = 1002, grid=(16, 16, 16), width: int = 2560, height: int = 1440,
            level: int = 3, square_weight: float = 0.5) -> Scene:
    """Mesh workload: grid[0]*grid[1]*grid[2] icospheres (20 * 4^level triangles each, 1280
    at level 3; 4096 spheres by default) on a jittered 3D grid in view, alpha 0.5 for every
    triangle, fp32 per-vertex colours. Triangles are paired into quads by the quad generator
    (lucid_generate_quads), run once on the sphere mesh and replicated (pairing depends only
    on the mesh's shape, which every sphere shares up to translation and scale)."""
    from .quads import generate_quads
    rng = np.random.default_rng(seed)
    sp, st = icosphere(level)
    sq = generate_quads(sp, st, square_weight)
    gx, gy, gz = grid
    n = gx * gy * gz
    cam = Camera(width, height)
    k = np.arange(n)
    i, j, l = k % gx, (k // gx) % gy, k // (gx * gy)
    depth = 4.0 + (l + rng.uniform(-0.3, 0.3, n)) * (40.0 / max(gz, 1))
    ndc = np.stack([(i + 0.5 + rng.uniform(-0.3, 0.3, n)) / gx * 1.8 - 0.9,
                    (j + 0.5 + rng.uniform(-0.3, 0.3, n)) / gy * 1.8 - 0.9], axis=1)
    centres = cam.ndc_to_view(ndc[:, None, :], depth[:, None])[:, 0, :]
    radius = depth * (1.8 / max(gx, gy)) * 0.45 * np.tan(np.radians(45.0))
    pos = (centres[:, None, :] + sp[None, :, :] * radius[:, None, None]).astype(np.float32)
    nv = len(sp)
    quads = (sq[None, :, :].astype(np.int64) + (np.arange(n) * nv)[:, None, None]).astype(np.uint32)
    quads = quads.reshape(-1, 4)
    inst = split_instances(len(quads), INST_HAS_VERTEX_COLORS)
    sc = Scene(width, height, pos.reshape(-1, 3), quads, inst, make_config(cam, len(inst)),
               name=f"c2_{n}x{len(st)}_{width}x{height}")
    sc.colors_rgb = rng.uniform(0.0, 1.0, (n * nv, 3)).astype(np.float32)
    sc.tri_alpha = np.full(2 * len(quads), 0.5, np.float32)
    return sc


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (768, 432), (1024, 1024))


def config3(seed: int = 1003, num_scenes: int = 64, min_tris: int = 5_000,
            max_tris: int = 500_000):
    """The batch workload: independent config0-generator scenes with log-normal triangle
    counts clipped to [min_tris, max_tris] (median ~50k) at resolutions drawn from
    {256x256, 512x512, 768x432, 1024x1024}."""
    rng = np.random.default_rng(seed)
    counts = np.clip(np.exp(rng.normal(np.log(50_000), 1.0, num_scenes)), min_tris, max_tris)
    res = rng.integers(0, len(BATCH_RESOLUTIONS), num_scenes)
    out = []
    for i in range(num_scenes):
        w, h = BATCH_RESOLUTIONS[res[i]]
        sc = config0(seed=seed * 1000 + i, num_tris=int(counts[i]), width=w, height=h)
        sc.name = f"c3_{i}_{int(counts[i])}_{w}x{h}"
        out.append(sc)
    return out


def config4(seed: int = 1004, num_tris: int = 200_000, width: int = 512,
            height: int = 512) -> Scene:
    """The differentiable workload: the config0 generator at 200k triangles (mean depth
    complexity ~160 fragments/pixel at 512x512), fp32 colours."""
    sc = config0(seed=seed, num_tris=num_tris, width=width, height=height)
    sc.name = f"c4_{num_tris}_{width}x{height}"
    return sc


def _lognormal_edges(rng, n, median_px, sigma, width):
    lengths = median_px * np.exp(rng.normal(0.0, sigma, (n, 3)))
    ang = rng.uniform(0.0, 2.0 * np.pi, (n, 3))
    off = np.stack([np.cos(ang), np.sin(ang)], axis=-1) * lengths[..., None]
    return off * (2.0 / width)  # pixels -> NDC (x and y use the same pixel scale below)


def config1(seed: int = 1001, num_tris: int = 1_000_000, width: int = 1920, height: int = 1080,
            num_blobs: int = 64, blob_sigma: float = 0.1, median_px: float = 6.0,
            edge_sigma: float = 1.03, blob_weight_sigma: float = 2.3,
            color_format: str = "f32") -> Scene:
    """1M soup triangles clustered in 64 Gaussian blobs (sigma 0.1 NDC) with log-normal
    populations, log-normal vertex offsets (median 6 px), view depth U[1,1000]. The two
    log-normal sigmas are calibrated to a mean depth complexity of ~32 and a maximum of
    ~500 fragments/pixel."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-0.8, 0.8, (num_blobs, 2))
    weights = np.exp(rng.normal(0.0, blob_weight_sigma, num_blobs))
    which = rng.choice(num_blobs, num_tris, p=weights / weights.sum())
    c = centers[which] + rng.normal(0.0, blob_sigma, (num_tris, 2))
    off = _lognormal_edges(rng, num_tris, median_px, edge_sigma, width)
    off[..., 1] *= width / height  # same pixel length in y
    ndc = c[:, None, :] + off
    depth = rng.uniform(1.0, 1000.0, num_tris)
    rgb = rng.uniform(0.0, 1.0, (num_tris, 3, 3))
    alpha = rng.uniform(0.2, 0.8, num_tris)
    return soup_scene(width, height, ndc, depth, rgb, alpha, color_format=color_format,
                      name=f"c1_{num_tris}_{width}x{height}")


def planes_scene(width: int, height: int, layers, color_format: str = "f32") -> Scene:
    """Screen-covering fronto-parallel quads (the reference's #planes setup,
    src/scene_setup.cpp:194-222, made screen-aligned). layers: list of
    (depth, (r, g, b), alpha). Each layer is a proper 2-triangle quad."""
    cam = Camera(width, height)
    n = len(layers)
    positions = []
    quads = []
    rgb = []
    tri_alpha = np.zeros(2 * n, np.float32)
    # A small rotation keeps every edge off the exact screen axes: an edge with a zero x
    # slope divides by zero in the reference's scanline setup (quad_setup.glsl:325-327) and
    # loses coverage; the reference's #planes setup tilts its camera for the same reason
    # (src/scene_setup.cpp:216-218).
    ang = 0.013
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    for i, (d, col, a) in enumerate(layers):
        ndc = np.array([[-1.2, -1.2], [1.2, -1.2], [1.2, 1.2], [-1.2, 1.2]]) @ rot.T
        p = cam.ndc_to_view(ndc, np.full(4, d)).astype(np.float32)
        positions.append(p)
        quads.append([4 * i, 4 * i + 1, 4 * i + 2, 4 * i + 3])
        rgb.append(np.tile(np.asarray(col, np.float32), (4, 1)))
        tri_alpha[2 * i] = tri_alpha[2 * i + 1] = a
    positions = np.concatenate(positions).astype(np.float32)
    inst = split_instances(n, INST_HAS_VERTEX_COLORS)
    sc = Scene(width, height, positions, np.asarray(quads, np.uint32), inst,
               make_config(cam, len(inst)), name=f"planes_{n}")
    rgb = np.concatenate(rgb).astype(np.float32)
    if color_format == "rgba8":
        sc.colors_rgba8 = pack_rgba8(rgb, np.repeat(tri_alpha[0::2], 4))
    else:
        sc.colors_rgb = rgb
        sc.tri_alpha = tri_alpha
    return sc


def quad_soup_scene(width: int, height: int, ndc: np.ndarray, depth: np.ndarray, rgb: np.ndarray,
                    alpha: np.ndarray, *, color_format: str = "f32", name: str = "",
                    **config_kw) -> Scene:
    """Independent proper quads (2 triangles each): ndc (N,4,2), depth (N,4) view depths,
    rgb (N,4,3), alpha (N,) per quad (both triangles)."""
    cam = Camera(width, height)
    n = ndc.shape[0]
    pos = cam.ndc_to_view(ndc.astype(np.float64), depth.astype(np.float64)).astype(np.float32)
    positions = pos.reshape(-1, 3)
    base = np.arange(n, dtype=np.uint32) * 4
    quads = np.stack([base, base + 1, base + 2, base + 3], axis=1).astype(np.uint32)
    inst = split_instances(n, INST_HAS_VERTEX_COLORS)
    sc = Scene(width, height, positions, quads, inst, make_config(cam, len(inst), **config_kw),
               name=name)
    rgbf = rgb.astype(np.float32).reshape(-1, 3)
    if color_format == "rgba8":
        sc.colors_rgba8 = pack_rgba8(rgbf, np.repeat(alpha.astype(np.float32), 4))
    else:
        sc.colors_rgb = rgbf
        sc.tri_alpha = np.repeat(alpha.astype(np.float32), 2)
    return sc


def tilted_planes(width: int, height: int, n: int, seed: int = 7, extent: float = 1.1,
                  color_format: str = "f32") -> Scene:
    """n large quads with random depth slopes: per-pixel order changes across the screen,
    so block-level pre-sorting is wrong in many places (exercises the exact fallback)."""
    rng = np.random.default_rng(seed)
    ang = 0.011  # keeps edges off the exact screen axes (see planes_scene)
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    corners = (np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64) * extent) @ rot.T
    ndc = np.repeat(corners[None], n, axis=0)
    z0 = rng.uniform(5.0, 15.0, (n, 1))
    slope = rng.uniform(-4.0, 4.0, (n, 2))
    depth = z0 + (corners[None, :, 0] * slope[:, 0:1] + corners[None, :, 1] * slope[:, 1:2])
    rgb = rng.uniform(0.0, 1.0, (n, 4, 3))
    alpha = rng.uniform(0.05, 0.3, n)
    return quad_soup_scene(width, height, ndc, depth, rgb, alpha, color_format=color_format,
                           name=f"tilted_{n}")


def block_stack(width: int, height: int, n: int, center_px=(12, 12), size_px: float = 6.0,
                seed: int = 3, color_format: str = "f32") -> Scene:
    """n small soup triangles stacked over one 8x8 block at random depths."""
    rng = np.random.default_rng(seed)
    cx = (np.float64(center_px[0]) + 0.5) / width * 2 - 1
    cy = (np.float64(center_px[1]) + 0.5) / height * 2 - 1
    s = size_px / width * 2
    base = np.array([[-s, -s], [s, -0.5 * s], [0.0, s]])
    jitter = rng.uniform(-0.3, 0.3, (n, 3, 2)) * s
    ndc = np.array([cx, cy])[None, None, :] + base[None] + jitter
    depth = rng.uniform(2.0, 50.0, n)
    rgb = rng.uniform(0.0, 1.0, (n, 3, 3))
    alpha = rng.uniform(0.02, 0.2, n)
    return soup_scene(width, height, ndc, depth, rgb, alpha, color_format=color_format,
                      name=f"stack_{n}")


# ----------------------------------------------------------------------------- textures


def checker_texture(width: int, height: int, seed: int = 0, alpha_range=(255, 255),
                    num_levels: Optional[int] = None) -> Texture:
    """A random-colour checkerboard of 4 x 4-texel cells with per-texel noise, and its box-
    filtered mip chain; alpha uniform in alpha_range."""
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, 256, ((height + 3) // 4, (width + 3) // 4, 3))
    ys, xs = np.mgrid[0:height, 0:width]
    rgb = cells[ys // 4, xs // 4] + rng.integers(-24, 25, (height, width, 3))
    a = rng.integers(alpha_range[0], alpha_range[1] + 1, (height, width, 1))
    ch = np.clip(np.concatenate([rgb, a], axis=-1), 0, 255).astype(np.uint8)
    return Texture.with_mips(pack_channels(ch), num_levels)


def textured_screen_plane(width: int, height: int, texture: Texture, *, uv_scale=(1.0, 1.0),
                          uv_offset=(0.0, 0.0), opaque: bool = True, uv_rect=None,
                          alpha: float = 1.0, depth: float = 4.0,
                          color_format: str = "f32") -> Scene:
    """One screen-covering fronto-parallel quad (slightly rotated, as planes_scene) whose
    texture coordinate is affine in the screen: uv = uv_offset + uv_scale * (ndc + 1) / 2, so
    that pixel (x, y) samples uv = uv_offset + uv_scale * ((x + 0.5) / W, (y + 0.5) / H).
    White vertex colours with the given alpha; the texture is the opaque atlas (alpha read
    as 1) or the transparent one; uv_rect (min.xy, size.xy) sets LUCID_INST_HAS_UV_RECT."""
    cam = Camera(width, height)
    ang = 0.013
    rot = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    ndc = np.array([[-1.2, -1.2], [1.2, -1.2], [1.2, 1.2], [-1.2, 1.2]]) @ rot.T
    positions = cam.ndc_to_view(ndc, np.full(4, depth)).astype(np.float32)
    uv = (np.asarray(uv_offset, np.float64)[None] +
          np.asarray(uv_scale, np.float64)[None] * (ndc + 1.0) * 0.5).astype(np.float32)
    flags = INST_HAS_VERTEX_COLORS | INST_HAS_VERTEX_TEX_COORDS | INST_HAS_ALBEDO_TEXTURE
    flags |= INST_TEX_OPAQUE if opaque else 0
    flags |= INST_HAS_UV_RECT if uv_rect is not None else 0
    inst = split_instances(1, flags)
    sc = Scene(width, height, positions, np.array([[0, 1, 2, 3]], np.uint32), inst,
               make_config(cam, len(inst)), name="textured_plane")
    rgb = np.ones((4, 3), np.float32)
    if color_format == "rgba8":
        sc.colors_rgba8 = pack_rgba8(rgb, np.full(4, alpha, np.float32))
    else:
        sc.colors_rgb = rgb
        sc.tri_alpha = np.full(2, alpha, np.float32)
    sc.tex_coords = uv
    if uv_rect is not None:
        sc.instance_uv_rects = np.asarray([uv_rect], np.float32)
    sc.textures = (texture, None) if opaque else (None, texture)
    return sc


def textured_planes(width: int, height: int, n: int, seed: int = 11,
                    color_format: str = "f32") -> Scene:
    """n tilted, overlapping quads with random depth slopes, each its own instance: albedo
    textures (opaque or transparent atlas, with or without a uv rectangle selecting one
    quadrant of the atlas), texture coordinates repeated 0.3-6 times across the quad (both
    magnified and minified levels, and anisotropic footprints), vertex colours and alpha, and
    every third instance with an instance colour; a quarter of the instances untextured."""
    rng = np.random.default_rng(seed)
    cam = Camera(width, height)
    ang = rng.uniform(0.01, 0.3, n)
    cx, cy = rng.uniform(-0.6, 0.6, (2, n))
    sx, sy = rng.uniform(0.3, 1.0, (2, n))
    corners = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64)
    ndc = np.zeros((n, 4, 2))
    depth = np.zeros((n, 4))
    for i in range(n):
        r = np.array([[np.cos(ang[i]), -np.sin(ang[i])], [np.sin(ang[i]), np.cos(ang[i])]])
        ndc[i] = (corners * [sx[i], sy[i]]) @ r.T + [cx[i], cy[i]]
        z0, s0, s1 = rng.uniform(3.0, 12.0), *rng.uniform(-2.0, 2.0, 2)
        depth[i] = z0 + corners[:, 0] * s0 + corners[:, 1] * s1
    pos = cam.ndc_to_view(ndc, depth).astype(np.float32).reshape(-1, 3)
    base = np.arange(n, dtype=np.uint32) * 4
    quads = np.stack([base, base + 1, base + 2, base + 3], axis=1).astype(np.uint32)
    inst = np.zeros(n, INSTANCE_DTYPE)
    inst["index_offset"] = np.arange(n) * 4
    inst["num_quads"] = 1
    rects = np.zeros((n, 4), np.float32)
    icol = np.full(n, 0xFFFFFFFF, np.uint32)
    uv = np.zeros((n, 4, 2), np.float32)
    for i in range(n):
        flags = INST_HAS_VERTEX_COLORS
        if i % 4 != 3:
            flags |= INST_HAS_VERTEX_TEX_COORDS | INST_HAS_ALBEDO_TEXTURE
            flags |= INST_TEX_OPAQUE if i % 2 == 0 else 0
            if i % 3 == 1:
                flags |= INST_HAS_UV_RECT
                qx, qy = rng.integers(0, 2, 2)
                rects[i] = (0.5 * qx, 0.5 * qy, 0.5, 0.5)
        if i % 3 == 0:
            flags |= INST_HAS_COLOR
            icol[i] = pack_rgba8(rng.uniform(0.3, 1.0, 3).astype(np.float32),
                                 np.float32(rng.uniform(0.5, 1.0)))
        inst["flags"][i] = flags
        reps = rng.uniform(0.3, 6.0, 2)
        off = rng.uniform(-1.0, 1.0, 2)
        uv[i] = off + reps * (corners + 1.0) * 0.5
    sc = Scene(width, height, pos, quads, inst, make_config(cam, n), name=f"textured_{n}")
    rgb = rng.uniform(0.4, 1.0, (n, 4, 3)).astype(np.float32).reshape(-1, 3)
    alpha = rng.uniform(0.3, 0.9, n).astype(np.float32)
    if color_format == "rgba8":
        sc.colors_rgba8 = pack_rgba8(rgb, np.repeat(alpha, 4))
    else:
        sc.colors_rgb = rgb
        sc.tri_alpha = np.repeat(alpha, 2)
    sc.instance_colors = icol
    sc.instance_uv_rects = rects
    sc.tex_coords = uv.reshape(-1, 2)
    sc.textures = (checker_texture(128, 64, seed=seed + 1),
                   checker_texture(64, 64, seed=seed + 2, alpha_range=(40, 255)))
    return sc
