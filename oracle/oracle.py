"""ctypes front-end of the C oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker. The product (lucid_amd) never imports it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblucid_oracle.so")

INFO_U32 = 1152
MODE_EXACT, MODE_REF3, MODE_COUNTS = 0, 1, 2
OPT_FP32_SAMPLES, OPT_ADDITIVE = 0x1, 0x2


class _Texture(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("num_levels", ctypes.c_int)]


class _Inputs(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("height", ctypes.c_int),
        ("config", ctypes.c_void_p),
        ("instances", ctypes.c_void_p), ("num_instances", ctypes.c_int),
        ("instance_colors", ctypes.c_void_p),
        ("quad_indices", ctypes.c_void_p),
        ("positions", ctypes.c_void_p), ("num_verts", ctypes.c_int),
        ("colors_rgba8", ctypes.c_void_p), ("colors_rgb", ctypes.c_void_p),
        ("tri_alpha", ctypes.c_void_p), ("normals", ctypes.c_void_p),
        ("mode", ctypes.c_int), ("opts", ctypes.c_uint32),
        ("tex_coords", ctypes.c_void_p), ("instance_uv_rects", ctypes.c_void_p),
        ("textures", _Texture * 2),
    ]


class _Outputs(ctypes.Structure):
    _fields_ = [
        ("info", ctypes.c_void_p), ("quad_aabb", ctypes.c_void_p),
        ("tri_records", ctypes.c_void_p), ("tri_normals", ctypes.c_void_p),
        ("quad_colors", ctypes.c_void_p),
        ("bin_quads", ctypes.c_void_p), ("bin_tris", ctypes.c_void_p),
        ("bin_capacity", ctypes.c_size_t),
        ("image", ctypes.c_void_p), ("image_f", ctypes.c_void_p),
        ("frag_counts", ctypes.c_void_p), ("bin_levels", ctypes.c_void_p),
        ("num_fragments", ctypes.c_uint64), ("num_ref_errors", ctypes.c_uint32),
        ("frags", ctypes.c_void_p), ("frags_cap", ctypes.c_size_t),
        ("num_frags_dumped", ctypes.c_size_t),
        ("dump_x0", ctypes.c_int), ("dump_y0", ctypes.c_int),
        ("dump_x1", ctypes.c_int), ("dump_y1", ctypes.c_int),
    ]


# One exact-mode sample (oracle_fragment): pixel, triangle, position in the pixel's sorted
# list, depth, barycentrics of corners 1 and 2, lighting factor per channel.
FRAGMENT_DTYPE = np.dtype([("x", np.uint32), ("y", np.uint32), ("tri", np.uint32),
                           ("order", np.uint32), ("inv_ray", np.float32),
                           ("bary", np.float32, 2), ("light", np.float32, 3)])


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.oracle_render.argtypes = [ctypes.POINTER(_Inputs), ctypes.POINTER(_Outputs)]
        _lib.oracle_render.restype = ctypes.c_int
        _lib.oracle_encode_aabb28.restype = ctypes.c_uint32
        _lib.oracle_encode_aabb28.argtypes = [ctypes.c_uint32] * 4
        _lib.oracle_encode_rgba8.restype = ctypes.c_uint32
        _lib.oracle_encode_rgba8.argtypes = [ctypes.c_float] * 4
        _lib.oracle_encode_normal_uint.restype = ctypes.c_uint32
        _lib.oracle_encode_normal_uint.argtypes = [ctypes.c_float] * 3
        _lib.oracle_f2i.restype = ctypes.c_int32
        _lib.oracle_f2i.argtypes = [ctypes.c_float]
        _lib.oracle_f2u.restype = ctypes.c_uint32
        _lib.oracle_f2u.argtypes = [ctypes.c_float]
        _lib.oracle_num_threads.restype = ctypes.c_int
        _lib.oracle_texture_grad.restype = None
        _lib.oracle_texture_grad.argtypes = [ctypes.POINTER(_Texture)] + [ctypes.c_float] * 6 + [
            ctypes.c_void_p]
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def _texture(tex, keep) -> _Texture:
    """oracle_texture of a lucid_amd.scenes.Texture (or an empty one for None)."""
    if tex is None:
        return _Texture(None, 0, 0, 0)
    packed = np.ascontiguousarray(tex.packed(), np.uint32)
    keep.append(packed)
    return _Texture(packed.ctypes.data, tex.width, tex.height, tex.num_levels)


def texture_grad(tex, s, v, dsx, dvx, dsy, dvy) -> np.ndarray:
    """The oracle's textureGrad of one atlas at one point (RGBA float32)."""
    keep = []
    t = _texture(tex, keep)
    out = np.zeros(4, np.float32)
    lib().oracle_texture_grad(ctypes.byref(t), s, v, dsx, dvx, dsy, dvy, out.ctypes.data)
    return out


def bin_count(width: int, height: int) -> int:
    return ((width + 31) // 32) * ((height + 31) // 32)


def render(scene, mode: int = MODE_EXACT, opts: int = 0, bin_capacity: Optional[int] = None,
           want_records: bool = True, want_fragments: bool = False,
           fragment_rect: Optional[tuple] = None) -> dict:
    """Runs the oracle on a lucid_amd.scenes.Scene; returns numpy outputs. With
    want_fragments (exact mode), res["fragments"] lists every sample (FRAGMENT_DTYPE),
    sorted by pixel (row-major) and blend order; fragment_rect = (x0, y0, x1, y1) keeps only
    the samples of pixels in [x0, x1) x [y0, y1)."""
    L = lib()
    w, h = scene.width, scene.height
    nb = bin_count(w, h)
    q = scene.num_quads
    if bin_capacity is None:
        # bin lists hold a quad once per bin it touches: start from 8 per quad and grow
        # (the oracle reports -3 when its lists overflow)
        cap = max(1 << 16, 8 * q)
        while True:
            try:
                return render(scene, mode, opts, cap, want_records, want_fragments,
                              fragment_rect)
            except RuntimeError as e:
                if str(e) != "oracle_render failed: -3" or cap >= 2 * q * nb:
                    raise
                cap = min(cap * 4, 2 * q * nb)
    keep = []

    def c(a, dtype):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a

    cfg = np.frombuffer(scene.config, np.uint8).copy()
    keep.append(cfg)
    inp = _Inputs(
        w, h, cfg.ctypes.data,
        _ptr(c(scene.instances, scene.instances.dtype)), len(scene.instances),
        _ptr(c(scene.instance_colors, np.uint32)),
        _ptr(c(scene.quad_indices, np.uint32)),
        _ptr(c(scene.positions, np.float32)), scene.num_verts,
        _ptr(c(scene.colors_rgba8, np.uint32)), _ptr(c(scene.colors_rgb, np.float32)),
        _ptr(c(scene.tri_alpha, np.float32)), _ptr(c(scene.normals, np.uint32)),
        mode, opts,
        _ptr(c(getattr(scene, "tex_coords", None), np.float32)),
        _ptr(c(getattr(scene, "instance_uv_rects", None), np.float32)),
        (_Texture * 2)(*[_texture(t, keep) for t in (getattr(scene, "textures", None) or
                                                       (None, None))]),
    )
    out_np = {
        "info": np.zeros(INFO_U32 + 10 * nb, np.uint32),
        "quad_aabb": np.zeros(q, np.uint32),
        "tri_records": np.zeros((2 * q, 20) if want_records else (0, 20), np.uint32),
        "tri_normals": np.zeros(2 * q, np.uint32),
        "quad_colors": np.zeros((q, 4), np.uint32),
        "bin_quads": np.zeros(bin_capacity, np.uint32),
        "bin_tris": np.zeros(bin_capacity, np.uint32),
        "image": np.zeros((h, w, 4), np.uint8),
        "image_f": np.zeros((h, w, 4), np.float32),
        "frag_counts": np.zeros((h, w), np.uint32),
        "bin_levels": np.zeros(nb, np.uint32),
    }
    out = _Outputs(
        out_np["info"].ctypes.data, out_np["quad_aabb"].ctypes.data,
        out_np["tri_records"].ctypes.data if want_records else None,
        out_np["tri_normals"].ctypes.data, out_np["quad_colors"].ctypes.data,
        out_np["bin_quads"].ctypes.data, out_np["bin_tris"].ctypes.data, bin_capacity,
        out_np["image"].ctypes.data, out_np["image_f"].ctypes.data,
        out_np["frag_counts"].ctypes.data, out_np["bin_levels"].ctypes.data, 0, 0,
        None, 0, 0, *(fragment_rect if fragment_rect is not None else (0, 0, 0, 0)),
    )
    rc = L.oracle_render(ctypes.byref(inp), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle_render failed: {rc}")
    if want_fragments:
        # second run with a buffer sized from the first run's fragment counts
        if fragment_rect is None:
            n = int(out.num_fragments)
        else:
            x0, y0, x1, y1 = fragment_rect
            n = int(out_np["frag_counts"][y0:y1, x0:x1].sum())
        frags = np.zeros(max(n, 1), FRAGMENT_DTYPE)
        out.frags, out.frags_cap, out.num_frags_dumped = frags.ctypes.data, n, 0
        rc = L.oracle_render(ctypes.byref(inp), ctypes.byref(out))
        if rc != 0 or out.num_frags_dumped != n:
            raise RuntimeError(f"oracle_render fragment dump failed: {rc}")
        frags = frags[:n]
        idx = np.lexsort((frags["order"], frags["x"], frags["y"]))
        out_np["fragments"] = frags[idx]
    out_np["num_fragments"] = int(out.num_fragments)
    out_np["num_ref_errors"] = int(out.num_ref_errors)
    out_np["bin_count"] = nb
    return out_np


def counts(info: np.ndarray, nb: int, k: int) -> np.ndarray:
    return info[INFO_U32 + k * nb: INFO_U32 + (k + 1) * nb]


def bin_lists(res: dict) -> tuple:
    """Per-bin sorted quad/tri lists as tuples of numpy arrays."""
    nb = res["bin_count"]
    info = res["info"]
    qc, qo = counts(info, nb, 0), counts(info, nb, 1)
    tc, to = counts(info, nb, 3), counts(info, nb, 4)
    ql = [np.sort(res["bin_quads"][qo[b]:qo[b] + qc[b]]) for b in range(nb)]
    tl = [np.sort(res["bin_tris"][to[b]:to[b] + tc[b]]) for b in range(nb)]
    return ql, tl
