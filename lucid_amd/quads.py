"""Quad generator binding (lucid_generate_quads): pairs edge-adjacent triangles into quads
as the reference's Scene::generateQuads does (src/scene.cpp:237-247,
src/quad_generator.cpp:8-201). Host code in liblucid_amd.so; no GPU needed."""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def generate_quads(positions: np.ndarray, tris: np.ndarray, square_weight: float = 0.5) -> np.ndarray:
    """positions (V, 3) float32, tris (T, 3) uint32 -> quads (Q, 4) uint32 (triangles
    (v0, v1, v2) and (v0, v2, v3); unpaired triangles as (a, b, c, c))."""
    lib = _lib.load()
    pos = np.ascontiguousarray(positions, np.float32)
    tri = np.ascontiguousarray(tris, np.uint32)
    out = np.zeros((max(len(tri), 1), 4), np.uint32)
    n = ctypes.c_int64()
    _lib.check(lib.lucid_generate_quads(pos.ctypes.data, len(pos), tri.ctypes.data, len(tri),
                                        float(square_weight), out.ctypes.data, ctypes.byref(n)),
               "lucid_generate_quads")
    return out[:n.value]
