"""Quad generator (lucid_generate_quads, host code in the library) against the Python
restatement of src/quad_generator.cpp in oracle/quadgen.py, plus closed-form and invariant
checks. No GPU needed."""

import sys
import os

import numpy as np
import pytest

from lucid_amd import scenes
from lucid_amd.quads import generate_quads

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import quadgen  # noqa: E402


def _tri_set(t):
    """Triangles as rotation-normalised tuples (winding kept)."""
    out = []
    for a, b, c in t:
        r = [(a, b, c), (b, c, a), (c, a, b)]
        out.append(min(r))
    return sorted(out)


def _quad_tris(q):
    tris = []
    for v0, v1, v2, v3 in q:
        tris.append((v0, v1, v2))
        if v2 != v3:
            tris.append((v0, v2, v3))
    return tris


def _grid(n, jitter, seed):
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(n + 1), np.arange(n + 1))
    pos = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1).astype(np.float32)
    pos[:, :2] += rng.uniform(-jitter, jitter, (len(pos), 2)).astype(np.float32)
    tris = []
    for y in range(n):
        for x in range(n):
            a, b, c, d = y * (n + 1) + x, y * (n + 1) + x + 1, (y + 1) * (n + 1) + x + 1, (y + 1) * (n + 1) + x
            tris += [(a, b, c), (a, c, d)] if (x + y) % 3 else [(a, b, d), (b, c, d)]
    return pos, np.array(tris, np.uint32)


@pytest.mark.parametrize("mesh", ["icosphere2", "icosphere3", "grid", "grid_jitter"])
def test_matches_restatement(mesh):
    if mesh.startswith("icosphere"):
        pos, tris = scenes.icosphere(int(mesh[-1]))
    else:
        pos, tris = _grid(12, 0.3 if mesh.endswith("jitter") else 0.0, seed=5)
    q = generate_quads(pos, tris, 0.5)
    r = quadgen.generate_quads(pos, tris, 0.5)
    assert np.array_equal(q, r)


def test_every_triangle_exactly_once_and_pairs_share_an_edge():
    pos, tris = scenes.icosphere(3)
    q = generate_quads(pos, tris, 0.5)
    assert _tri_set(_quad_tris(q)) == _tri_set(tris.tolist())
    for v0, v1, v2, v3 in q:
        if v2 != v3:
            assert len({v0, v1, v2, v3}) == 4  # two triangles sharing edge (v0, v2)


def test_square_grid_pairs_into_its_squares():
    """Closed form: an axis-aligned grid of squares, each split along a diagonal, pairs every
    triangle with its square partner (squareness 1), leaving no degenerate quad."""
    n = 10
    xs, ys = np.meshgrid(np.arange(n + 1), np.arange(n + 1))
    pos = np.stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)], 1).astype(np.float32)
    tris = []
    for y in range(n):
        for x in range(n):
            a, b, c, d = y * (n + 1) + x, y * (n + 1) + x + 1, (y + 1) * (n + 1) + x + 1, (y + 1) * (n + 1) + x
            tris += [(a, b, c), (a, c, d)]
    q = generate_quads(pos, np.array(tris, np.uint32), 0.5)
    assert len(q) == n * n and np.all(q[:, 2] != q[:, 3])
    for quad in q:
        p = pos[quad]
        assert abs(quadgen.squareness(p) - 1.0) < 1e-6


def test_order_follows_triangles_and_unpaired_are_degenerate():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 5, 0], [6, 5, 0], [5, 6, 0]], np.float32)
    tris = np.array([[3, 4, 5], [0, 1, 2]], np.uint32)  # two isolated triangles
    q = generate_quads(pos, tris, 0.5)
    assert q.tolist() == [[3, 4, 5, 5], [0, 1, 2, 2]]
