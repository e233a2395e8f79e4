"""CPU ORACLE — TEST INFRASTRUCTURE ONLY: a pure-Python restatement of the reference's quad
generator (src/quad_generator.cpp:8-201), used by tests to check lucid_generate_quads.

Arithmetic is float32 in the reference's operation order (squareness and score), and ties of
the greedy score are broken by candidate index, as in the library.
"""

from __future__ import annotations

import heapq

import numpy as np

f32 = np.float32


def _normalize(v):
    d = f32(f32(v[0] * v[0]) + f32(v[1] * v[1])) + f32(v[2] * v[2])
    inv = f32(f32(1.0) / np.sqrt(f32(d)))
    return (f32(v[0] * inv), f32(v[1] * inv), f32(v[2] * inv))


def squareness(corners):
    """(4 - sum |dot| of consecutive edge directions) / 4 (quad_generator.cpp:8-17)."""
    e = []
    for i in range(4):
        a, b = corners[i], corners[(i + 1) % 4]
        e.append(_normalize((f32(b[0] - a[0]), f32(b[1] - a[1]), f32(b[2] - a[2]))))
    s = f32(0.0)
    for i in range(4):
        x, y = e[i], e[(i + 1) % 4]
        d = f32(f32(f32(x[0] * y[0]) + f32(x[1] * y[1])) + f32(x[2] * y[2]))
        s = f32(s + f32(abs(d)))
    return f32(f32(f32(4.0) - s) * f32(0.25))


def generate_quads(positions: np.ndarray, tris: np.ndarray, square_weight: float = 0.5) -> np.ndarray:
    pos = np.asarray(positions, np.float32)
    tris = [tuple(int(v) for v in t) for t in np.asarray(tris)]
    nt = len(tris)
    # triNeighbours (:19-45): the first triangle of a directed edge owns it
    edge = {}
    for i, t in enumerate(tris):
        for j in range(3):
            edge.setdefault((t[j], t[(j + 1) % 3]), i)
    nbr = [[-1, -1, -1] for _ in range(nt)]
    for i, t in enumerate(tris):
        for j in range(3):
            o = edge.get((t[(j + 1) % 3], t[j]))
            if o is not None and o != i:
                nbr[i][j] = o
    # quadNodes (:56-110)
    cand = []  # [tris, verts, conflicts, squareness]
    tri_quads = [[-1, -1, -1] for _ in range(nt)]
    for t0 in range(nt):
        a = tris[t0]
        for i in range(3):
            t1 = nbr[t0][i]
            if t1 == -1 or tri_quads[t0][i] != -1:
                continue
            if t0 not in nbr[t1]:
                continue
            j = nbr[t1].index(t0)
            opp = next((v for v in tris[t1] if v not in a), None)
            if opp is None:
                continue
            verts = (a[i], opp, a[(i + 1) % 3], a[(i + 2) % 3])
            sq = squareness([pos[v] for v in verts])
            q = len(cand)
            tri_quads[t0][i] = q
            tri_quads[t1][j] = q
            cand.append([(t0, t1), verts, [], sq])
    for t in range(nt):
        for i in range(3):
            q0, q1 = tri_quads[t][i], tri_quads[t][(i + 1) % 3]
            if q0 != -1 and q1 != -1:
                for x, y in ((q0, q1), (q1, q0)):
                    c = cand[x][2]
                    if y not in c and len(c) < 4:
                        c.append(y)
    # genQuads (:118-175): greedy independent set on degree - squareness * weight
    w = f32(square_weight)
    degree = [len(c[2]) for c in cand]
    score = lambda q: f32(f32(degree[q]) - f32(cand[q][3] * w))
    heap = [(score(q), q) for q in range(len(cand))]
    heapq.heapify(heap)
    current = {q: s for s, q in heap}
    state = [0] * len(cand)
    tri_done = [False] * nt
    while heap:
        s, q = heapq.heappop(heap)
        if current.get(q) != s or state[q]:
            continue  # stale entry or already decided
        state[q] = 2
        tri_done[cand[q][0][1]] = True
        for n in cand[q][2]:
            state[n] = 1
            for n2 in cand[n][2]:
                if not state[n2]:
                    degree[n2] -= 1
                    current[n2] = score(n2)
                    heapq.heappush(heap, (current[n2], n2))
    # output in triangle order (:177-200)
    out = []
    for t in range(nt):
        if tri_done[t]:
            continue
        sel = next((q for q in tri_quads[t] if q != -1 and state[q] == 2), -1)
        if sel < 0:
            a = tris[t]
            out.append((a[0], a[1], a[2], a[2]))
        else:
            out.append(cand[sel][1])
    return np.array(out, np.uint32).reshape(-1, 4)
