"""Dense float64 PyTorch compositing reference for the backward pass (test infrastructure).

Given the oracle's exact fragment lists (every covered sample of every pixel, in blend order,
with its barycentrics and lighting factors), ``dense_image`` rebuilds the fp32-sample forward
of data/shaders/shared/shading.glsl:107-184,276-283 as differentiable float64 torch ops of the
per-vertex colours and per-triangle alpha:
  x   = w0 c[v0] + b0 c[v1] + b1 c[v2]                (w0 = 1 - b0 - b1)
  s   = clamp(L2S(S2L(x) * light), 0, 1)             (finalShading, funcs.glsl:261-271)
  out = clamp(sum_i s_i a_i T_i + T_end bg, 0, 1),   T_i = prod_{j<i} (1 - a_j)
torch.autograd of this is the reference gradient; central finite differences pin it.
"""

from __future__ import annotations

import numpy as np


def corner_vertices(scene) -> np.ndarray:
    """Vertex ids of the three corners of every triangle id (2 per quad): triangle 2q+s uses
    quad vertices (0, 1+s, 2+s), addressed through the instances."""
    q = scene.num_quads
    qi = np.zeros((q, 4), np.int64)
    inst = scene.instances
    done = 0
    for ins in inst:
        n = int(ins["num_quads"])
        base = int(ins["index_offset"]) // 4
        qi[done:done + n] = scene.quad_indices[base:base + n].astype(np.int64) + int(ins["vertex_offset"])
        done += n
    cv = np.zeros((2 * q, 3), np.int64)
    cv[0::2] = qi[:, [0, 1, 2]]
    cv[1::2] = qi[:, [0, 2, 3]]
    return cv


def _srgb_to_linear(torch, x):
    hi = torch.clamp((x + 0.055) / 1.055, min=1e-12) ** 2.4
    return torch.where(x < 0.04045, x / 12.92, hi)


def _linear_to_srgb(torch, y):
    hi = 1.055 * torch.clamp(y, min=1e-12) ** (1.0 / 2.4) - 0.055
    return torch.where(y < 0.0031308, 12.92 * y, hi)


class DenseCompositor:
    def __init__(self, scene, fragments: np.ndarray, background=(0.0, 0.0, 0.0)):
        import torch
        self.torch = torch
        self.w, self.h = scene.width, scene.height
        f = fragments
        cv = corner_vertices(scene)
        t = f["tri"].astype(np.int64)
        self.tri = torch.from_numpy(t)
        self.v = torch.from_numpy(cv[t])  # (F, 3)
        b = f["bary"].astype(np.float64)
        self.b0 = torch.from_numpy(b[:, 0])
        self.b1 = torch.from_numpy(b[:, 1])
        self.light = torch.from_numpy(f["light"].astype(np.float64))
        pix = f["y"].astype(np.int64) * self.w + f["x"].astype(np.int64)
        self.pix = torch.from_numpy(pix)
        # segments: fragments are grouped by pixel, in blend order within a pixel
        start = np.ones(len(pix), bool)
        start[1:] = pix[1:] != pix[:-1]
        self.seg_start = torch.from_numpy(np.flatnonzero(start))
        self.seg_id = torch.from_numpy(np.cumsum(start) - 1)
        self.seg_pix = self.pix[self.seg_start]
        self.bg = torch.tensor(background, dtype=torch.float64)

    def image(self, colors, tri_alpha):
        """(H, W, 3) float64 image from colors [V, 3] and tri_alpha [2Q] (float64 tensors)."""
        torch = self.torch
        c = colors
        w0 = (1.0 - self.b0) - self.b1
        x = (w0[:, None] * c[self.v[:, 0]] + self.b0[:, None] * c[self.v[:, 1]] +
             self.b1[:, None] * c[self.v[:, 2]])
        s = torch.clamp(_linear_to_srgb(torch, _srgb_to_linear(torch, x) * self.light), 0.0, 1.0)
        a = tri_alpha[self.tri]
        la = torch.log1p(-a)
        # exclusive segmented cumulative sum of log(1 - a)
        cs = torch.cumsum(la, 0)
        seg_base = (cs - la)[self.seg_start]  # inclusive sum before each segment's first
        excl = (cs - la) - seg_base[self.seg_id]
        T = torch.exp(excl)
        contrib = s * (a * T)[:, None]
        npx = self.w * self.h
        C = torch.zeros((npx, 3), dtype=torch.float64).index_add(0, self.pix, contrib)
        seg_total = torch.zeros(len(self.seg_start), dtype=torch.float64).index_add(
            0, self.seg_id, la)
        T_end = torch.ones(npx, dtype=torch.float64).index_copy(0, self.seg_pix,
                                                               torch.exp(seg_total))
        out = torch.clamp(C + T_end[:, None] * self.bg[None, :], 0.0, 1.0)
        return out.reshape(self.h, self.w, 3)
