"""torch.autograd binding of the differentiable renderer.

Forward: ``lucid_render`` with fp32 samples (no RGBA8 quantisation of shaded samples, which
has no useful derivative). Backward: ``lucid_backward_ex``, the hand-written adjoint of sorted
front-to-back compositing; it recomputes the per-pixel sorted fragment lists instead of
storing them, and reuses the forward frame's bins and block lists when no other frame was
rendered in between (the saved colour and alpha tensors guard against in-place changes).
The reference renderer (src/lucid_renderer.h:21-96) has no backward pass; this is the
differentiable extension of the same forward.
"""

from __future__ import annotations

from . import _lib
from .renderer import GpuScene, LucidRenderer


def _torch():
    import torch
    return torch


class LucidRenderFunction:
    """Namespace for the autograd.Function (created lazily so that importing lucid_amd does
    not import torch)."""
    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            torch = _torch()

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, colors_rgb, tri_alpha, renderer: LucidRenderer,
                            gscene: GpuScene):
                    colors = colors_rgb.detach().contiguous().to(torch.float32)
                    alpha = tri_alpha.detach().contiguous().to(torch.float32)
                    out = renderer.alloc_outputs(rgba8=False, rgba32f=True, counts=False)
                    # by default the backward recomputes the per-pixel sorted fragment lists;
                    # with renderer.record_backward the frame stores each pixel's blend order
                    # for it instead (when autograd records this op)
                    record = bool(getattr(renderer, "record_backward", False)) and (
                        ctx.needs_input_grad[0] or ctx.needs_input_grad[1])
                    renderer.render(gscene, out, colors_rgb=colors, tri_alpha=alpha,
                                    record_backward=record)
                    ctx.renderer, ctx.gscene = renderer, gscene
                    ctx.save_for_backward(colors, alpha)
                    return out["image_f"]

                @staticmethod
                def backward(ctx, grad_image):
                    colors, alpha = ctx.saved_tensors
                    g_col, g_alpha = ctx.renderer.backward(ctx.gscene, grad_image, colors, alpha,
                                                           reuse_frame=True)
                    return g_col, g_alpha, None, None

            cls._fn = _Fn
        return cls._fn


def render_image(renderer: LucidRenderer, gscene: GpuScene, colors_rgb, tri_alpha):
    """Differentiable exact-OIT render: (H, W, 4) fp32 image from per-vertex RGB [V, 3] and
    per-triangle alpha [2 Q] (one per triangle id). The renderer must use fp32 samples."""
    if not renderer.opts & _lib.OPT_FP32_SAMPLES:
        raise ValueError("the differentiable path needs a renderer created with OPT_FP32_SAMPLES")
    return LucidRenderFunction.get().apply(colors_rgb, tri_alpha, renderer, gscene)


def render_image_logits(renderer: LucidRenderer, gscene: GpuScene, colors_rgb, opacity_logits):
    """As render_image, with per-triangle opacity given as logits (alpha = sigmoid)."""
    torch = _torch()
    return render_image(renderer, gscene, colors_rgb, torch.sigmoid(opacity_logits))
