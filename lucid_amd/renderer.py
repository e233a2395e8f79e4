"""Python mirror of the reference's LucidRenderer (src/lucid_renderer.h:21-96).

``LucidRenderer(view_size, opts)`` ≙ ``LucidRenderer::exConstruct`` and
``render(scene)`` ≙ ``LucidRenderer::render`` (src/lucid_renderer.cpp:319-350);
``verify_info``/``get_stats`` ≙ ``verifyInfo``/``getStats`` (:629-836).
Device memory and streams come from PyTorch; all compute runs in
liblucid_amd.so's HIP kernels through the C ABI.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .scenes import Scene

BIN_LEVEL_EMPTY, BIN_LEVEL_LOW, BIN_LEVEL_HIGH = 0, 2, 4


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("lucid_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch


class GpuScene:
    """Device-resident copy of a Scene (the C ABI's device arrays)."""

    def __init__(self, scene: Scene, device: int = 0):
        torch = _torch()
        dev = torch.device("cuda", device)
        self.scene = scene

        def up(a, dtype):
            if a is None:
                return None
            return torch.from_numpy(np.ascontiguousarray(a).view(dtype).reshape(-1)).to(dev)

        self.quad_indices = up(scene.quad_indices.astype(np.uint32), np.int32)
        self.positions = up(scene.positions.astype(np.float32), np.float32)
        self.colors_rgba8 = up(None if scene.colors_rgba8 is None else
                               scene.colors_rgba8.astype(np.uint32), np.int32)
        self.colors_rgb = up(None if scene.colors_rgb is None else
                             scene.colors_rgb.astype(np.float32), np.float32)
        self.tri_alpha = up(None if scene.tri_alpha is None else
                            scene.tri_alpha.astype(np.float32), np.float32)
        self.normals = up(None if scene.normals is None else
                          scene.normals.astype(np.uint32), np.int32)
        self.tex_coords = up(None if scene.tex_coords is None else
                             scene.tex_coords.astype(np.float32), np.float32)
        # atlases: (device texels, width, height, levels) or None
        self.textures = []
        for t in (scene.textures or (None, None)):
            self.textures.append(None if t is None else
                                 (up(t.packed().astype(np.uint32), np.int32), t.width, t.height,
                                  t.num_levels))
        self.instance_uv_rects = (None if scene.instance_uv_rects is None else
                                  np.ascontiguousarray(scene.instance_uv_rects, np.float32))
        self.instances = np.ascontiguousarray(scene.instances)
        self.instance_colors = (None if scene.instance_colors is None else
                                np.ascontiguousarray(scene.instance_colors, np.uint32))
        self.config = np.frombuffer(scene.config, np.uint8).copy()

    def inputs(self, colors_rgb=None, tri_alpha=None) -> _lib.FrameInputs:
        """The C ABI's frame inputs; colors_rgb / tri_alpha (contiguous fp32 device
        tensors) replace the scene's own, as the differentiable path does."""
        def p(t):
            return None if t is None else t.data_ptr()

        def tex(t):
            return _lib.Texture(None, 0, 0, 0) if t is None else _lib.Texture(p(t[0]), *t[1:])
        return _lib.FrameInputs(
            self.instances.ctypes.data, len(self.instances),
            None if self.instance_colors is None else self.instance_colors.ctypes.data,
            self.config.ctypes.data,
            p(self.quad_indices), p(self.positions), int(self.scene.num_verts),
            p(self.colors_rgba8),
            p(colors_rgb if colors_rgb is not None else self.colors_rgb),
            p(tri_alpha if tri_alpha is not None else self.tri_alpha), p(self.normals),
            p(self.tex_coords),
            None if self.instance_uv_rects is None else self.instance_uv_rects.ctypes.data,
            tex(self.textures[0]), tex(self.textures[1]))


class LucidRenderer:
    def __init__(self, view_size, opts: int = 0, device: int = 0, max_quads: int = 0):
        torch = _torch()
        self._torch = torch
        self.lib = _lib.load()
        self.width, self.height = int(view_size[0]), int(view_size[1])
        self.opts = opts
        self.device = device
        # autograd (lucid_amd.autograd): False = the backward recomputes every pixel's sorted
        # fragment list; True = differentiable frames store each pixel's blend order for the
        # backward (about one byte per fragment), which then skips that walk
        self.record_backward = False
        torch.cuda.set_device(device)
        handle = ctypes.c_void_p()
        ci = _lib.CreateInfo(self.width, self.height, opts, device, max_quads)
        _lib.check(self.lib.lucid_create(ctypes.byref(ci), ctypes.byref(handle)), "lucid_create")
        self.handle = handle
        self.bin_count = self.lib.lucid_bin_count(handle)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.lucid_destroy(h)
            self.handle = None

    # accessors (src/lucid_renderer.h:43-49)
    def bin_size(self) -> int:
        return self.lib.lucid_bin_size(self.handle)

    def block_size(self) -> int:
        return self.lib.lucid_block_size(self.handle)

    def max_visible_quads(self) -> int:
        return self.lib.lucid_max_visible_quads(self.handle)

    def max_scene_quads(self) -> int:
        return self.lib.lucid_max_scene_quads(self.handle)

    # tuning (performance only; see lucid_tuning in include/lucid_c.h)
    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        _lib.check(self.lib.lucid_get_tuning(self.handle, ctypes.byref(t)), "lucid_get_tuning")
        return t.as_dict()

    def set_tuning(self, slice_words: int = 1, presort_min: int = 128,
                   samples_per_wave: int = 1024, window: int = 1) -> None:
        t = _lib.Tuning(slice_words, presort_min, samples_per_wave, window)
        _lib.check(self.lib.lucid_set_tuning(self.handle, ctypes.byref(t)), "lucid_set_tuning")

    def autotune(self, gscene: "GpuScene", outputs: Optional[dict] = None, iters: int = 3,
                 stream=None) -> dict:
        """Times every tuning candidate on this scene and keeps the fastest."""
        torch = self._torch
        if outputs is None:
            outputs = self.alloc_outputs()
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        best = _lib.Tuning()
        inp = gscene.inputs()
        _lib.check(self.lib.lucid_autotune(self.handle, ctypes.byref(inp),
                                           ctypes.byref(self._outputs(outputs)),
                                           ctypes.c_void_p(st.cuda_stream), iters,
                                           ctypes.byref(best)), "lucid_autotune")
        return best.as_dict()

    @staticmethod
    def _outputs(outputs: dict) -> _lib.FrameOutputs:
        return _lib.FrameOutputs(
            outputs["image"].data_ptr() if outputs.get("image") is not None else None,
            outputs["image_f"].data_ptr() if outputs.get("image_f") is not None else None,
            outputs["frag_counts"].data_ptr() if outputs.get("frag_counts") is not None else None)

    def alloc_outputs(self, rgba8: bool = True, rgba32f: bool = False, counts: bool = True):
        torch = self._torch
        dev = torch.device("cuda", self.device)
        h, w = self.height, self.width
        return {
            "image": torch.empty((h, w, 4), dtype=torch.uint8, device=dev) if rgba8 else None,
            "image_f": torch.empty((h, w, 4), dtype=torch.float32, device=dev) if rgba32f else None,
            "frag_counts": torch.empty((h, w), dtype=torch.int32, device=dev) if counts else None,
        }

    def render_async(self, gscene: GpuScene, outputs: dict, stream=None, colors_rgb=None,
                     tri_alpha=None, record_backward: bool = False) -> None:
        """Enqueues one frame on `stream` (torch stream or None = current stream). With
        `record_backward` (fp32-sample renderers) the frame also records the backward's
        first pass for a following backward(..., reuse_frame=True)."""
        torch = self._torch
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        o = _lib.FrameOutputs(
            outputs["image"].data_ptr() if outputs.get("image") is not None else None,
            outputs["image_f"].data_ptr() if outputs.get("image_f") is not None else None,
            outputs["frag_counts"].data_ptr() if outputs.get("frag_counts") is not None else None)
        inp = gscene.inputs(colors_rgb, tri_alpha)
        flags = _lib.RENDER_RECORD_BACKWARD if record_backward else 0
        _lib.check(self.lib.lucid_render_ex(self.handle, ctypes.byref(inp), ctypes.byref(o), flags,
                                            ctypes.c_void_p(st.cuda_stream)), "lucid_render_ex")

    def finish(self) -> int:
        return self.lib.lucid_finish(self.handle)

    def frame_recorded(self) -> bool:
        """True if the last finished frame holds a backward record (render(...,
        record_backward=True) whose record fitted the backward buffers)."""
        return bool(self.lib.lucid_frame_recorded(self.handle))

    def render(self, gscene: GpuScene, outputs: Optional[dict] = None, stream=None,
               max_retries: int = 4, colors_rgb=None, tri_alpha=None,
               record_backward: bool = False) -> dict:
        """Renders one frame and waits for it; re-renders after an internal buffer grew."""
        if outputs is None:
            outputs = self.alloc_outputs()
        for _ in range(max_retries + 1):
            self.render_async(gscene, outputs, stream, colors_rgb, tri_alpha, record_backward)
            rc = self.finish()
            if rc == _lib.LUCID_OK:
                return outputs
            if rc != _lib.LUCID_ERR_CAPACITY:
                _lib.check(rc, "lucid_finish")
        raise _lib.LucidError(_lib.LUCID_ERR_CAPACITY, "render")

    def backward(self, gscene: GpuScene, grad_image, colors_rgb=None, tri_alpha=None,
                 stream=None, max_retries: int = 4, reuse_frame: bool = False):
        """Adjoint of the fp32-sample forward: dL/d(image_f) (H x W x 4 fp32 device tensor)
        -> (dL/d(colors_rgb) [V, 3], dL/d(tri_alpha) [2 Q]). The per-pixel sorted fragment
        lists are recomputed. With `reuse_frame`, the setup, bins and block lists of the
        renderer's last finished frame are reused when it was rendered from the same inputs
        (whose contents must be unchanged), and so is the first pass if that frame recorded
        it (render(..., record_backward=True)); the result is identical either way."""
        torch = self._torch
        dev = torch.device("cuda", self.device)
        st = stream if stream is not None else torch.cuda.current_stream(self.device)
        g_img = grad_image.contiguous().to(device=dev, dtype=torch.float32)
        if tuple(g_img.shape) != (self.height, self.width, 4):
            raise ValueError(f"grad_image must be {(self.height, self.width, 4)}")
        nv = int(gscene.scene.num_verts)
        g_col = torch.empty((nv, 3), dtype=torch.float32, device=dev)
        g_alpha = torch.empty((2 * gscene.scene.num_quads,), dtype=torch.float32, device=dev)
        inp = gscene.inputs(colors_rgb, tri_alpha)
        flags = _lib.BACKWARD_REUSE_FRAME if reuse_frame else 0
        for _ in range(max_retries + 1):
            rc = self.lib.lucid_backward_ex(self.handle, ctypes.byref(inp), g_img.data_ptr(),
                                            g_col.data_ptr(), g_alpha.data_ptr(), flags,
                                            ctypes.c_void_p(st.cuda_stream))
            if rc == _lib.LUCID_OK:
                return g_col, g_alpha
            if rc != _lib.LUCID_ERR_CAPACITY:
                _lib.check(rc, "lucid_backward_ex")
        raise _lib.LucidError(_lib.LUCID_ERR_CAPACITY, "lucid_backward")

    # ------------------------------------------------------------ readback / stats

    def read_info(self) -> np.ndarray:
        n = _lib.INFO_U32 + 10 * self.bin_count
        buf = np.zeros(n, np.uint32)
        _lib.check(self.lib.lucid_read_info(self.handle, buf.ctypes.data, n), "lucid_read_info")
        return buf

    def verify_info(self) -> int:
        rc = self.lib.lucid_verify_info(self.handle)
        if rc < 0:
            _lib.check(rc, "lucid_verify_info")
        return rc

    def read_bin_lists(self):
        nq, nt = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(self.lib.lucid_read_bin_lists(self.handle, None, 0, None, 0, ctypes.byref(nq),
                                                 ctypes.byref(nt)), "lucid_read_bin_lists")
        q = np.zeros(max(1, nq.value), np.uint32)
        t = np.zeros(max(1, nt.value), np.uint32)
        _lib.check(self.lib.lucid_read_bin_lists(self.handle, q.ctypes.data, q.size, t.ctypes.data,
                                                 t.size, ctypes.byref(nq), ctypes.byref(nt)),
                   "lucid_read_bin_lists")
        return q[:nq.value], t[:nt.value]

    def read_setup(self, num_quads: int):
        recs = np.zeros((2 * num_quads, 20), np.uint32)
        aabb = np.zeros(num_quads, np.uint32)
        _lib.check(self.lib.lucid_read_setup(self.handle, recs.ctypes.data, 2 * num_quads,
                                             aabb.ctypes.data, num_quads), "lucid_read_setup")
        return recs, aabb

    STAGES = ("setup", "binning", "rowgen_low", "rowgen_high_sort", "block_raster", "fallback")

    def stage_times(self) -> dict:
        """Per-stage GPU ms of the last frame (needs opts & OPT_TIMERS)."""
        ms = np.zeros(len(self.STAGES), np.float32)
        n = self.lib.lucid_stage_times(self.handle, ms.ctypes.data, len(ms))
        if n < 0:
            _lib.check(n, "lucid_stage_times")
        return {name: float(ms[i]) for i, name in enumerate(self.STAGES[:n])}

    def get_stats(self) -> dict:
        """Counters of the last frame (src/lucid_renderer.cpp:709-836)."""
        info = self.read_info()
        nb = self.bin_count
        cnt = info[_lib.INFO_U32:]
        quad_counts, tri_counts = cnt[0:nb], cnt[3 * nb:4 * nb]
        levels = info[5:10].astype(np.int64)
        return {
            "input_quads": int(info[0]),
            "visible_quads": int(info[1]) + int(info[2]),
            "visible_small": int(info[1]), "visible_large": int(info[2]),
            "rejected": {"other": int(info[32]), "backface": int(info[33]),
                         "frustum": int(info[34]), "between_samples": int(info[35])},
            "bin_quads": int(quad_counts.sum()), "bin_tris": int(tri_counts.sum()),
            "max_quads_per_bin": int(quad_counts.max(initial=0)),
            "max_tris_per_bin": int(tri_counts.max(initial=0)),
            "empty_bins": int(levels[BIN_LEVEL_EMPTY]), "low_bins": int(levels[BIN_LEVEL_LOW]),
            "high_bins": int(levels[BIN_LEVEL_HIGH]),
            "promoted_bins": int(levels.sum() - nb),
            "fragments": int(info[60]), "block_tris": int(info[61]),
            "fallback_pixels": int(info[62]), "oversized_blocks": int(info[63]),
        }
