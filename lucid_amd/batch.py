"""Batched rendering: one frame of many scenes (each with its own renderer, of any size)
recorded into one hipGraph and launched at once (lucid_batch_* in include/lucid_c.h).

The reference renders one scene per frame (src/lucid_renderer.cpp:319-350); this is the
batch extension of the same forward. Frames are spread over forked streams inside the
graph so that small scenes run side by side; the graph is recorded again only when a
frame's constants or buffers change.
"""

from __future__ import annotations

import ctypes
from typing import List, Sequence

from . import _lib
from .renderer import GpuScene, LucidRenderer


class LucidBatch:
    def __init__(self, renderers: Sequence[LucidRenderer]):
        self.lib = _lib.load()
        self.renderers = list(renderers)
        arr = (ctypes.c_void_p * len(self.renderers))(*[r.handle.value for r in self.renderers])
        handle = ctypes.c_void_p()
        _lib.check(self.lib.lucid_batch_create(arr, len(self.renderers), ctypes.byref(handle)),
                   "lucid_batch_create")
        self.handle = handle
        self._keep = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.lucid_batch_destroy(h)
            self.handle = None

    def render_async(self, scenes: Sequence[GpuScene], outputs: Sequence[dict], stream=None):
        import torch
        n = len(self.renderers)
        if len(scenes) != n or len(outputs) != n:
            raise ValueError("one scene and one output set per renderer")
        st = stream if stream is not None else torch.cuda.current_stream()
        ins = (_lib.FrameInputs * n)(*[g.inputs() for g in scenes])
        outs = (_lib.FrameOutputs * n)(*[LucidRenderer._outputs(o) for o in outputs])
        self._keep = (ins, outs, scenes)  # host arrays referenced by the inputs
        _lib.check(self.lib.lucid_batch_render(self.handle, ins, outs,
                                               ctypes.c_void_p(st.cuda_stream)),
                   "lucid_batch_render")

    def finish(self) -> int:
        return self.lib.lucid_batch_finish(self.handle)

    def render(self, scenes: Sequence[GpuScene], outputs: Sequence[dict], stream=None,
               max_retries: int = 4) -> List[dict]:
        for _ in range(max_retries + 1):
            self.render_async(scenes, outputs, stream)
            rc = self.finish()
            if rc == _lib.LUCID_OK:
                return list(outputs)
            if rc != _lib.LUCID_ERR_CAPACITY:
                _lib.check(rc, "lucid_batch_finish")
        raise _lib.LucidError(_lib.LUCID_ERR_CAPACITY, "batch render")
