"""ctypes bindings of liblucid_amd.so (include/lucid_c.h).

The library is built in-tree (``lucid_amd/liblucid_amd.so``) by
``__graft_entry__.build()`` / ``make -C lucid_amd/csrc``. Loading fails
loudly when it is missing: there is no CPU fallback on the product path.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblucid_amd.so")

LUCID_OK = 0
LUCID_ERR_INVALID_ARG = -1
LUCID_ERR_OUT_OF_MEMORY = -2
LUCID_ERR_HIP = -3
LUCID_ERR_NO_DEVICE = -4
LUCID_ERR_CAPACITY = -5
LUCID_ERR_UNSUPPORTED = -6

OPT_TIMERS = 1 << 4
OPT_ADDITIVE_BLENDING = 1 << 5
OPT_VISUALIZE_ERRORS = 1 << 6
OPT_ALPHA_THRESHOLD = 1 << 7
OPT_FP32_SAMPLES = 1 << 16
OPT_UNSORTED = 1 << 17

BACKWARD_REUSE_FRAME = 1
RENDER_RECORD_BACKWARD = 1

INFO_U32 = 1152

# Every symbol include/lucid_c.h declares.
EXPORTED = (
    "lucid_error_string", "lucid_create", "lucid_destroy", "lucid_render", "lucid_render_ex",
    "lucid_frame_recorded", "lucid_finish",
    "lucid_read_info", "lucid_verify_info", "lucid_read_bin_lists", "lucid_read_setup",
    "lucid_stage_times", "lucid_backward", "lucid_backward_ex",
    "lucid_set_tuning", "lucid_get_tuning", "lucid_autotune",
    "lucid_batch_create", "lucid_batch_destroy", "lucid_batch_render", "lucid_batch_finish",
    "lucid_generate_quads",
    "lucid_bin_count", "lucid_bin_size", "lucid_block_size", "lucid_max_visible_quads",
    "lucid_max_scene_quads", "lucid_render_noalloc",
)


class CreateInfo(ctypes.Structure):
    _fields_ = [
        ("view_width", ctypes.c_int), ("view_height", ctypes.c_int),
        ("opts", ctypes.c_uint32), ("device", ctypes.c_int), ("max_quads", ctypes.c_int64),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [("slice_words", ctypes.c_int), ("presort_min", ctypes.c_int),
                ("samples_per_wave", ctypes.c_int), ("window", ctypes.c_int)]

    def as_dict(self):
        return {"slice_words": self.slice_words, "presort_min": self.presort_min,
                "samples_per_wave": self.samples_per_wave, "window": self.window}


class Texture(ctypes.Structure):
    _fields_ = [("texels", ctypes.c_void_p), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("num_levels", ctypes.c_int)]


class FrameInputs(ctypes.Structure):
    _fields_ = [
        ("instances", ctypes.c_void_p), ("num_instances", ctypes.c_int),
        ("instance_colors", ctypes.c_void_p), ("config", ctypes.c_void_p),
        ("quad_indices", ctypes.c_void_p), ("positions", ctypes.c_void_p),
        ("num_verts", ctypes.c_int64),
        ("colors_rgba8", ctypes.c_void_p), ("colors_rgb", ctypes.c_void_p),
        ("tri_alpha", ctypes.c_void_p), ("normals", ctypes.c_void_p),
        ("tex_coords", ctypes.c_void_p), ("instance_uv_rects", ctypes.c_void_p),
        ("opaque_texture", Texture), ("transparent_texture", Texture),
    ]


class FrameOutputs(ctypes.Structure):
    _fields_ = [
        ("image_rgba8", ctypes.c_void_p), ("image_rgba32f", ctypes.c_void_p),
        ("frag_counts", ctypes.c_void_p),
    ]


class LucidError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = load().lucid_error_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


_lib = None


def load(path: str = LIB_PATH):
    """Loads the HIP library (after torch, so that both share one HIP runtime)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C lucid_amd/csrc` (there is no CPU fallback)")
    try:
        import torch  # noqa: F401  (binds liblucid_amd to torch's libamdhip64)
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    vp, i32, u32, i64, sz = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int64,
                             ctypes.c_size_t)
    lib.lucid_error_string.restype = ctypes.c_char_p
    lib.lucid_error_string.argtypes = [i32]
    lib.lucid_create.restype = i32
    lib.lucid_create.argtypes = [ctypes.POINTER(CreateInfo), ctypes.POINTER(vp)]
    lib.lucid_destroy.restype = None
    lib.lucid_destroy.argtypes = [vp]
    lib.lucid_render.restype = i32
    lib.lucid_render.argtypes = [vp, ctypes.POINTER(FrameInputs), ctypes.POINTER(FrameOutputs), vp]
    lib.lucid_render_ex.restype = i32
    lib.lucid_render_ex.argtypes = [vp, ctypes.POINTER(FrameInputs), ctypes.POINTER(FrameOutputs),
                                    ctypes.c_uint32, vp]
    lib.lucid_frame_recorded.restype = i32
    lib.lucid_frame_recorded.argtypes = [vp]
    lib.lucid_finish.restype = i32
    lib.lucid_finish.argtypes = [vp]
    lib.lucid_read_info.restype = i32
    lib.lucid_read_info.argtypes = [vp, vp, sz]
    lib.lucid_verify_info.restype = i32
    lib.lucid_verify_info.argtypes = [vp]
    lib.lucid_read_bin_lists.restype = i32
    lib.lucid_read_bin_lists.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(sz),
                                         ctypes.POINTER(sz)]
    lib.lucid_read_setup.restype = i32
    lib.lucid_read_setup.argtypes = [vp, vp, sz, vp, sz]
    lib.lucid_stage_times.restype = i32
    lib.lucid_stage_times.argtypes = [vp, vp, i32]
    lib.lucid_backward.restype = i32
    lib.lucid_backward.argtypes = [vp, ctypes.POINTER(FrameInputs), vp, vp, vp, vp]
    lib.lucid_backward_ex.restype = i32
    lib.lucid_backward_ex.argtypes = [vp, ctypes.POINTER(FrameInputs), vp, vp, vp,
                                      ctypes.c_uint32, vp]
    lib.lucid_set_tuning.restype = i32
    lib.lucid_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.lucid_get_tuning.restype = i32
    lib.lucid_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.lucid_autotune.restype = i32
    lib.lucid_autotune.argtypes = [vp, ctypes.POINTER(FrameInputs), ctypes.POINTER(FrameOutputs),
                                   vp, i32, ctypes.POINTER(Tuning)]
    lib.lucid_batch_create.restype = i32
    lib.lucid_batch_create.argtypes = [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)]
    lib.lucid_batch_destroy.restype = None
    lib.lucid_batch_destroy.argtypes = [vp]
    lib.lucid_batch_render.restype = i32
    lib.lucid_batch_render.argtypes = [vp, ctypes.POINTER(FrameInputs), ctypes.POINTER(FrameOutputs),
                                       vp]
    lib.lucid_batch_finish.restype = i32
    lib.lucid_batch_finish.argtypes = [vp]
    lib.lucid_render_noalloc.restype = i32
    lib.lucid_render_noalloc.argtypes = [i32, i32, ctypes.c_uint32, ctypes.POINTER(Tuning),
                                         ctypes.POINTER(FrameInputs), ctypes.POINTER(FrameOutputs),
                                         vp]
    lib.lucid_generate_quads.restype = i32
    lib.lucid_generate_quads.argtypes = [vp, i64, vp, i64, ctypes.c_float, vp, ctypes.POINTER(i64)]
    for name in ("lucid_bin_count", "lucid_bin_size", "lucid_block_size"):
        getattr(lib, name).restype = i32
        getattr(lib, name).argtypes = [vp]
    for name in ("lucid_max_visible_quads", "lucid_max_scene_quads"):
        getattr(lib, name).restype = i64
        getattr(lib, name).argtypes = [vp]
    _lib = lib
    return lib


def check(code: int, what: str = "") -> None:
    if code != LUCID_OK:
        raise LucidError(code, what)
