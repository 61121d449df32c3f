"""ctypes binding of libspme_hip.so (the C ABI of include/spme.h and
include/spme_ext.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
if the shared library has not been built (spme_amd/build.py).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspme_hip.so")

SPME_FP64, SPME_FP32, SPME_FP64_PLAIN, SPME_FP32_SUM64 = 0, 1, 2, 3
SPME_RECIP, SPME_REAL, SPME_SELF, SPME_FULL = 1, 2, 4, 7

# Every symbol the public headers declare.
EXPORTED = [
    "cutoff2xi", "create_spme_engine", "destroy_spme_engine", "compute_spme",
    "spme_last_error", "spme_engine_create", "spme_engine_destroy",
    "spme_engine_set_positions", "spme_engine_apply", "spme_engine_info",
    "spme_engine_debug_stage", "spme_engine_set_timing", "spme_engine_get_timing",
    "spme_mlp_nparam", "spme_mlp_forward", "spme_mlp_backward", "spme_mlp_adam",
    "spme_engine_set_tiles", "spme_engine_get_tiles", "spme_engine_autotune",
    "spme_group_create", "spme_group_destroy", "spme_group_set_tiles",
    "spme_group_get_tiles", "spme_group_autotune", "spme_group_apply",
    "spme_block_gram_worksize", "spme_block_gram", "spme_block_update",
    "spme_block_gram_noalloc_max", "spme_block_gram_noalloc", "spme_engine_set_radii",
    "spme_stream_probe",
]

TILE_FIELDS = ["spread_cc", "spread_zc", "interp_zc", "spread_persist",
               "interp_persist", "real_persist"]

STAGES = ["permute", "spread", "fft_x_fwd", "fft_y_fwd", "fft_z_infl",
          "fft_y_inv", "fft_x_inv", "interp", "permute_real", "real"]


class SpmeParams(ctypes.Structure):
    _fields_ = [("xi", ctypes.c_double), ("dim", ctypes.c_int),
                ("porder", ctypes.c_int), ("np", ctypes.c_int),
                ("box_size", ctypes.c_double), ("rc", ctypes.c_double),
                ("precision", ctypes.c_int), ("max_rhs", ctypes.c_int),
                ("radii", ctypes.c_int)]


class SpmeInfo(ctypes.Structure):
    _fields_ = [("K", ctypes.c_int), ("P", ctypes.c_int), ("N", ctypes.c_int),
                ("precision", ctypes.c_int), ("max_rhs", ctypes.c_int),
                ("RS", ctypes.c_int), ("comp_stride", ctypes.c_long),
                ("L", ctypes.c_double), ("xi", ctypes.c_double),
                ("rc", ctypes.c_double), ("mesh_bytes", ctypes.c_long),
                ("stream", ctypes.c_void_p), ("real_entries", ctypes.c_long)]


class SpmeTiles(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in TILE_FIELDS]

    def as_dict(self):
        return {n: getattr(self, n) for n in TILE_FIELDS}


class SpmeError(RuntimeError):
    pass


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    # torch bundles its own HIP runtime with the same soname
    # (libamdhip64.so.7) as the one this library links; the process keeps
    # whichever is loaded first.  Load torch's first so that one runtime
    # serves both and torch's streams and buffers are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise SpmeError(f"{LIB_PATH} is missing: build it with "
                        "`python -m spme_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    vp, d, i, l = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_long
    dp = ctypes.POINTER(ctypes.c_double)
    L.cutoff2xi.argtypes = [d, d, d, dp]
    L.create_spme_engine.argtypes = [d, i, i, i, d, i, ctypes.POINTER(vp)]
    L.destroy_spme_engine.argtypes = [vp]
    L.compute_spme.argtypes = [vp, vp, i, d, vp, i, d, vp, i]
    L.spme_last_error.restype = ctypes.c_char_p
    L.spme_engine_create.argtypes = [ctypes.POINTER(SpmeParams), ctypes.POINTER(vp)]
    L.spme_engine_destroy.argtypes = [vp]
    L.spme_engine_set_positions.argtypes = [vp, vp, i, vp]
    L.spme_engine_set_radii.argtypes = [vp, vp, i, vp]
    L.spme_engine_apply.argtypes = [vp, i, d, vp, l, d, vp, l, i, vp]
    L.spme_engine_info.argtypes = [vp, ctypes.POINTER(SpmeInfo)]
    L.spme_engine_debug_stage.argtypes = [vp, i, i, vp, l, vp]
    L.spme_engine_set_timing.argtypes = [vp, i]
    L.spme_engine_get_timing.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int)]
    f = ctypes.c_float
    L.spme_mlp_nparam.restype = i
    L.spme_mlp_forward.argtypes = [vp, vp, i, vp, vp, vp, vp]
    L.spme_mlp_backward.argtypes = [vp, vp, vp, vp, vp, i, vp, vp, vp, vp, vp]
    L.spme_mlp_adam.argtypes = [vp, vp, vp, vp, f, f, f, f, i, vp]
    tp = ctypes.POINTER(SpmeTiles)
    L.spme_engine_set_tiles.argtypes = [vp, tp]
    L.spme_engine_get_tiles.argtypes = [vp, i, tp]
    L.spme_engine_autotune.argtypes = [vp, i, i, tp, vp]
    L.spme_group_create.argtypes = [i, ctypes.POINTER(vp), i, ctypes.POINTER(vp)]
    L.spme_group_destroy.argtypes = [vp]
    L.spme_group_set_tiles.argtypes = [vp, tp]
    L.spme_group_get_tiles.argtypes = [vp, tp]
    L.spme_group_autotune.argtypes = [vp, i, i, tp, vp]
    L.spme_group_apply.argtypes = [vp, d, ctypes.POINTER(vp), ctypes.POINTER(l), d,
                                   ctypes.POINTER(vp), ctypes.POINTER(l), i, i, vp]
    L.spme_block_gram_worksize.argtypes = [l, i, i]
    L.spme_block_gram_worksize.restype = l
    L.spme_block_gram.argtypes = [i, l, i, i, vp, l, vp, l, vp, vp, l, vp]
    L.spme_block_update.argtypes = [i, l, i, i, vp, l, vp, d, vp, l, vp]
    L.spme_block_gram_noalloc_max.argtypes = []
    L.spme_block_gram_noalloc_max.restype = l
    L.spme_block_gram_noalloc.argtypes = [i, l, i, i, vp, l, vp, l, vp, vp]
    L.spme_stream_probe.argtypes = [i, i, vp, vp, l, ctypes.c_float, i, i, vp,
                                    ctypes.POINTER(ctypes.c_float)]
    _lib = L
    return L


def last_error() -> str:
    return load().spme_last_error().decode()


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise SpmeError(f"{what}: {last_error()} (code {rc})")
