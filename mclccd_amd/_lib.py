"""ctypes binding of the native library (include/mclg.h -> mclccd_amd/libmclg.so).

There is no fallback: if the library is missing or cannot be loaded, importing
this module raises.  Build it with ``python -m mclccd_amd.build`` (or
``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmclg.so")

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_STATE = -4
ERR_UNSUPPORTED = -5
ERR_CAPACITY = -6
NOALLOC_MAX_HITS = 1 << 20

KEEP_CANDIDATES = 1
KEEP_LEAF_PAIRS = 2
CHECK_ORDER = 4
ORIENT_INDEX = 0
ORIENT_KDBVH = 1
PRECISION_F64 = 0
PRECISION_F32 = 1
NUM_TIMINGS = 13
TIMING_NAMES = ("update_pack", "update_reptri", "update_leaf_boxes", "update_morton", "update_sort",
                "update_lbvh", "traverse_pack", "traverse", "expand", "accd", "adjoint", "accd_vf", "accd_ee")


class Options(C.Structure):
    _fields_ = [("box_eta", C.c_double), ("vf_ccd_eta", C.c_double), ("ee_ccd_eta", C.c_double),
                ("reptri", C.c_int), ("threaded", C.c_int), ("discrete", C.c_int), ("continuous", C.c_int),
                ("vf_one_sided", C.c_int), ("ee_robust", C.c_int)]


class Counts(C.Structure):
    _fields_ = [("leaf_pairs", C.c_int64), ("candidates", C.c_int64), ("hits", C.c_int64),
                ("discrete", C.c_int64), ("dist_evals", C.c_int64),
                ("vf_candidates", C.c_int64), ("ee_dist_evals", C.c_int64), ("order_dependent", C.c_int64),
                ("traverse_steps", C.c_int64), ("suspended", C.c_int64)]


class DeviceResults(C.Structure):
    _fields_ = [("hit_sten4", C.c_void_p), ("hit_is_vf", C.c_void_p), ("hit_toi", C.c_void_p),
                ("discrete2", C.c_void_p), ("vf_sten4", C.c_void_p), ("vf_toi", C.c_void_p), ("n_vf", C.c_int64),
                ("ee_sten4", C.c_void_p), ("ee_toi", C.c_void_p), ("n_ee", C.c_int64), ("leaf_pairs2", C.c_void_p)]


class MclgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mclg error {code}: {msg}")
        self.code = code


def _share_torch_hip_runtime():
    """PyTorch-ROCm ships its own HIP runtime.  Two HIP runtimes in one process
    cannot both use the GPU (the first to initialise holds it), so when torch
    is importable it is imported before the library: libmclg.so's libamdhip64
    dependency then resolves (same SONAME) to the runtime torch already loaded,
    whichever of torch and this package the caller uses first."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m mclccd_amd.build`")
    _share_torch_hip_runtime()
    L = C.CDLL(LIB_PATH)
    P, I, U, I64 = C.c_void_p, C.c_int, C.c_uint, C.c_int64
    sig = {
        "mclg_default_options": (None, [P]),
        "mclg_create": (I, [I, P]),
        "mclg_destroy": (None, [P]),
        "mclg_set_options": (I, [P, P]),
        "mclg_get_options": (I, [P, P]),
        "mclg_last_error": (C.c_char_p, [P]),
        "mclg_set_stream": (I, [P, P]),
        "mclg_update": (I, [P, P, P, I, P, I, I, P]),
        "mclg_update_device": (I, [P, P, P, I, P, I, I, P]),
        "mclg_update_refit": (I, [P, P, P, I, P, I, I, P]),
        "mclg_update_refit_device": (I, [P, P, P, I, P, I, I, P]),
        "mclg_traverse": (I, [P, P, P, P, U]),
        "mclg_traverse_device": (I, [P, P, P, P, U]),
        "mclg_get_counts": (I, [P, P]),
        "mclg_copy_hits": (I, [P, P, P, P, I64]),
        "mclg_copy_discrete": (I, [P, P, I64]),
        "mclg_copy_candidates": (I, [P, P, P, P, I64]),
        "mclg_copy_leaf_pairs": (I, [P, P, I64]),
        "mclg_copy_order_dependent": (I, [P, P, I64]),
        "mclg_get_device_results": (I, [P, P]),
        "mclg_copy_leaves": (I, [P, P, P, P, I64]),
        "mclg_copy_tree": (I, [P, P, P, P, P, I64]),
        "mclg_toi_grad": (I, [P, P, P, P, I]),
        "mclg_toi_grad_noalloc": (I, [P, P, P, P]),
        "mclg_batch_create": (I, [I, P]),
        "mclg_batch_destroy": (None, [P]),
        "mclg_batch_last_error": (C.c_char_p, [P]),
        "mclg_batch_set_scenes": (I, [P, I, P, P, P, P, I]),
        "mclg_batch_run": (I, [P, P, P, I]),
        "mclg_batch_last_ms": (C.c_double, [P]),
        "mclg_batch_get_counts": (I, [P, P]),
        "mclg_batch_scene_counts": (I, [P, P, P]),
        "mclg_batch_copy_hits": (I, [P, P, P, P, P, I64]),
        "mclg_batch_copy_discrete": (I, [P, P, P, I64]),
        "mclg_batch_set_params": (I, [P, I, I, I, I]),
        "mclg_batch_set_active": (I, [P, P, I]),
        "mclg_batch_get_params": (I, [P, P]),
        "mclg_batch_autotune": (I, [P, C.c_char_p, I]),
        "mclg_set_rank": (I, [P, P, I]),
        "mclg_set_orientation": (I, [P, I]),
        "mclg_get_orientation": (I, [P, P]),
        "mclg_set_precision": (I, [P, I]),
        "mclg_get_precision": (I, [P, P]),
        "mclg_kdbvh_order": (I, [I, P, I, P, P]),
        "mclg_copy_events": (I, [P, I, P, P, I64, P]),
        "mclg_get_timings": (I, [P, P, I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()
