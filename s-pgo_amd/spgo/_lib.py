"""ctypes binding of ``libspgo.so`` (the C ABI declared in include/spgo.h).

The library is loaded from ``spgo/lib/libspgo.so``; there is no other
implementation behind these calls -- if the library is missing or fails to
load, every call raises.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspgo.so")

SPGO_MODE_REFERENCE = 0
SPGO_MODE_LM = 1
SPGO_ROBUST_CAUCHY = 0
SPGO_ROBUST_HUBER = 1
SPGO_ROBUST_NONE = 2
SPGO_PREC_FP64 = 0
SPGO_PREC_MIXED = 1
SPGO_FACTOR_AUTO = 0
SPGO_FACTOR_DENSE = 1
SPGO_FACTOR_ENVELOPE = 2
SPGO_FACTOR_ENVELOPE_SPLIT = 3
SPGO_FACTOR_PCG = 4
SPGO_FACTOR_CR = 5
SPGO_COORD_AUTO = 0
SPGO_COORD_INCREMENTAL = 1
SPGO_COORD_FRAMES = 2

ERRORS = {-1: "invalid argument", -2: "capacity exceeded", -3: "unsupported layout",
          -4: "device error", -5: "invalid state"}


class Options(ctypes.Structure):
    _fields_ = [("mode", c_int32), ("robust", c_int32), ("near_boost", c_int32), ("precision", c_int32),
                ("robust_param", c_double), ("eps", c_double), ("lambda0", c_double),
                ("lam_up", c_double), ("lam_down", c_double), ("lam_min", c_double), ("lam_max", c_double),
                ("alpha_min", c_double), ("alpha_max", c_double),
                ("reserved0", c_int32), ("use_graph", c_int32),
                ("chunk_landmarks", c_int32), ("group_chunks", c_int32),
                ("factorization", c_int32), ("pcg_max_iter", c_int32), ("coordinates", c_int32),
                ("reserved", c_int32 * 5),
                ("pcg_tol", c_double)]


class IterStats(ctypes.Structure):
    _fields_ = [("cost_candidate", c_double), ("cost", c_double), ("lambda_", c_double),
                ("accepted", c_int32), ("chol_ok", c_int32), ("pcg_iters", c_int32), ("pcg_fallback", c_int32)]


class LayoutInfo(ctypes.Structure):
    _fields_ = [("n_poses", c_int32), ("n_slots", c_int32), ("n_measurements", c_int32),
                ("n_landmarks", c_int32), ("n_groups", c_int32), ("n_chunks", c_int32),
                ("max_span", c_int32), ("n_anchors_used", c_int32), ("factorization", c_int32),
                ("profile_depth", c_int32), ("reduced_size", c_int32), ("coordinates", c_int32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_problems", c_int32), ("n_poses", c_int32), ("n_landmarks", c_int32),
                ("n_measurements", c_int32), ("n_groups", c_int32), ("max_depth", c_int32),
                ("max_span", c_int32), ("chunk_landmarks", c_int32), ("group_chunks", c_int32),
                ("reserved", c_int32 * 7)]


F32P = POINTER(c_float)
F64P = POINTER(c_double)
I32P = POINTER(c_int32)

# name -> (restype, argtypes); every symbol declared in include/spgo.h
SIGNATURES = {
    "spgo_default_options": (c_int32, [POINTER(Options)]),
    "spgo_create": (c_int32, [c_int32, c_int32, POINTER(Options), POINTER(c_void_p)]),
    "spgo_destroy": (c_int32, [c_void_p]),
    "spgo_set_options": (c_int32, [c_void_p, POINTER(Options)]),
    "spgo_get_options": (c_int32, [c_void_p, POINTER(Options)]),
    "spgo_load_calibration": (c_int32, [c_void_p, F32P, F32P]),
    "spgo_get_calibration": (c_int32, [c_void_p, F32P, F32P]),
    "spgo_write_observation": (c_int32, [c_void_p, c_int32, c_int32, c_int32, F32P, F32P]),
    "spgo_write_observations_bulk": (c_int32, [c_void_p, c_int32, I32P, I32P, I32P, F32P]),
    "spgo_get_observation": (c_int32, [c_void_p, c_int32, c_int32, F32P, F32P]),
    "spgo_step": (c_int32, [c_void_p, c_int32]),
    "spgo_get_incremental_pose": (c_int32, [c_void_p, c_int32, F32P]),
    "spgo_write_incremental_pose": (c_int32, [c_void_p, c_int32, F32P]),
    "spgo_load_inverse_depths": (c_int32, [c_void_p, F32P]),
    "spgo_get_inverse_depths": (c_int32, [c_void_p, F32P]),
    "spgo_init_inverse_depths_from_disparity": (c_int32, [c_void_p, I32P]),
    "spgo_set_step_size": (c_int32, [c_void_p, c_float]),
    "spgo_get_state_f64": (c_int32, [c_void_p, F64P, F64P]),
    "spgo_set_state_f64": (c_int32, [c_void_p, F64P, F64P]),
    "spgo_get_stats": (c_int32, [c_void_p, POINTER(IterStats), c_int32, I32P, F64P]),
    "spgo_get_lambda": (c_int32, [c_void_p, F64P]),
    "spgo_set_lambda": (c_int32, [c_void_p, c_double]),
    "spgo_linearize": (c_int32, [c_void_p, c_double, F64P, F64P, F64P]),
    "spgo_reduced_band": (c_int32, [c_void_p, c_double, c_int32, F64P, F64P, POINTER(c_double)]),
    "spgo_device_synchronize": (c_int32, []),
    "spgo_reduced_band_noalloc": (c_int32, [c_void_p, c_double, c_int32, F64P, F64P, POINTER(c_double)]),
    "spgo_problem_info": (c_int32, [c_void_p, POINTER(LayoutInfo)]),
    "spgo_last_step_ms": (c_int32, [c_void_p, F64P]),
    "spgo_profile_phases": (c_int32, [c_void_p, c_int32, F64P]),
    "spgo_last_error": (c_char_p, [c_void_p]),
    "spgo_set_preconditioner": (c_int32, [c_void_p, c_void_p]),
    "spgo_batch_create": (c_int32, [c_int32, POINTER(c_void_p), POINTER(Options), POINTER(c_void_p)]),
    "spgo_batch_destroy": (c_int32, [c_void_p]),
    "spgo_batch_set_options": (c_int32, [c_void_p, POINTER(Options)]),
    "spgo_batch_get_options": (c_int32, [c_void_p, POINTER(Options)]),
    "spgo_batch_reload": (c_int32, [c_void_p]),
    "spgo_batch_reset_states": (c_int32, [c_void_p]),
    "spgo_batch_step": (c_int32, [c_void_p, c_int32]),
    "spgo_batch_get_state_f64": (c_int32, [c_void_p, c_int32, F64P, F64P]),
    "spgo_batch_store": (c_int32, [c_void_p]),
    "spgo_batch_get_stats": (c_int32, [c_void_p, c_int32, POINTER(IterStats), c_int32, I32P, F64P]),
    "spgo_batch_last_step_ms": (c_int32, [c_void_p, F64P]),
    "spgo_batch_get_info": (c_int32, [c_void_p, POINTER(BatchInfo)]),
    "spgo_batch_profile_phases": (c_int32, [c_void_p, c_int32, F64P]),
    "spgo_batch_autotune": (c_int32, [c_void_p, c_int32, I32P, I32P, F64P]),
    "spgo_batch_last_error": (c_char_p, [c_void_p]),
    "spgo_lpcg_create": (c_int32, [c_int32, c_uint64, POINTER(c_void_p)]),
    "spgo_lpcg_destroy": (c_int32, [c_void_p]),
    "spgo_lpcg_n_params": (c_int32, [c_void_p, I32P]),
    "spgo_lpcg_get_params": (c_int32, [c_void_p, F32P]),
    "spgo_lpcg_set_params": (c_int32, [c_void_p, F32P]),
    "spgo_lpcg_load_systems": (c_int32, [c_void_p, c_int32, I32P, F64P, F64P]),
    "spgo_lpcg_loss_grad": (c_int32, [c_void_p, F64P, F32P]),
    "spgo_lpcg_train": (c_int32, [c_void_p, c_int32, c_float, F32P]),
    "spgo_lpcg_solve": (c_int32, [c_void_p, c_int32, c_double, c_int32, I32P, F64P, F64P]),
    "spgo_lpcg_last_ms": (c_int32, [c_void_p, F64P]),
    "spgo_lpcg_get_features": (c_int32, [c_void_p, F32P, F64P]),
    "spgo_lpcg_last_error": (c_char_p, [c_void_p]),
}

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libspgo.so once and attach the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libspgo.so not built at {path}: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# Every live handle owner (Solver, Batch, LearnedPCG), so that what is still
# open when the interpreter exits is destroyed by an atexit handler -- before
# module teardown and before the HIP runtime's own exit handlers -- instead of
# from __del__ during finalisation, when the order of teardown is not defined.
_live = weakref.WeakSet()
_CLOSE_ORDER = {"Batch": 0, "Solver": 1, "LearnedPCG": 2}   # a batch uses solvers, a solver a network


def register(obj) -> None:
    _live.add(obj)


@atexit.register
def close_all() -> None:
    """Destroy every handle still open, batches first, then solvers, then
    networks, and wait for the device."""
    objs = sorted(list(_live), key=lambda o: _CLOSE_ORDER.get(type(o).__name__, 1))
    for o in objs:
        try:
            o.close()
        except Exception:
            pass
    if _lib is not None and objs:
        try:
            _lib.spgo_device_synchronize()
        except Exception:
            pass


def check(rc: int, handle=None, batch: bool = False, kind: str = "") -> None:
    if rc != 0:
        lib = load()
        fn = lib.spgo_last_error
        if batch or kind == "batch":
            fn = lib.spgo_batch_last_error
        elif kind == "lpcg":
            fn = lib.spgo_lpcg_last_error
        msg = fn(handle)
        msg = msg.decode() if msg else ""
        raise RuntimeError(f"spgo error {rc} ({ERRORS.get(rc, '?')}): {msg}")


def default_options() -> Options:
    o = Options()
    check(load().spgo_default_options(ctypes.byref(o)))
    return o
