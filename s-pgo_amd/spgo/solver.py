"""Python host side of the LM stereo bundle-adjustment step.

``Solver`` owns one ``libspgo`` handle.  Its methods mirror the reference's
Cython ``CudaSolver`` (src/CythonAPI/Solver.pyx:16-107) and add bulk
ingestion, float64 state access, options and per-iteration statistics.
Arrays are copied in and out; the library owns all device memory.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from ._lib import BatchInfo, F32P, F64P, I32P, IterStats, LayoutInfo, Options


def _f32(a, n: Optional[int] = None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
    if n is not None and a.size < n:
        raise ValueError(f"expected at least {n} values, got {a.size}")
    return a


def _p32(a: np.ndarray):
    return a.ctypes.data_as(F32P)


def make_options(mode: str = "lm", robust: str = "cauchy", robust_param: Optional[float] = None,
                 precision: str = "fp64", near_boost: bool = True, **kw) -> Options:
    """Options with readable names; ``kw`` sets any other field directly."""
    o = _lib.default_options()
    o.mode = {"reference": _lib.SPGO_MODE_REFERENCE, "lm": _lib.SPGO_MODE_LM}[mode]
    o.robust = {"cauchy": _lib.SPGO_ROBUST_CAUCHY, "huber": _lib.SPGO_ROBUST_HUBER,
                "none": _lib.SPGO_ROBUST_NONE}[robust]
    if robust_param is not None:
        o.robust_param = robust_param
    elif robust == "huber":
        o.robust_param = 2.0
    o.precision = {"fp64": _lib.SPGO_PREC_FP64, "mixed": _lib.SPGO_PREC_MIXED}[precision]
    o.near_boost = 1 if near_boost else 0
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Solver:
    """One stereo bundle-adjustment problem on the GPU."""

    def __init__(self, N: int, M: int, options: Optional[Options] = None):
        self._lib = _lib.load()
        _lib.register(self)
        self.N = int(N)
        self.M = int(M)
        h = ctypes.c_void_p()
        rc = self._lib.spgo_create(self.N, self.M, ctypes.byref(options) if options is not None else None,
                                   ctypes.byref(h))
        _lib.check(rc, None)
        self._h = h

    # -- lifetime -----------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.spgo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        _lib.check(rc, self._h)

    # -- options ------------------------------------------------------------
    def set_options(self, options: Options) -> None:
        self._check(self._lib.spgo_set_options(self._h, ctypes.byref(options)))

    def get_options(self) -> Options:
        o = Options()
        self._check(self._lib.spgo_get_options(self._h, ctypes.byref(o)))
        return o

    # -- calibration ----------------------------------------------------------
    def load_calibration(self, intrinsics, T_r_to_l) -> None:
        i = _f32(intrinsics, 8)
        t = _f32(T_r_to_l, 16)
        self._check(self._lib.spgo_load_calibration(self._h, _p32(i), _p32(t)))

    def get_calibration(self):
        i = np.empty(8, np.float32)
        t = np.empty(16, np.float32)
        self._check(self._lib.spgo_get_calibration(self._h, _p32(i), _p32(t)))
        return i, t

    # -- measurements -------------------------------------------------------------
    def write_observation(self, anchor: int, target: int, feat: int, left, right) -> None:
        l3 = _f32(left, 3)
        r3 = _f32(right, 3)
        self._check(self._lib.spgo_write_observation(self._h, int(anchor), int(target), int(feat), _p32(l3),
                                                     _p32(r3)))

    def write_observations(self, anchor, target, feat, uvLR) -> None:
        a = np.ascontiguousarray(np.asarray(anchor, np.int32).reshape(-1))
        t = np.ascontiguousarray(np.asarray(target, np.int32).reshape(-1))
        f = np.ascontiguousarray(np.asarray(feat, np.int32).reshape(-1))
        uv = np.ascontiguousarray(np.asarray(uvLR, np.float32).reshape(-1, 4))
        if not (a.size == t.size == f.size == uv.shape[0]):
            raise ValueError("measurement arrays differ in length")
        self._check(self._lib.spgo_write_observations_bulk(self._h, a.size, a.ctypes.data_as(I32P),
                                                           t.ctypes.data_as(I32P), f.ctypes.data_as(I32P),
                                                           uv.ctypes.data_as(F32P)))

    def get_observation(self, frame: int, feat: int):
        l3 = np.empty(3, np.float32)
        r3 = np.empty(3, np.float32)
        self._check(self._lib.spgo_get_observation(self._h, int(frame), int(feat), _p32(l3), _p32(r3)))
        return l3, r3

    # -- state ----------------------------------------------------------------
    def step(self, iterations: int) -> None:
        self._check(self._lib.spgo_step(self._h, int(iterations)))

    def get_incremental_pose(self, k: int) -> np.ndarray:
        T = np.empty(16, np.float32)
        self._check(self._lib.spgo_get_incremental_pose(self._h, int(k), _p32(T)))
        return T.reshape(4, 4)

    def write_incremental_pose(self, k: int, T) -> None:
        t = _f32(T, 16)
        self._check(self._lib.spgo_write_incremental_pose(self._h, int(k), _p32(t)))

    def load_inverse_depths(self, alpha) -> None:
        a = _f32(alpha, self.N * self.M)
        self._check(self._lib.spgo_load_inverse_depths(self._h, _p32(a)))

    def get_inverse_depths(self) -> np.ndarray:
        a = np.empty(self.N * self.M, np.float32)
        self._check(self._lib.spgo_get_inverse_depths(self._h, _p32(a)))
        return a

    def init_inverse_depths_from_disparity(self) -> int:
        """Inverse depths from the anchor frame's stereo pair (least-squares
        triangulation along the left ray, Optimizer.py:87-114); returns the
        number of landmarks set."""
        n = ctypes.c_int32(0)
        self._check(self._lib.spgo_init_inverse_depths_from_disparity(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_step_size(self, s: float) -> None:
        self._check(self._lib.spgo_set_step_size(self._h, float(s)))

    def get_state(self):
        """(poses (N,4,4) float64, inverse depths (N*M,) float64)."""
        X = np.empty((self.N, 4, 4), np.float64)
        al = np.empty(self.N * self.M, np.float64)
        self._check(self._lib.spgo_get_state_f64(self._h, X.ctypes.data_as(F64P), al.ctypes.data_as(F64P)))
        return X, al

    def set_state(self, poses=None, alpha=None) -> None:
        Xp = al = None
        if poses is not None:
            X = np.ascontiguousarray(np.asarray(poses, np.float64).reshape(self.N, 4, 4))
            Xp = X.ctypes.data_as(F64P)
        if alpha is not None:
            a = np.ascontiguousarray(np.asarray(alpha, np.float64).reshape(-1))
            if a.size != self.N * self.M:
                raise ValueError("alpha must have N*M values")
            al = a.ctypes.data_as(F64P)
        self._check(self._lib.spgo_set_state_f64(self._h, Xp, al))

    # -- diagnostics --------------------------------------------------------------
    def stats(self):
        """(initial cost, list of per-iteration dicts) of the last step."""
        n = ctypes.c_int32()
        c0 = ctypes.c_double()
        self._check(self._lib.spgo_get_stats(self._h, None, 0, ctypes.byref(n), ctypes.byref(c0)))
        buf = (IterStats * max(n.value, 1))()
        self._check(self._lib.spgo_get_stats(self._h, buf, n.value, ctypes.byref(n), ctypes.byref(c0)))
        recs = [dict(cost_candidate=buf[i].cost_candidate, cost=buf[i].cost, lam=buf[i].lambda_,
                     accepted=bool(buf[i].accepted), chol_ok=bool(buf[i].chol_ok),
                     pcg_iters=int(buf[i].pcg_iters), pcg_fallback=bool(buf[i].pcg_fallback)) for i in range(n.value)]
        return c0.value, recs

    def get_lambda(self) -> float:
        v = ctypes.c_double()
        self._check(self._lib.spgo_get_lambda(self._h, ctypes.byref(v)))
        return v.value

    def set_lambda(self, lam: float) -> None:
        self._check(self._lib.spgo_set_lambda(self._h, float(lam)))

    def linearize(self, lam: float = 0.0):
        """(H (6N,6N), g (6N,), cost) of the reduced pose system at the current state."""
        P = 6 * self.N
        H = np.empty((P, P), np.float64)
        g = np.empty(P, np.float64)
        c = ctypes.c_double()
        self._check(self._lib.spgo_linearize(self._h, float(lam), H.ctypes.data_as(F64P), g.ctypes.data_as(F64P),
                                             ctypes.byref(c)))
        return H, g, c.value

    def reduced_band(self, W: int, lam: float = 0.0):
        """(Hband (N, W+1, 6, 6), g (N, 6), cost): the block band of the reduced
        system the step factors, in the solver's own coordinates, at the
        current state with damping ``lam``."""
        Hb = np.empty((self.N, W + 1, 6, 6), np.float64)
        g = np.empty((self.N, 6), np.float64)
        c = ctypes.c_double()
        self._check(self._lib.spgo_reduced_band(self._h, float(lam), int(W), Hb.ctypes.data_as(F64P),
                                                g.ctypes.data_as(F64P), ctypes.byref(c)))
        return Hb, g, c.value

    def reduced_band_noalloc(self, W: int, lam: float = 0.0):
        """``reduced_band`` without device allocation (the band is staged in the
        library's scratch array); needs a solver that has stepped once."""
        Hb = np.empty((self.N, W + 1, 6, 6), np.float64)
        g = np.empty((self.N, 6), np.float64)
        c = ctypes.c_double()
        self._check(self._lib.spgo_reduced_band_noalloc(self._h, float(lam), int(W), Hb.ctypes.data_as(F64P),
                                                        g.ctypes.data_as(F64P), ctypes.byref(c)))
        return Hb, g, c.value

    def layout_info(self) -> dict:
        info = LayoutInfo()
        self._check(self._lib.spgo_problem_info(self._h, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in LayoutInfo._fields_}

    PHASES = ("form_system", "cholesky", "solve", "pose_update", "assemble", "accept")

    def profile_phases(self, iterations: int) -> dict:
        """Mean device ms per LM iteration of each phase, from HIP events on the
        solver's stream over ``iterations`` directly launched iterations."""
        ms = np.zeros(len(self.PHASES), np.float64)
        self._check(self._lib.spgo_profile_phases(self._h, int(iterations), ms.ctypes.data_as(F64P)))
        return dict(zip(self.PHASES, ms.tolist()))

    def last_step_ms(self) -> float:
        v = ctypes.c_double()
        self._check(self._lib.spgo_last_step_ms(self._h, ctypes.byref(v)))
        return v.value

    def set_preconditioner(self, net) -> None:
        """The learned preconditioner (an ``lpcg.LearnedPCG``) of the PCG
        factorisation, or None for block Jacobi.  The solver keeps a reference."""
        self._check(self._lib.spgo_set_preconditioner(self._h, net._h if net is not None else None))
        self._net = net


def solver_from_problem(p, options: Optional[Options] = None) -> Solver:
    """Solver loaded with a ``problems.Problem``: calibration, measurements in
    bulk, initial poses and inverse depths."""
    s = Solver(p.n_poses, p.m_per_frame, options)
    s.load_calibration(p.intrinsics, p.T_r_to_l)
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    for k in range(p.n_poses):
        s.write_incremental_pose(k, p.X0[k].reshape(16))
    s.load_inverse_depths(p.alpha0)
    return s


class Batch:
    """Independent problems stepped together: one launch per LM phase for all.

    The problems are ``Solver`` objects loaded through their own API (they must
    share one stereo calibration); the batch holds references to them.  The
    first ``step`` reads their measurements and states; later steps continue
    from the batch's device state until ``reload``.  ``store`` writes the
    batch's states back into the solvers.
    """

    def __init__(self, solvers, options: Optional[Options] = None):
        self._lib = _lib.load()
        _lib.register(self)
        self.solvers = list(solvers)
        hs = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        rc = self._lib.spgo_batch_create(len(self.solvers), hs, ctypes.byref(options) if options is not None
                                         else None, ctypes.byref(h))
        _lib.check(rc, None, batch=True)
        self._h = h

    def _check(self, rc: int) -> None:
        _lib.check(rc, self._h, batch=True)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.spgo_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self) -> int:
        return len(self.solvers)

    def set_options(self, o: Options) -> None:
        self._check(self._lib.spgo_batch_set_options(self._h, ctypes.byref(o)))

    def get_options(self) -> Options:
        o = Options()
        self._check(self._lib.spgo_batch_get_options(self._h, ctypes.byref(o)))
        return o

    def reload(self) -> None:
        self._check(self._lib.spgo_batch_reload(self._h))

    def reset_states(self) -> None:
        """Restart from the solvers' states (measurements kept); damping reset."""
        self._check(self._lib.spgo_batch_reset_states(self._h))

    def step(self, iterations: int) -> None:
        self._check(self._lib.spgo_batch_step(self._h, int(iterations)))

    def get_state(self, p: int):
        """(poses (N,4,4) float64, inverse depths (N*M,) float64) of problem p."""
        s = self.solvers[p]
        X = np.empty((s.N, 4, 4), np.float64)
        al = np.empty(s.N * s.M, np.float64)
        self._check(self._lib.spgo_batch_get_state_f64(self._h, int(p), X.ctypes.data_as(F64P),
                                                       al.ctypes.data_as(F64P)))
        return X, al

    def store(self) -> None:
        self._check(self._lib.spgo_batch_store(self._h))

    def stats(self, p: int):
        """(initial cost, per-iteration dicts) of problem p in the last step."""
        n = ctypes.c_int32()
        c0 = ctypes.c_double()
        self._check(self._lib.spgo_batch_get_stats(self._h, int(p), None, 0, ctypes.byref(n), ctypes.byref(c0)))
        buf = (IterStats * max(n.value, 1))()
        self._check(self._lib.spgo_batch_get_stats(self._h, int(p), buf, n.value, ctypes.byref(n),
                                                   ctypes.byref(c0)))
        recs = [dict(cost_candidate=buf[i].cost_candidate, cost=buf[i].cost, lam=buf[i].lambda_,
                     accepted=bool(buf[i].accepted), chol_ok=bool(buf[i].chol_ok),
                     pcg_iters=int(buf[i].pcg_iters), pcg_fallback=bool(buf[i].pcg_fallback)) for i in range(n.value)]
        return c0.value, recs

    def last_step_ms(self) -> float:
        v = ctypes.c_double()
        self._check(self._lib.spgo_batch_last_step_ms(self._h, ctypes.byref(v)))
        return v.value

    def info(self) -> dict:
        i = BatchInfo()
        self._check(self._lib.spgo_batch_get_info(self._h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_ if k != "reserved"}

    def profile_phases(self, iterations: int) -> dict:
        """Mean device ms per LM iteration of each phase over the batch (the
        envelope kernel's solves are inside "cholesky")."""
        ms = np.zeros(len(Solver.PHASES), np.float64)
        self._check(self._lib.spgo_batch_profile_phases(self._h, int(iterations), ms.ctypes.data_as(F64P)))
        d = dict(zip(Solver.PHASES, ms.tolist()))
        d.pop("solve")
        return d

    def autotune(self, iterations: int = 3) -> dict:
        """Time every candidate tile shape and keep the fastest (see
        spgo_batch_autotune); returns the chosen shape and its device time."""
        cl = ctypes.c_int32()
        gc = ctypes.c_int32()
        ms = ctypes.c_double()
        self._check(self._lib.spgo_batch_autotune(self._h, int(iterations), ctypes.byref(cl), ctypes.byref(gc),
                                                  ctypes.byref(ms)))
        return {"chunk_landmarks": cl.value, "group_chunks": gc.value, "ms": ms.value}

