"""Learned preconditioner for CG on the reduced pose system (host side).

``LearnedPCG`` owns one ``spgo_lpcg`` handle of libspgo.so: the network
parameters with their Adam state and a loaded set of block-banded systems.
Systems come from the solver itself: ``Solver.reduced_band`` returns the
blocks H_{i,i-d}, d <= W, of the reduced system the step factors and its
right-hand side g at the solver's current state, in the solver's own
coordinates (``band_system`` does the same for a dense H given on the host).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import F32P, F64P, I32P

WMAX = 16


def band_width(H: np.ndarray) -> int:
    """Largest block offset d with a nonzero block H_{i,i-d}."""
    n = H.shape[0] // 6
    B = np.abs(H).reshape(n, 6, n, 6).max(axis=(1, 3))
    i, j = np.nonzero(B)
    return int(np.max(i - j)) if i.size else 0


def band_system(H: np.ndarray, g: np.ndarray, W: int):
    """(Hband (n, W+1, 6, 6), g (n, 6)) of a dense reduced system whose blocks
    beyond offset W are zero."""
    n = H.shape[0] // 6
    if band_width(H) > W:
        raise ValueError(f"system has blocks beyond offset {W}")
    Hb = np.zeros((n, W + 1, 6, 6))
    for d in range(min(W, n - 1) + 1):
        for i in range(d, n):
            Hb[i, d] = H[6 * i:6 * i + 6, 6 * (i - d):6 * (i - d) + 6]
    return Hb, np.ascontiguousarray(g.reshape(n, 6), dtype=np.float64)


class LearnedPCG:
    """Network, training and PCG on block-banded reduced systems (GPU)."""

    def __init__(self, band_blocks: int, seed: int = 0):
        self._lib = _lib.load()
        _lib.register(self)
        self.W = int(band_blocks)
        h = ctypes.c_void_p()
        _lib.check(self._lib.spgo_lpcg_create(self.W, int(seed), ctypes.byref(h)), None, kind="lpcg")
        self._h = h
        n = ctypes.c_int32()
        self._check(self._lib.spgo_lpcg_n_params(self._h, ctypes.byref(n)))
        self.n_params = n.value
        self.n_nodes: list = []

    def _check(self, rc: int) -> None:
        _lib.check(rc, self._h, kind="lpcg")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.spgo_lpcg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self._check(self._lib.spgo_lpcg_get_params(self._h, p.ctypes.data_as(F32P)))
        return p

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1))
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        self._check(self._lib.spgo_lpcg_set_params(self._h, p.ctypes.data_as(F32P)))

    def load(self, systems: Sequence) -> None:
        """systems: (Hband (n, W+1, 6, 6), g (n, 6)) pairs."""
        nn = np.array([s[0].shape[0] for s in systems], np.int32)
        for Hb, g in systems:
            if Hb.shape[1:] != (self.W + 1, 6, 6) or g.shape != (Hb.shape[0], 6):
                raise ValueError("system shapes do not match the band width")
        Hb = np.ascontiguousarray(np.concatenate([s[0] for s in systems]), dtype=np.float64)
        g = np.ascontiguousarray(np.concatenate([s[1] for s in systems]), dtype=np.float64)
        self._check(self._lib.spgo_lpcg_load_systems(self._h, len(systems), nn.ctypes.data_as(I32P),
                                                     Hb.ctypes.data_as(F64P), g.ctypes.data_as(F64P)))
        self.n_nodes = nn.tolist()

    def loss_grad(self):
        loss = ctypes.c_double()
        grad = np.empty(self.n_params, np.float32)
        self._check(self._lib.spgo_lpcg_loss_grad(self._h, ctypes.byref(loss), grad.ctypes.data_as(F32P)))
        return loss.value, grad

    def train(self, steps: int, lr: float = 1e-3) -> np.ndarray:
        losses = np.empty(max(int(steps), 1), np.float32)
        self._check(self._lib.spgo_lpcg_train(self._h, int(steps), float(lr), losses.ctypes.data_as(F32P)))
        return losses[:steps]

    def solve(self, learned: bool = True, tol: float = 1e-8, max_iter: int = 5000):
        """(iterations per system, relative residual per system, x (T, 6))."""
        S = len(self.n_nodes)
        it = np.empty(S, np.int32)
        rel = np.empty(S, np.float64)
        x = np.empty((sum(self.n_nodes), 6), np.float64)
        self._check(self._lib.spgo_lpcg_solve(self._h, 1 if learned else 0, float(tol), int(max_iter),
                                              it.ctypes.data_as(I32P), rel.ctypes.data_as(F64P),
                                              x.ctypes.data_as(F64P)))
        return it, rel, x

    def last_ms(self):
        ms = np.zeros(2, np.float64)
        self._check(self._lib.spgo_lpcg_last_ms(self._h, ms.ctypes.data_as(F64P)))
        return {"preconditioner": ms[0], "pcg": ms[1]}

    def features(self):
        T = sum(self.n_nodes)
        X = np.empty((T, 40), np.float32)
        C = np.empty((T, 6, 6), np.float64)
        self._check(self._lib.spgo_lpcg_get_features(self._h, X.ctypes.data_as(F32P), C.ctypes.data_as(F64P)))
        return X, C


def solver_systems(problem_list, W: int, states_per_problem: int = 2, lam: float = 1e-4, iters_between: int = 2,
                   options=None, restarts: int = 1, perturb=(0.01, 0.05), seed: int = 0) -> list:
    """Reduced systems generated by the solver, in its own coordinates (the
    block band of offsets <= W that the CG preconditioner sees, as
    ``Solver.reduced_band``).  For each problem: ``restarts`` starting states
    (the first the problem's initial state, the others its true state with
    every pose perturbed by a twist of RMS size drawn in ``perturb`` rad / m and
    inverse depths by up to 10 %), from each the system at the start and after
    every ``iters_between`` further LM iterations, ``states_per_problem``
    systems per start."""
    from .solver import make_options, solver_from_problem
    from .problems import _exp_se3
    rng = np.random.default_rng(seed)
    out = []
    for p in problem_list:
        s = solver_from_problem(p, options if options is not None else make_options(mode="lm"))
        used = np.unique(p.meas_l)
        for r in range(restarts):
            if r:
                sc = rng.uniform(perturb[0], perturb[1])
                xi = rng.normal(0.0, sc / np.sqrt(3.0), size=(p.n_poses, 6))
                X = p.X_true @ _exp_se3(xi)
                al = p.alpha_true.copy()
                al[used] *= 1.0 + rng.uniform(-0.1, 0.1, size=used.size)
                s.set_state(X, al)
            for k in range(states_per_problem):
                if k:
                    s.step(iters_between)
                Hb, g, _ = s.reduced_band(W, lam)
                out.append((Hb, g))
        s.close()
    return out
