"""FP64 CPU oracle for the S-PGO Levenberg-Marquardt stereo bundle-adjustment step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product path (``spgo`` / ``build_ip``) never
imports, links or falls back to it.

It restates, in float64 numpy, the per-iteration semantics of the reference GPU
step ``updateState`` (src/Optimization/Solver_cuda.cu:557-795), built the way
the reference builds it: an explicit weighted pose Jacobian ``J_T`` with one
2x6 block per (measurement, camera, state i in [anchor, target)), an inverse
depth Jacobian ``J_alpha``, the products ``A = J_T^T J_T``, ``B = J_T^T J_alpha``,
``C = diag(J_alpha^T J_alpha)``, the landmark Schur complement, a Cholesky solve,
back-substitution and the SE(3) retraction.  Sparse matrices replace the
reference's dense buffers; the arithmetic is the same.

Two modes:
* ``reference``: lambda = 0 and every step accepted -- the reference GPU path
  (no damping, no cost, Solver_cuda.cu:594-782).
* ``lm``: lambda added to the diagonal of the reduced pose system (the unused
  ``add_damping_to_schur``, Solver_cuda.cu:348-359), the robust cost evaluated
  at the candidate, accept if it decreased (lambda *= lam_down) else reject
  (lambda *= lam_up).

Pinning (tests/test_oracle_*.py): closed-form Lie functions against
``scipy.spatial.transform`` and ``scipy.linalg.expm``; Jacobians against
central finite differences of the oracle's own residual; the gradient against
a finite difference of the cost; the Schur solve against a dense solve of the
full normal equations; the reference's deterministic fixture
(PythonUtils/utils.py:83-129, PythonUtils/Simulator.py:94-177) as a
known-answer test; convergence to ground truth on noise-free problems.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

NUMERICAL_TOLERANCE = 1e-4            # include/lie.h:4
CAUCHY_C2 = 9.0                       # include/LMutils.h:48
EPS_C = 1e-4                          # include/LMutils.h:50
ALPHA_MIN = 1.0 / 30.0                # include/LMutils.h:54
ALPHA_MAX = 1.0 / 0.2                 # include/LMutils.h:55


# ---------------------------------------------------------------------------
# Lie group math (src/LieUtils/{SO3,SE3,utils}.cu), vectorised over (..., 3)
# ---------------------------------------------------------------------------
def skew(v: np.ndarray) -> np.ndarray:
    """[v]x  (src/LieUtils/utils.cu:21-36)."""
    v = np.asarray(v, dtype=np.float64)
    S = np.zeros(v.shape[:-1] + (3, 3))
    S[..., 0, 1] = -v[..., 2]
    S[..., 0, 2] = v[..., 1]
    S[..., 1, 0] = v[..., 2]
    S[..., 1, 2] = -v[..., 0]
    S[..., 2, 0] = -v[..., 1]
    S[..., 2, 1] = v[..., 0]
    return S


def exp_so3(psi: np.ndarray) -> np.ndarray:
    """Rodrigues; below |psi| = 1e-4 the second-order form I + S + S^2/2
    (src/LieUtils/SO3.cu:5-32)."""
    psi = np.asarray(psi, dtype=np.float64)
    S = skew(psi)
    S2 = S @ S
    th = np.linalg.norm(psi, axis=-1)[..., None, None]
    big = th > NUMERICAL_TOLERANCE
    ths = np.where(big, th, 1.0)
    c1 = np.where(big, np.sin(ths) / ths, 1.0)
    c2 = np.where(big, (1.0 - np.cos(ths)) / (ths * ths), 0.5)
    return np.eye(3) + c1 * S + c2 * S2


def left_jacobian_so3(psi: np.ndarray) -> np.ndarray:
    """Jl(psi) = I + (1-cos)/th^2 S + (th-sin)/th^3 S^2; below 1e-4 the
    coefficients 1/2 and 1/6 (src/LieUtils/SO3.cu:82-107)."""
    psi = np.asarray(psi, dtype=np.float64)
    S = skew(psi)
    S2 = S @ S
    th = np.linalg.norm(psi, axis=-1)[..., None, None]
    big = th > NUMERICAL_TOLERANCE
    ths = np.where(big, th, 1.0)
    c1 = np.where(big, (1.0 - np.cos(ths)) / (ths * ths), 0.5)
    c2 = np.where(big, (ths - np.sin(ths)) / (ths ** 3), 1.0 / 6.0)
    return np.eye(3) + c1 * S + c2 * S2


def exp_se3(xi: np.ndarray) -> np.ndarray:
    """Exp(psi, rho) = [Exp_SO3(psi), Jl(psi) rho] (src/LieUtils/SE3.cu:3-38)."""
    xi = np.asarray(xi, dtype=np.float64)
    T = np.zeros(xi.shape[:-1] + (4, 4))
    T[..., :3, :3] = exp_so3(xi[..., :3])
    T[..., :3, 3] = (left_jacobian_so3(xi[..., :3]) @ xi[..., 3:, None])[..., 0]
    T[..., 3, 3] = 1.0
    return T


def invert_pose(T: np.ndarray) -> np.ndarray:
    """Rigid inverse [R^T, -R^T t] (src/LieUtils/utils.cu:38-69)."""
    T = np.asarray(T, dtype=np.float64)
    Rt = np.swapaxes(T[..., :3, :3], -1, -2)
    out = np.zeros(T.shape[:-2] + (4, 4))
    out[..., :3, :3] = Rt
    out[..., :3, 3] = -(Rt @ T[..., :3, 3, None])[..., 0]
    out[..., 3, 3] = 1.0
    return out


def project_rotations(T: np.ndarray) -> np.ndarray:
    """Poses with their rotation block replaced by the nearest rotation (polar
    factor U V^T of the SVD).  The library does this to every pose it uploads
    and to T_r_to_l; float32 inputs are rotations only to ~1e-7, and the
    reference (which never projects) computes in float32 anyway."""
    T = np.array(T, dtype=np.float64, copy=True)
    U, _, Vt = np.linalg.svd(T[..., :3, :3])
    R = U @ Vt
    ok = np.linalg.det(T[..., :3, :3]) > 1e-6
    T[..., :3, :3] = np.where(ok[..., None, None], R, T[..., :3, :3])
    return T


def global_poses(X: np.ndarray, T_rl: np.ndarray):
    """G_0 = I, G_{k+1} = G_k X_k^-1, G^R_k = G_k T_rl, serial
    (Solver_cuda.cu:440-493)."""
    N = X.shape[0]
    G = np.zeros((N + 1, 4, 4))
    G[0] = np.eye(4)
    for k in range(N):
        G[k + 1] = G[k] @ invert_pose(X[k])
    GR = G @ T_rl
    return G, GR


# ---------------------------------------------------------------------------
# robust kernels
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Options:
    mode: str = "lm"                 # "reference" or "lm"
    robust: str = "cauchy"           # "cauchy" | "huber" | "none"
    robust_param: float = CAUCHY_C2  # c^2 for cauchy, delta (px) for huber
    near_boost: bool = True          # x4 weight when |t - a| < 2 (Solver_cuda.cu:226-228)
    eps: float = EPS_C
    lambda0: float = 1e-4
    lam_up: float = 10.0
    lam_down: float = 0.1
    lam_min: float = 1e-10
    lam_max: float = 1e10
    alpha_min: float = ALPHA_MIN
    alpha_max: float = ALPHA_MAX
    step_size: float = 1.0


def robust(s: np.ndarray, opts: Options):
    """rho(s) and rho'(s) of the squared residual norm s.

    cauchy: rho = c2 log(1 + s/c2), rho' = 1/(1 + s/c2); the reference weight
    sqrt(1/(1 + s/9)) (Solver_cuda.cu:222) is sqrt(rho').
    huber:  rho = s (sqrt(s) <= d) else 2 d sqrt(s) - d^2; rho' = 1 or d/sqrt(s).
    """
    if opts.robust == "cauchy":
        c2 = opts.robust_param
        return c2 * np.log1p(s / c2), 1.0 / (1.0 + s / c2)
    if opts.robust == "huber":
        d = opts.robust_param
        rs = np.sqrt(s)
        inl = rs <= d
        rho = np.where(inl, s, 2.0 * d * rs - d * d)
        drho = np.where(inl, 1.0, d / np.where(inl, 1.0, rs))
        return rho, drho
    if opts.robust == "none":
        return s.copy(), np.ones_like(s)
    raise ValueError(opts.robust)


# ---------------------------------------------------------------------------
# problem view
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ProblemArrays:
    N: int
    L: int
    a: np.ndarray        # (O,) anchor frame
    t: np.ndarray        # (O,) target frame
    l: np.ndarray        # (O,) landmark slot
    z: np.ndarray        # (O, 2, 2) observed pixel [cam][xy]
    ray: np.ndarray      # (O, 2) anchor-left pixel defining the landmark ray
    Kc: np.ndarray       # (2, 3, 3) intrinsics per camera
    T_rl: np.ndarray     # (4, 4)


def problem_arrays(N: int, L: int, intrinsics, T_r_to_l, meas_a, meas_t, meas_l, uv) -> ProblemArrays:
    """Measurement arrays with the reference's observation-table semantics.

    ``writeObservations`` stores the pixels in ``_observations[target][feat]``
    (Solver.cpp:40-56); a measurement reads its pixels from
    ``observations[target][feat]`` and its ray from the left pixel of
    ``observations[anchor][feat]`` (Solver_cuda.cu:147, 217), an all-zero
    entry if that cell was never written.  Later writes to a cell win.
    """
    a = np.asarray(meas_a, dtype=np.int64)
    t = np.asarray(meas_t, dtype=np.int64)
    l = np.asarray(meas_l, dtype=np.int64)
    uv = np.asarray(uv, dtype=np.float32).astype(np.float64).reshape(-1, 4)
    table: Dict = {}
    for m in range(a.size):
        table[(int(t[m]), int(l[m]))] = uv[m]
    z = np.empty((a.size, 2, 2))
    ray = np.zeros((a.size, 2))
    for m in range(a.size):
        e = table[(int(t[m]), int(l[m]))]
        z[m, 0] = e[0:2]
        z[m, 1] = e[2:4]
        h = table.get((int(a[m]), int(l[m])))
        if h is not None:
            ray[m] = h[0:2]
    intr = np.asarray(intrinsics, dtype=np.float32).astype(np.float64).reshape(8)
    Kc = np.zeros((2, 3, 3))
    for c in range(2):
        fx, fy, cx, cy = intr[4 * c: 4 * c + 4]
        Kc[c] = [[fx, 0, cx], [0, fy, cy], [0, 0, 1]]
    T = project_rotations(np.asarray(T_r_to_l, dtype=np.float32).astype(np.float64).reshape(4, 4))
    return ProblemArrays(N=int(N), L=int(L), a=a, t=t, l=l, z=z, ray=ray, Kc=Kc, T_rl=T)


def from_problem(p) -> ProblemArrays:
    """ProblemArrays of a ``spgo.problems.Problem``."""
    return problem_arrays(p.n_poses, p.n_slots, p.intrinsics, p.T_r_to_l,
                          p.meas_a, p.meas_t, p.meas_l, p.uv)


# ---------------------------------------------------------------------------
# residuals and Jacobians (Solver_cuda.cu:49-310, LMutils.cu:538-628)
# ---------------------------------------------------------------------------
def _landmark_in_anchor(pa: ProblemArrays, alpha: np.ndarray) -> np.ndarray:
    """t_feat_in_anchor = Kl^-1 [p; 1] / alpha (Solver_cuda.cu:147-157)."""
    Kl = pa.Kc[0]
    x = (pa.ray[:, 0] - Kl[0, 2]) / Kl[0, 0]
    y = (pa.ray[:, 1] - Kl[1, 2]) / Kl[1, 1]
    al = alpha[pa.l]
    return np.stack([x / al, y / al, 1.0 / al], axis=-1)


def residuals(pa: ProblemArrays, X: np.ndarray, alpha: np.ndarray, opts: Options):
    """Unweighted residual e = z - zhat (O, 2cams, 2), weights w (O, 2) and
    per-(measurement, camera) cost 1/2 kappa rho(|e|^2)."""
    G, GR = global_poses(X, pa.T_rl)
    Pa = _landmark_in_anchor(pa, alpha)
    e = np.empty((pa.a.size, 2, 2))
    w = np.empty((pa.a.size, 2))
    cost = np.empty((pa.a.size, 2))
    kappa = np.where((np.abs(pa.t - pa.a) < 2) & opts.near_boost, 16.0, 1.0)
    for c in range(2):
        Gt = (G if c == 0 else GR)[pa.t]
        Tat = invert_pose(Gt) @ G[pa.a]
        Pt = (Tat[:, :3, :3] @ Pa[..., None])[..., 0] + Tat[:, :3, 3]
        K = pa.Kc[c]
        zh = np.stack([K[0, 0] * Pt[:, 0] / Pt[:, 2] + K[0, 2],
                       K[1, 1] * Pt[:, 1] / Pt[:, 2] + K[1, 2]], -1)
        e[:, c] = pa.z[:, c] - zh
        s = np.sum(e[:, c] ** 2, axis=-1)
        rho, drho = robust(s, opts)
        w[:, c] = np.sqrt(kappa * drho)
        cost[:, c] = 0.5 * kappa * rho
    return e, w, cost


@dataclasses.dataclass
class Linearization:
    cost: float
    r: np.ndarray            # (4O,) weighted residual, row 4m + 2c + xy
    J_T: sp.csr_matrix       # (4O, 6N) weighted pose Jacobian
    J_a: sp.csr_matrix       # (4O, L) weighted inverse-depth Jacobian
    A: np.ndarray            # (6N, 6N)
    g_T: np.ndarray          # (6N,)
    C: np.ndarray            # (L,)
    g_a: np.ndarray          # (L,)
    B: sp.csr_matrix         # (6N, L)


def jacobians(pa: ProblemArrays, X: np.ndarray, alpha: np.ndarray, opts: Options):
    """Weighted J_T, J_alpha and r exactly as the reference lays them out.

    Per measurement m and camera c (Solver_cuda.cu:113-309):
      M = K_c Pi'(P_t) (2x3), dz/dalpha = M R_{a->t} (-P_a / alpha),
      for i in [a, t): dz/dxi_i = M [-R_{i->t} [P_i]x, R_{i->t}] with
      T_{i->t} = (G^c_t)^-1 G_i and P_i = G_i^-1 G_a P_a
      (compute_del_tn_del_xi, LMutils.cu:552-582); rows 4m + 2c + xy.
    """
    N, L, O = pa.N, pa.L, pa.a.size
    G, GR = global_poses(X, pa.T_rl)
    Ginv = invert_pose(G)
    Pa = _landmark_in_anchor(pa, alpha)
    al = alpha[pa.l]
    e, w, cost = residuals(pa, X, alpha, opts)
    r = np.empty(4 * O)
    ja = np.empty(4 * O)
    rows_T, cols_T, vals_T = [], [], []
    span = np.maximum(pa.t - pa.a, 0)
    maxspan = int(span.max()) if O else 0
    for c in range(2):
        Gc = G if c == 0 else GR
        Gt_inv = invert_pose(Gc[pa.t])
        Tat = Gt_inv @ G[pa.a]
        Pt = (Tat[:, :3, :3] @ Pa[..., None])[..., 0] + Tat[:, :3, 3]
        K = pa.Kc[c]
        zx, zy, zz = Pt[:, 0], Pt[:, 1], Pt[:, 2]
        M = np.zeros((O, 2, 3))
        M[:, 0, 0] = K[0, 0] / zz
        M[:, 0, 2] = -K[0, 0] * zx / (zz * zz)
        M[:, 1, 1] = K[1, 1] / zz
        M[:, 1, 2] = -K[1, 1] * zy / (zz * zz)
        dza = (M @ (Tat[:, :3, :3] @ (-Pa / al[:, None])[..., None]))[..., 0]   # (O, 2)
        wc = w[:, c]
        for xy in range(2):
            row = 4 * np.arange(O) + 2 * c + xy
            r[row] = wc * e[:, c, xy]
            ja[row] = wc * dza[:, xy]
        for d in range(maxspan):
            sel = np.nonzero(span > d)[0]
            if sel.size == 0:
                continue
            i = pa.a[sel] + d
            T_it = Gt_inv[sel] @ G[i]
            P_i = (Ginv[i] @ G[pa.a[sel]])
            P_i = (P_i[:, :3, :3] @ Pa[sel][..., None])[..., 0] + P_i[:, :3, 3]
            R_it = T_it[:, :3, :3]
            D = np.concatenate([-R_it @ skew(P_i), R_it], axis=-1)           # (n, 3, 6)
            J = wc[sel, None, None] * (M[sel] @ D)                            # (n, 2, 6)
            for xy in range(2):
                rr = np.repeat(4 * sel + 2 * c + xy, 6)
                cc = (6 * i[:, None] + np.arange(6)[None, :]).reshape(-1)
                rows_T.append(rr)
                cols_T.append(cc)
                vals_T.append(J[:, xy, :].reshape(-1))
    if rows_T:
        rows_T = np.concatenate(rows_T)
        cols_T = np.concatenate(cols_T)
        vals_T = np.concatenate(vals_T)
    else:
        rows_T = cols_T = np.zeros(0, np.int64)
        vals_T = np.zeros(0)
    J_T = sp.csr_matrix((vals_T, (rows_T, cols_T)), shape=(4 * O, 6 * N))
    J_a = sp.csr_matrix((ja, (np.arange(4 * O), np.repeat(pa.l, 4))), shape=(4 * O, L))
    return float(cost.sum()), r, J_T, J_a


def linearize(pa: ProblemArrays, X: np.ndarray, alpha: np.ndarray, opts: Options) -> Linearization:
    """A = J_T^T J_T, g_T = J_T^T r, C = diag(J_a^T J_a), g_a = J_a^T r,
    B = J_T^T J_a (Solver_cuda.cu:636-712 with the atomics of 240-241, 306)."""
    cost, r, J_T, J_a = jacobians(pa, X, alpha, opts)
    JTt = J_T.T.tocsr()
    A = (JTt @ J_T).toarray()
    g_T = JTt @ r
    C = np.asarray(J_a.multiply(J_a).sum(axis=0)).reshape(-1)
    g_a = J_a.T @ r
    B = (JTt @ J_a).tocsr()
    return Linearization(cost=cost, r=r, J_T=J_T, J_a=J_a, A=A, g_T=g_T, C=C, g_a=g_a, B=B)


def cost_only(pa: ProblemArrays, X: np.ndarray, alpha: np.ndarray, opts: Options) -> float:
    return float(residuals(pa, X, alpha, opts)[2].sum())


# ---------------------------------------------------------------------------
# reduced system, solve, back-substitution, update (Solver_cuda.cu:322-374, 497-553,
# linear_solvers.cu:195-218)
# ---------------------------------------------------------------------------
def reduced_system(lin: Linearization, lam: float, eps: float):
    """H = A - B diag(1/(C+eps)) B^T + lam I, g = g_T - B (g_a/(C+eps))."""
    Cinv = 1.0 / (lin.C + eps)
    BCi = lin.B @ sp.diags(Cinv)
    H = lin.A - (BCi @ lin.B.T).toarray()
    g = lin.g_T - BCi @ lin.g_a
    if lam != 0.0:
        H = H + lam * np.eye(H.shape[0])
    return H, g


def solve_pose_step(H: np.ndarray, g: np.ndarray):
    """Cholesky H = L L^T and delta = cholesky_solve(g) (linear_solvers.cu:195-218);
    returns (delta, ok).  A failed factorisation returns ok=False (the
    reference ignores the info flag, linear_solvers.cu:205)."""
    try:
        c = sla.cho_factor(H, lower=True, check_finite=True)
    except (np.linalg.LinAlgError, ValueError):
        return np.zeros_like(g), False
    return sla.cho_solve(c, g), True


def back_substitute(lin: Linearization, delta: np.ndarray, eps: float) -> np.ndarray:
    """delta_alpha = (g_a - B^T delta) / (C + eps) (Solver_cuda.cu:361-374)."""
    return (lin.g_a - lin.B.T @ delta) / (lin.C + eps)


def retract(X: np.ndarray, alpha: np.ndarray, delta: np.ndarray, dalpha: np.ndarray, opts: Options):
    """X_k <- X_k Exp(s delta_k); alpha <- clamp(alpha + dalpha) with no step size
    on alpha (Solver_cuda.cu:511-551: line 551 overwrites 549)."""
    N = X.shape[0]
    Xn = X @ exp_se3(opts.step_size * delta.reshape(N, 6))
    an = np.clip(alpha + dalpha, opts.alpha_min, opts.alpha_max)
    return Xn, an


@dataclasses.dataclass
class IterRecord:
    cost_candidate: float
    cost: float          # cost of the current state after the accept/reject decision
    lam: float           # lambda used for this iteration's solve
    accepted: bool
    chol_ok: bool


def lm_solve(pa: ProblemArrays, X0: np.ndarray, alpha0: np.ndarray, iterations: int,
             opts: Optional[Options] = None, lam: Optional[float] = None):
    """Run ``iterations`` LM iterations from (X0, alpha0).

    Returns (X, alpha, lam, initial_cost, [IterRecord]).  In ``reference`` mode
    lambda is 0 and every candidate is accepted, which is the reference GPU
    step repeated ``iterations`` times (Solver_cuda.cu:594-782).
    """
    opts = opts or Options()
    X = project_rotations(np.asarray(X0, dtype=np.float64))
    alpha = np.asarray(alpha0, dtype=np.float64).copy()
    ref = opts.mode == "reference"
    if lam is None:
        lam = 0.0 if ref else opts.lambda0
    lin = linearize(pa, X, alpha, opts)
    c0 = lin.cost
    recs: List[IterRecord] = []
    for _ in range(iterations):
        H, g = reduced_system(lin, lam, opts.eps)
        delta, ok = solve_pose_step(H, g)
        if ok:
            dalpha = back_substitute(lin, delta, opts.eps)
        else:
            dalpha = np.zeros_like(alpha)
        Xc, ac = retract(X, alpha, delta, dalpha, opts)
        lin_c = linearize(pa, Xc, ac, opts)
        lam_used = lam
        if ref:
            accept = True
        else:
            accept = lin_c.cost < lin.cost
            lam = max(lam * opts.lam_down, opts.lam_min) if accept else min(lam * opts.lam_up, opts.lam_max)
        if accept:
            X, alpha, lin = Xc, ac, lin_c
        recs.append(IterRecord(lin_c.cost, lin.cost, lam_used, accept, ok))
    return X, alpha, lam, c0, recs


# ---------------------------------------------------------------------------
# accuracy helpers (CUDAmanager.py:100-109)
# ---------------------------------------------------------------------------
def log_so3(R: np.ndarray) -> np.ndarray:
    """Rotation vector of R (float64, principal branch)."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_matrix(R).as_rotvec()


def pose_errors(X_true: np.ndarray, X_est: np.ndarray) -> np.ndarray:
    """Rotation angle (rad) and translation norm (m) of X_true^-1 X_est per pose."""
    E = invert_pose(X_true) @ np.asarray(X_est, dtype=np.float64)
    rot = np.linalg.norm(log_so3(E[:, :3, :3]), axis=-1)
    tr = np.linalg.norm(E[:, :3, 3], axis=-1)
    return np.stack([rot, tr], -1)


def depth_from_disparity(intrinsics, T_rl, pl, pr):
    """Least-squares stereo depth along the left ray (PythonUtils/Optimizer.py:87-114):
    b_c = K_c^-1 p_c (homogeneous pixels, n x 3), v = R_rl b_r, N = [v]x,
    depth = ((N b_l) . (N t_rl)) / |N b_l|^2.  Returns depths (n,)."""
    intr = np.asarray(intrinsics, dtype=np.float64)
    T = np.asarray(T_rl, dtype=np.float64).reshape(4, 4)
    pl = np.asarray(pl, dtype=np.float64)
    pr = np.asarray(pr, dtype=np.float64)
    Kl = np.array([[intr[0], 0, intr[2]], [0, intr[1], intr[3]], [0, 0, 1.0]])
    Kr = np.array([[intr[4], 0, intr[6]], [0, intr[5], intr[7]], [0, 0, 1.0]])
    bl = pl @ np.linalg.inv(Kl).T
    br = pr @ np.linalg.inv(Kr).T
    v = br @ T[:3, :3].T
    nb = np.cross(v, bl)
    nt = np.cross(v, T[:3, 3][None, :])
    return np.sum(nb * nt, axis=1) / np.sum(nb * nb, axis=1)
