"""Seeded synthetic stereo bundle-adjustment problems.

The scene follows the structure of the reference simulator
(PythonUtils/Simulator.py:13-91, PythonUtils/StereoSetup.py:36-73): stereo
keyframes on a trajectory that look at the origin, landmarks in a box, each
landmark stored as an inverse depth along the ray of its left pixel in a host
(anchor) keyframe and observed by both cameras of later keyframes.  The
measurement enumeration matches CUDAmanager.py:69-79 (host frame first, then
later frames, both cameras valid).  Unlike the reference simulator every random
draw comes from one ``numpy.random.default_rng(seed)`` stream, so a problem is
fully determined by its parameters.

Conventions
-----------
* ``X[k]`` (k = 0..N-1) is the incremental pose ``T_curr_to_next`` of the
  reference: it maps frame-k coordinates into frame k+1 (Simulator.py:68-80),
  ``G_{k+1} = G_k X_k^-1`` with ``G_0 = I``.
* The right camera pose in the left camera frame is ``T_r_to_l``
  (StereoSetup.py:26-33); ``G^R = G T_r_to_l``.
* Intrinsics are packed ``[fxL, fyL, cxL, cyL, fxR, fyR, cxR, cyR]``
  (CUDAmanager.py:86-91).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np


@dataclasses.dataclass
class Problem:
    """One stereo bundle-adjustment problem in the reference's API terms."""

    n_poses: int                 # N incremental poses, N+1 frames
    m_per_frame: int             # M, the CudaSolver(N, M) capacity parameter
    intrinsics: np.ndarray       # float32 (8,)
    T_r_to_l: np.ndarray         # float32 (16,) row-major
    meas_a: np.ndarray           # int32 (O,) anchor frame per measurement
    meas_t: np.ndarray           # int32 (O,) target frame per measurement
    meas_l: np.ndarray           # int32 (O,) landmark id per measurement
    uv: np.ndarray               # float32 (O, 4): uL, vL, uR, vR
    X0: np.ndarray               # float32 (N, 4, 4) initial incremental poses
    alpha0: np.ndarray           # float32 (N*M,) initial inverse depths
    X_true: np.ndarray           # float64 (N, 4, 4)
    alpha_true: np.ndarray       # float64 (N*M,)
    n_landmarks: int             # landmarks actually used (ids 0..n_landmarks-1)
    outlier: Optional[np.ndarray] = None   # bool (O,) injected outliers
    name: str = ""

    @property
    def n_slots(self) -> int:
        return self.n_poses * self.m_per_frame

    @property
    def n_obs(self) -> int:
        return int(self.meas_a.shape[0])


# --------------------------------------------------------------------------
# small SE(3) helpers (float64, generator-only)
# --------------------------------------------------------------------------
def _rotvec_to_matrix(w: np.ndarray) -> np.ndarray:
    """Rodrigues' formula for a batch of rotation vectors (..., 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1)[..., None, None]
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1] = -w[..., 2]
    K[..., 0, 2] = w[..., 1]
    K[..., 1, 0] = w[..., 2]
    K[..., 1, 2] = -w[..., 0]
    K[..., 2, 0] = -w[..., 1]
    K[..., 2, 1] = w[..., 0]
    small = th < 1e-12
    ths = np.where(small, 1.0, th)
    a = np.where(small, 1.0, np.sin(ths) / ths)
    b = np.where(small, 0.5, (1.0 - np.cos(ths)) / (ths * ths))
    return np.eye(3) + a * K + b * (K @ K)


def _exp_se3(xi: np.ndarray) -> np.ndarray:
    """Exp of a batch of twists (psi, rho) with t = Jl(psi) rho."""
    xi = np.asarray(xi, dtype=np.float64)
    psi, rho = xi[..., :3], xi[..., 3:]
    th = np.linalg.norm(psi, axis=-1)[..., None, None]
    K = np.zeros(psi.shape[:-1] + (3, 3))
    K[..., 0, 1] = -psi[..., 2]
    K[..., 0, 2] = psi[..., 1]
    K[..., 1, 0] = psi[..., 2]
    K[..., 1, 2] = -psi[..., 0]
    K[..., 2, 0] = -psi[..., 1]
    K[..., 2, 1] = psi[..., 0]
    small = th < 1e-8
    ths = np.where(small, 1.0, th)
    c1 = np.where(small, 0.5, (1 - np.cos(ths)) / ths ** 2)
    c2 = np.where(small, 1.0 / 6.0, (ths - np.sin(ths)) / ths ** 3)
    Jl = np.eye(3) + c1 * K + c2 * (K @ K)
    T = np.zeros(xi.shape[:-1] + (4, 4))
    T[..., :3, :3] = _rotvec_to_matrix(psi)
    T[..., :3, 3] = (Jl @ rho[..., None])[..., 0]
    T[..., 3, 3] = 1.0
    return T


def _inv(T: np.ndarray) -> np.ndarray:
    R = T[..., :3, :3]
    t = T[..., :3, 3]
    Ti = np.zeros_like(T)
    Rt = np.swapaxes(R, -1, -2)
    Ti[..., :3, :3] = Rt
    Ti[..., :3, 3] = -(Rt @ t[..., None])[..., 0]
    Ti[..., 3, 3] = 1.0
    return Ti


def look_at_origin(position: np.ndarray) -> np.ndarray:
    """Camera-to-global pose of a pinhole camera (x right, y down, z forward)
    at ``position`` looking at the origin with world z up."""
    f = -position / np.linalg.norm(position)
    up = np.array([0.0, 0.0, 1.0])
    x = np.cross(f, up)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    T = np.eye(4)
    T[:3, 0], T[:3, 1], T[:3, 2] = x, y, f
    T[:3, 3] = position
    return T


# --------------------------------------------------------------------------
# generator
# --------------------------------------------------------------------------
def make_problem(n_poses: int, n_landmarks: int, obs_per_landmark, seed: int,
                 fx: float = 500.0, fy: float = 500.0, cx: float = 320.0, cy: float = 240.0,
                 width: int = 640, height: int = 480, baseline: float = 0.12,
                 radius: float = 2.0, box: float = 4.0, noise_px: float = 1.0,
                 rot_perturb: float = 0.05, trans_perturb: float = 0.1,
                 alpha_perturb: float = 0.1, outlier_frac: float = 0.0,
                 outlier_px: float = 50.0, min_depth: float = 0.2,
                 m_per_frame: Optional[int] = None, name: str = "") -> Problem:
    """Seeded stereo problem on a circular trajectory.

    * ``n_poses + 1`` frames evenly spaced on a circle of ``radius`` metres in
      the z=0 plane, each looking at the origin.
    * Landmarks uniform in a cube of side ``box`` centred at the origin; each
      is hosted by frame ``a`` and observed (both cameras valid: depth above
      ``min_depth``, pixel inside the image) in the ``obs_per_landmark``
      consecutive frames ``a .. a+obs_per_landmark-1``.  Candidates without such
      a run are redrawn.  ``obs_per_landmark`` may be a pair ``(lo, hi)``: each
      candidate then draws its own count uniformly in ``lo..hi``.
    * Pixel noise N(0, noise_px^2) on every coordinate; a fraction
      ``outlier_frac`` of the non-host stereo observations get an extra
      U(-outlier_px, outlier_px) error per coordinate.
    * Initial incremental poses ``X_true[k] Exp(xi)`` with the rotation and
      translation parts of xi drawn N(0, s^2/3) per axis (RMS norms
      ``rot_perturb`` rad and ``trans_perturb`` m); initial inverse depths
      ``alpha_true (1 + U(-alpha_perturb, alpha_perturb))``.
    """
    rng = np.random.default_rng(seed)
    N = int(n_poses)
    F = N + 1
    if np.ndim(obs_per_landmark) == 0:
        K = int(obs_per_landmark)
        k_lo = k_hi = K
    else:
        k_lo, k_hi = (int(v) for v in obs_per_landmark)
        K = None
    if k_lo < 1 or k_hi > F or k_lo > k_hi:
        raise ValueError("obs_per_landmark must be in [1, n_poses + 1]")

    theta = 2.0 * math.pi * np.arange(F) / F
    G = np.stack([look_at_origin(np.array([radius * math.cos(a), radius * math.sin(a), 0.0]))
                  for a in theta])                       # camera-to-global, left camera
    T_rl = np.eye(4)
    T_rl[0, 3] = baseline                                # right camera in the left frame
    GR = G @ T_rl
    Ginv = _inv(G)
    GRinv = _inv(GR)

    Kmat = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1.0]])

    def visible(P):
        # both cameras of every frame see P (n,3): depth above min_depth, pixel
        # inside the image -> (n, F); in chunks of candidates to stay in cache
        out = np.empty((P.shape[0], F), dtype=bool)
        for c0 in range(0, P.shape[0], 8192):
            Q = P[c0:c0 + 8192]
            ok = np.ones((Q.shape[0], F), dtype=bool)
            for Tinv in (Ginv, GRinv):
                R, t = Tinv[:, :3, :3], Tinv[:, :3, 3]
                Pc = [Q[:, 0:1] * R[None, :, i, 0] + Q[:, 1:2] * R[None, :, i, 1] + Q[:, 2:3] * R[None, :, i, 2]
                      + t[None, :, i] for i in range(3)]
                z = Pc[2]
                zs = np.where(np.abs(z) < 1e-12, 1e-12, z)
                u = fx * Pc[0] / zs + cx
                v = fy * Pc[1] / zs + cy
                ok &= (z > min_depth) & (u >= 0) & (u < width) & (v >= 0) & (v < height)
            out[c0:c0 + 8192] = ok
        return out

    def draw(k: int, need: int):
        # `need` landmarks with a run of k consecutive visible frames
        pts, hosts = [], []
        n_start = F - k + 1
        while need > 0:
            m = max(4 * need, 1024)
            cand = rng.uniform(-box / 2, box / 2, size=(m, 3))
            ok = visible(cand)
            # run of k consecutive visible frames starting at s, s in [0, F-k]
            c = np.concatenate([np.zeros((m, 1), np.int64), np.cumsum(ok, axis=1)], axis=1)
            run = (c[:, k:k + n_start] - c[:, 0:n_start]) == k          # (m, n_start)
            has = run.any(axis=1)
            # choose a start uniformly among valid ones
            score = np.where(run, rng.random(run.shape), -1.0)
            start = np.argmax(score, axis=1)
            idx = np.nonzero(has)[0][:need]
            pts.append(cand[idx])
            hosts.append(start[idx])
            need -= idx.size
        return np.concatenate(pts), np.concatenate(hosts)

    if K is not None:
        pts, hosts = draw(K, int(n_landmarks))
        pts, hosts = [pts], [hosts]
        counts = [np.full(hosts[0].size, K, np.int64)]
    else:
        # every landmark's count drawn first, so the counts are exactly uniform
        kd = rng.integers(k_lo, k_hi + 1, size=int(n_landmarks))
        pts, hosts, counts = [], [], []
        for k in range(k_lo, k_hi + 1):
            nk = int((kd == k).sum())
            if nk:
                pk, hk = draw(k, nk)
                pts.append(pk)
                hosts.append(hk)
                counts.append(np.full(nk, k, np.int64))

    P = np.concatenate(pts)[:n_landmarks]
    host = np.concatenate(hosts)[:n_landmarks].astype(np.int64)
    kl = np.concatenate(counts)[:n_landmarks].astype(np.int64)
    order = np.argsort(host, kind="stable")       # landmark ids grouped by host frame
    P, host, kl = P[order], host[order], kl[order]
    L = P.shape[0]

    # measurements: landmark-major, frames host..host+k-1 (CUDAmanager.py:69-79)
    meas_l = np.repeat(np.arange(L), kl)
    meas_a = np.repeat(host, kl)
    first = np.concatenate([[0], np.cumsum(kl)[:-1]])
    meas_t = meas_a + (np.arange(meas_l.size) - np.repeat(first, kl))
    Pm = P[meas_l]
    Pcl = np.einsum('nij,nj->ni', Ginv[meas_t, :3, :3], Pm) + Ginv[meas_t, :3, 3]
    Pcr = np.einsum('nij,nj->ni', GRinv[meas_t, :3, :3], Pm) + GRinv[meas_t, :3, 3]
    uv = np.empty((meas_l.size, 4))
    uv[:, 0] = fx * Pcl[:, 0] / Pcl[:, 2] + cx
    uv[:, 1] = fy * Pcl[:, 1] / Pcl[:, 2] + cy
    uv[:, 2] = fx * Pcr[:, 0] / Pcr[:, 2] + cx
    uv[:, 3] = fy * Pcr[:, 1] / Pcr[:, 2] + cy
    uv += rng.normal(0.0, noise_px, size=uv.shape)
    outlier = np.zeros(meas_l.size, dtype=bool)
    if outlier_frac > 0:
        outlier = (rng.random(meas_l.size) < outlier_frac) & (meas_t != meas_a)
        uv[outlier] += rng.uniform(-outlier_px, outlier_px, size=(int(outlier.sum()), 4))

    # true inverse depth along the host-left ray: alpha = 1 / z_host
    z_host = (np.einsum('nij,nj->ni', Ginv[host, :3, :3], P) + Ginv[host, :3, 3])[:, 2]
    M = m_per_frame if m_per_frame is not None else int(math.ceil(L / N))
    if N * M < L:
        raise ValueError("m_per_frame too small for the landmark count")
    alpha_true = np.ones(N * M)
    alpha_true[:L] = 1.0 / z_host

    # true incremental poses X_k = G_{k+1}^-1 G_k (Simulator.py:68-80)
    X_true = Ginv[1:] @ G[:-1]
    xi = np.concatenate([rng.normal(0.0, rot_perturb / math.sqrt(3.0), size=(N, 3)),
                         rng.normal(0.0, trans_perturb / math.sqrt(3.0), size=(N, 3))], axis=1)
    X0 = X_true @ _exp_se3(xi)
    alpha0 = alpha_true.copy()
    alpha0[:L] *= 1.0 + rng.uniform(-alpha_perturb, alpha_perturb, size=L)

    intr = np.array([fx, fy, cx, cy, fx, fy, cx, cy], dtype=np.float32)
    return Problem(n_poses=N, m_per_frame=M, intrinsics=intr,
                   T_r_to_l=T_rl.reshape(16).astype(np.float32),
                   meas_a=meas_a.astype(np.int32), meas_t=meas_t.astype(np.int32),
                   meas_l=meas_l.astype(np.int32), uv=uv.astype(np.float32),
                   X0=X0.astype(np.float32), alpha0=alpha0.astype(np.float32),
                   X_true=X_true, alpha_true=alpha_true, n_landmarks=L,
                   outlier=outlier, name=name)


# Named workloads (sizes from the project's benchmark definitions).
CONFIGS = {
    # 10 keyframes, 500 landmarks, 4 observing keyframes each, 10 iterations
    "c1_small": dict(n_poses=10, n_landmarks=500, obs_per_landmark=4, seed=1001),
    # 64 keyframes, 20k landmarks, 8 stereo observations each, 20 iterations
    "c2_fp64": dict(n_poses=64, n_landmarks=20000, obs_per_landmark=8, seed=2002),
    # 256 keyframes, 100k landmarks, 10 observations, 5% outliers (Huber)
    "c3_mixed": dict(n_poses=256, n_landmarks=100000, obs_per_landmark=10, seed=3003,
                     outlier_frac=0.05, outlier_px=50.0),
}


def make_long_problem(n_poses: int, n_landmarks: int, obs_per_landmark: int, seed: int,
                      fx: float = 500.0, fy: float = 500.0, cx: float = 320.0, cy: float = 240.0,
                      width: int = 640, height: int = 480, baseline: float = 0.12, radius: float = 2.0,
                      depth_range=(1.0, 4.0), noise_px: float = 1.0, rot_perturb: float = 0.05,
                      trans_perturb: float = 0.1, alpha_perturb: float = 0.1, min_depth: float = 0.2,
                      name: str = "") -> Problem:
    """Seeded stereo problem for long trajectories (memory linear in the
    landmark count).  Same circle of ``n_poses + 1`` keyframes looking at the
    origin as ``make_problem``; each landmark is drawn in its host frame's view
    (pixel uniform in the image, depth uniform in ``depth_range``) and kept if
    both cameras see it in the ``obs_per_landmark`` consecutive frames from the
    host on.  Noise and initial perturbations as in ``make_problem``."""
    rng = np.random.default_rng(seed)
    N = int(n_poses)
    F = N + 1
    K = int(obs_per_landmark)
    if K < 1 or K > F:
        raise ValueError("obs_per_landmark must be in [1, n_poses + 1]")
    theta = 2.0 * math.pi * np.arange(F) / F
    G = np.stack([look_at_origin(np.array([radius * math.cos(a), radius * math.sin(a), 0.0])) for a in theta])
    T_rl = np.eye(4)
    T_rl[0, 3] = baseline
    Ginv = _inv(G)
    GRinv = _inv(G @ T_rl)

    def proj(Tinv, P):
        Pc = np.einsum('nij,nj->ni', Tinv[:, :3, :3], P) + Tinv[:, :3, 3]
        z = Pc[:, 2]
        zs = np.where(np.abs(z) < 1e-12, 1e-12, z)
        return np.stack([fx * Pc[:, 0] / zs + cx, fy * Pc[:, 1] / zs + cy], -1), z

    pts, hosts = [], []
    need = int(n_landmarks)
    while need > 0:
        m = max(2 * need, 1024)
        a = rng.integers(0, F - K + 1, size=m)
        u = rng.uniform(0, width, size=m)
        v = rng.uniform(0, height, size=m)
        d = rng.uniform(depth_range[0], depth_range[1], size=m)
        Pc = np.stack([(u - cx) / fx * d, (v - cy) / fy * d, d], -1)
        P = np.einsum('nij,nj->ni', G[a, :3, :3], Pc) + G[a, :3, 3]
        ok = np.ones(m, bool)
        for k in range(K):
            t = a + k
            pl, zl = proj(Ginv[t], P)
            pr, zr = proj(GRinv[t], P)
            ok &= ((zl > min_depth) & (zr > min_depth) & (pl[:, 0] >= 0) & (pl[:, 0] < width)
                   & (pl[:, 1] >= 0) & (pl[:, 1] < height) & (pr[:, 0] >= 0) & (pr[:, 0] < width)
                   & (pr[:, 1] >= 0) & (pr[:, 1] < height))
        idx = np.nonzero(ok)[0][:need]
        pts.append(P[idx])
        hosts.append(a[idx])
        need -= idx.size
    P = np.concatenate(pts)
    host = np.concatenate(hosts).astype(np.int64)
    order = np.argsort(host, kind="stable")
    P, host = P[order], host[order]
    L = P.shape[0]
    meas_l = np.repeat(np.arange(L), K)
    meas_a = np.repeat(host, K)
    meas_t = meas_a + np.tile(np.arange(K), L)
    pl, _ = proj(Ginv[meas_t], P[meas_l])
    pr, _ = proj(GRinv[meas_t], P[meas_l])
    uv = np.concatenate([pl, pr], 1) + rng.normal(0.0, noise_px, size=(meas_l.size, 4))
    z_host = (np.einsum('nij,nj->ni', Ginv[host, :3, :3], P) + Ginv[host, :3, 3])[:, 2]
    M = int(math.ceil(L / N))
    alpha_true = np.ones(N * M)
    alpha_true[:L] = 1.0 / z_host
    X_true = Ginv[1:] @ G[:-1]
    xi = np.concatenate([rng.normal(0.0, rot_perturb / math.sqrt(3.0), size=(N, 3)),
                         rng.normal(0.0, trans_perturb / math.sqrt(3.0), size=(N, 3))], axis=1)
    X0 = X_true @ _exp_se3(xi)
    alpha0 = alpha_true.copy()
    alpha0[:L] *= 1.0 + rng.uniform(-alpha_perturb, alpha_perturb, size=L)
    intr = np.array([fx, fy, cx, cy, fx, fy, cx, cy], dtype=np.float32)
    return Problem(n_poses=N, m_per_frame=M, intrinsics=intr, T_r_to_l=T_rl.reshape(16).astype(np.float32),
                   meas_a=meas_a.astype(np.int32), meas_t=meas_t.astype(np.int32), meas_l=meas_l.astype(np.int32),
                   uv=uv.astype(np.float32), X0=X0.astype(np.float32), alpha0=alpha0.astype(np.float32),
                   X_true=X_true, alpha_true=alpha_true, n_landmarks=L,
                   outlier=np.zeros(meas_l.size, dtype=bool), name=name)


# --------------------------------------------------------------------------
# on-disk format: one .npz of plain arrays (no pickled objects)
# --------------------------------------------------------------------------
_NPZ_FIELDS = ("meas_a", "meas_t", "meas_l", "uv", "intrinsics", "T_r_to_l", "X0", "alpha0", "X_true", "alpha_true")
NPZ_VERSION = 1


def save_problem(path: str, p: Problem, seed: Optional[int] = None) -> None:
    """Write a problem as an .npz: measurement ids (int32 SoA), pixels (float32
    O x 4: uL vL uR vR), calibration, initial and true states, sizes and the
    generator seed.  ``load_problem`` reads it back without unpickling."""
    arrays = {k: np.asarray(getattr(p, k)) for k in _NPZ_FIELDS}
    arrays["outlier"] = np.asarray(p.outlier if p.outlier is not None else np.zeros(p.n_obs, bool))
    arrays["sizes"] = np.array([NPZ_VERSION, p.n_poses, p.m_per_frame, p.n_landmarks,
                                -1 if seed is None else int(seed)], np.int64)
    arrays["name"] = np.frombuffer(p.name.encode(), np.uint8)
    np.savez(path, **arrays)


def load_problem(path: str) -> Problem:
    with np.load(path, allow_pickle=False) as z:
        sizes = z["sizes"]
        if int(sizes[0]) != NPZ_VERSION:
            raise ValueError(f"{path}: format version {int(sizes[0])}, expected {NPZ_VERSION}")
        kw = {k: z[k] for k in _NPZ_FIELDS}
        return Problem(n_poses=int(sizes[1]), m_per_frame=int(sizes[2]), n_landmarks=int(sizes[3]),
                       outlier=z["outlier"], name=bytes(z["name"]).decode(), **kw)


# Batch of independent sliding windows: keyframes per window uniform in 8..32,
# landmarks uniform in 200..2000, observations per landmark uniform in 3..8.
WINDOW_BATCH = dict(n_windows=512, n_range=(8, 32), l_range=(200, 2000), k_range=(3, 8), seed=4004)


def make_window_batch(n_windows: int, n_range=(8, 32), l_range=(200, 2000), k_range=(3, 8),
                      seed: int = 4004, **kw) -> list:
    """``n_windows`` independent problems from the same generator, each with its
    own keyframe count, landmark count and seed (drawn from one stream)."""
    rng = np.random.default_rng(seed)
    out = []
    for w in range(int(n_windows)):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        L = int(rng.integers(l_range[0], l_range[1] + 1))
        s = int(rng.integers(0, 2**31 - 1))
        out.append(make_problem(n, L, (k_range[0], k_range[1]), s, name=f"window{w}", **kw))
    return out


def make_config(name: str, **overrides) -> Problem:
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    return make_problem(name=name, **kw)


def map_value_to_index(v: int, x: int, n: int) -> int:
    """Anchor frame of landmark ``v`` of ``x`` when ``n`` frames host them in
    equal shares (PythonUtils/Optimizer.py:12-18)."""
    vc = max(0, min(v, x))
    return min(int((vc / x) * n), n - 1)


def make_manager_problem(n: int, m: int, seed: int, arc: float = 0.6, radius: float = 2.0,
                         depth_range=(1.0, 4.0), noise_px: float = 1.0, fx: float = 500.0, fy: float = 500.0,
                         cx: float = 320.0, cy: float = 240.0, width: int = 640, height: int = 480,
                         baseline: float = 0.12, min_depth: float = 0.2, name: str = "") -> Problem:
    """A problem enumerated the way the reference's harness feeds its solver
    (CUDAmanager.py:69-79): ``n`` incremental poses (``n + 1`` frames), ``n m``
    landmarks, landmark ``l`` hosted by frame ``map_value_to_index(l, n m, n)``
    and written for every frame from its host to the last in which both
    cameras see it.  A landmark's frames need not be consecutive and may span
    the whole trajectory.  Frames sit on an arc of ``arc`` radians of a circle
    of ``radius`` metres, looking at the origin; each landmark is drawn in its
    host's view (pixel uniform in the image, depth uniform in ``depth_range``).
    The initial state is the harness's: identity incremental poses and inverse
    depths 1 (include/Solver.hpp:36-44)."""
    rng = np.random.default_rng(seed)
    N, F, L = int(n), int(n) + 1, int(n) * int(m)
    theta = arc * (np.arange(F) / max(F - 1, 1) - 0.5)
    G = np.stack([look_at_origin(np.array([radius * math.cos(a), radius * math.sin(a), 0.0])) for a in theta])
    T_rl = np.eye(4)
    T_rl[0, 3] = baseline
    Ginv = _inv(G)
    GRinv = _inv(G @ T_rl)
    host = np.array([map_value_to_index(l, L, N) for l in range(L)], np.int64)
    u = rng.uniform(0, width, size=L)
    v = rng.uniform(0, height, size=L)
    d = rng.uniform(depth_range[0], depth_range[1], size=L)
    Pc = np.stack([(u - cx) / fx * d, (v - cy) / fy * d, d], -1)
    P = np.einsum('nij,nj->ni', G[host, :3, :3], Pc) + G[host, :3, 3]

    def proj(Tinv, Q):
        Qc = np.einsum('ij,nj->ni', Tinv[:3, :3], Q) + Tinv[:3, 3]
        z = Qc[:, 2]
        zs = np.where(np.abs(z) < 1e-12, 1e-12, z)
        return np.stack([fx * Qc[:, 0] / zs + cx, fy * Qc[:, 1] / zs + cy], -1), z

    ma, mt, ml, uvs = [], [], [], []
    for f in range(F):
        pl, zl = proj(Ginv[f], P)
        pr, zr = proj(GRinv[f], P)
        ok = ((host <= f) & (zl > min_depth) & (zr > min_depth) & (pl[:, 0] >= 0) & (pl[:, 0] < width)
              & (pl[:, 1] >= 0) & (pl[:, 1] < height) & (pr[:, 0] >= 0) & (pr[:, 0] < width)
              & (pr[:, 1] >= 0) & (pr[:, 1] < height))
        ok[host == f] = True        # the host always sees its own landmarks
        idx = np.nonzero(ok)[0]
        ma.append(host[idx])
        mt.append(np.full(idx.size, f, np.int64))
        ml.append(idx)
        uvs.append(np.concatenate([pl[idx], pr[idx]], 1))
    ma, mt, ml, uv = (np.concatenate(x) for x in (ma, mt, ml, uvs))
    order = np.lexsort((mt, ml))                # landmark-major, frames ascending
    ma, mt, ml, uv = ma[order], mt[order], ml[order], uv[order]
    uv = uv + rng.normal(0.0, noise_px, size=uv.shape)
    z_host = (np.einsum('nij,nj->ni', Ginv[host, :3, :3], P) + Ginv[host, :3, 3])[:, 2]
    X_true = Ginv[1:] @ G[:-1]
    intr = np.array([fx, fy, cx, cy, fx, fy, cx, cy], dtype=np.float32)
    return Problem(n_poses=N, m_per_frame=int(m), intrinsics=intr, T_r_to_l=T_rl.reshape(16).astype(np.float32),
                   meas_a=ma.astype(np.int32), meas_t=mt.astype(np.int32), meas_l=ml.astype(np.int32),
                   uv=uv.astype(np.float32), X0=np.tile(np.eye(4, dtype=np.float32), (N, 1, 1)),
                   alpha0=np.ones(L, np.float32), X_true=X_true, alpha_true=1.0 / z_host, n_landmarks=L,
                   outlier=np.zeros(ma.size, dtype=bool), name=name)


def make_revisit_problem(n_poses: int, n_landmarks: int, obs_per_landmark: int, seed: int,
                         revisit_frac: float = 0.2, revisit_obs: int = 4, fx: float = 500.0, fy: float = 500.0,
                         cx: float = 320.0, cy: float = 240.0, width: int = 640, height: int = 480,
                         baseline: float = 0.12, radius: float = 2.0, depth_range=(1.0, 4.0),
                         noise_px: float = 1.0, rot_perturb: float = 0.05, trans_perturb: float = 0.1,
                         alpha_perturb: float = 0.1, min_depth: float = 0.2, name: str = "") -> Problem:
    """A trajectory with loop revisits: the first ``L = round((1 - revisit_frac)
    (n_poses + 1))`` frames go once round the circle of ``make_long_problem``;
    the last ``revisit_frac`` of the frames go round it again (1 cm further out),
    so frame ``f + L`` revisits the place of frame ``f``.  Each landmark is drawn
    in its host frame's view and observed by the ``obs_per_landmark``
    consecutive frames from the host on; a landmark whose host is revisited is
    also observed by up to ``revisit_obs`` consecutive frames from the host's
    revisit on (those that see it in both cameras), far after its anchor.
    Noise and initial perturbations as in ``make_problem``."""
    rng = np.random.default_rng(seed)
    N, F, K = int(n_poses), int(n_poses) + 1, int(obs_per_landmark)
    Lp = int(round((1.0 - revisit_frac) * F))
    f = np.arange(F)
    theta = 2.0 * math.pi * (f % Lp) / Lp
    rad = radius + 0.01 * (f // Lp)
    G = np.stack([look_at_origin(np.array([r * math.cos(a), r * math.sin(a), 0.0])) for r, a in zip(rad, theta)])
    T_rl = np.eye(4)
    T_rl[0, 3] = baseline
    Ginv = _inv(G)
    GRinv = _inv(G @ T_rl)

    def visible(fr, P):
        ok = np.ones(P.shape[0], bool)
        uv = []
        for Tinv in (Ginv[fr], GRinv[fr]):
            Pc = np.einsum('nij,nj->ni', Tinv[:, :3, :3], P) + Tinv[:, :3, 3]
            z = Pc[:, 2]
            zs = np.where(np.abs(z) < 1e-12, 1e-12, z)
            u = fx * Pc[:, 0] / zs + cx
            v = fy * Pc[:, 1] / zs + cy
            ok &= (z > min_depth) & (u >= 0) & (u < width) & (v >= 0) & (v < height)
            uv.append(np.stack([u, v], -1))
        return ok, np.concatenate(uv, 1)

    pts, hosts = [], []
    need = int(n_landmarks)
    while need > 0:
        m = max(2 * need, 1024)
        a = rng.integers(0, F - K + 1, size=m)
        u = rng.uniform(0, width, size=m)
        v = rng.uniform(0, height, size=m)
        d = rng.uniform(depth_range[0], depth_range[1], size=m)
        Pc = np.stack([(u - cx) / fx * d, (v - cy) / fy * d, d], -1)
        P = np.einsum('nij,nj->ni', G[a, :3, :3], Pc) + G[a, :3, 3]
        ok = np.ones(m, bool)
        for k in range(K):
            ok &= visible(a + k, P)[0]
        idx = np.nonzero(ok)[0][:need]
        pts.append(P[idx])
        hosts.append(a[idx])
        need -= idx.size
    P = np.concatenate(pts)
    host = np.concatenate(hosts).astype(np.int64)
    order = np.argsort(host, kind="stable")
    P, host = P[order], host[order]
    L = P.shape[0]
    ml, mt = [np.repeat(np.arange(L), K)], [(host[:, None] + np.arange(K)[None, :]).reshape(-1)]
    for j in range(int(revisit_obs)):
        b = host + Lp + j
        cand = np.nonzero(b <= N)[0]
        ok, _ = visible(b[cand], P[cand])
        ml.append(cand[ok])
        mt.append(b[cand[ok]])
    ml, mt = np.concatenate(ml), np.concatenate(mt)
    o = np.lexsort((mt, ml))
    ml, mt = ml[o], mt[o]
    ma = host[ml]
    _, uv = visible(mt, P[ml])
    uv = uv + rng.normal(0.0, noise_px, size=uv.shape)
    z_host = (np.einsum('nij,nj->ni', Ginv[host, :3, :3], P) + Ginv[host, :3, 3])[:, 2]
    M = int(math.ceil(L / N))
    alpha_true = np.ones(N * M)
    alpha_true[:L] = 1.0 / z_host
    X_true = Ginv[1:] @ G[:-1]
    xi = np.concatenate([rng.normal(0.0, rot_perturb / math.sqrt(3.0), size=(N, 3)),
                         rng.normal(0.0, trans_perturb / math.sqrt(3.0), size=(N, 3))], axis=1)
    X0 = X_true @ _exp_se3(xi)
    alpha0 = alpha_true.copy()
    alpha0[:L] *= 1.0 + rng.uniform(-alpha_perturb, alpha_perturb, size=L)
    intr = np.array([fx, fy, cx, cy, fx, fy, cx, cy], dtype=np.float32)
    return Problem(n_poses=N, m_per_frame=M, intrinsics=intr, T_r_to_l=T_rl.reshape(16).astype(np.float32),
                   meas_a=ma.astype(np.int32), meas_t=mt.astype(np.int32), meas_l=ml.astype(np.int32),
                   uv=uv.astype(np.float32), X0=X0.astype(np.float32), alpha0=alpha0.astype(np.float32),
                   X_true=X_true, alpha_true=alpha_true, n_landmarks=L,
                   outlier=np.zeros(ml.size, dtype=bool), name=name)
