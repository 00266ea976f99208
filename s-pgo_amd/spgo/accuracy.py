"""Estimation accuracy against ground truth.

``pose_errors`` is the per-pose error of CUDAmanager.compute_estimation_errors
(CUDAmanager.py:100-109), ||Log(X_true^-1 X_est)|| with xi = (psi, rho), split
into its rotation angle and translation part; ``depth_errors`` the relative
inverse-depth errors.  Host-side numpy, for reports and convergence studies.
"""
from __future__ import annotations

import numpy as np


def _skew(w: np.ndarray) -> np.ndarray:
    S = np.zeros(w.shape[:-1] + (3, 3))
    S[..., 0, 1], S[..., 0, 2] = -w[..., 2], w[..., 1]
    S[..., 1, 0], S[..., 1, 2] = w[..., 2], -w[..., 0]
    S[..., 2, 0], S[..., 2, 1] = -w[..., 1], w[..., 0]
    return S


def log_so3(R: np.ndarray) -> np.ndarray:
    """Rotation vectors of rotation matrices (..., 3, 3)."""
    R = np.asarray(R, dtype=np.float64)
    c = np.clip((np.trace(R, axis1=-2, axis2=-1) - 1.0) / 2.0, -1.0, 1.0)
    th = np.arccos(c)
    v = np.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0], R[..., 1, 0] - R[..., 0, 1]], -1)
    s = np.sin(th)
    small = th < 1e-6
    scale = np.where(small, 0.5 + th ** 2 / 12.0, th / (2.0 * np.where(small, 1.0, s)))
    w = v * scale[..., None]
    # near pi the antisymmetric part vanishes: take the axis from R + I
    near_pi = np.pi - th < 1e-4
    if np.any(near_pi):
        for idx in zip(*np.nonzero(near_pi)) if R.ndim > 2 else [()]:
            B = (R[idx] + np.eye(3)) / 2.0
            k = int(np.argmax(np.diag(B)))
            axis = B[:, k] / np.sqrt(max(B[k, k], 1e-300))
            axis *= np.sign(np.dot(axis, v[idx]) if np.any(v[idx]) else 1.0) or 1.0
            w[idx] = axis * th[idx]
    return w


def _left_jacobian_inv(psi: np.ndarray) -> np.ndarray:
    th = np.linalg.norm(psi, axis=-1)
    S = _skew(psi)
    small = th < 1e-6
    ths = np.where(small, 1.0, th)
    b = np.where(small, 1.0 / 12.0, (1.0 - ths * np.sin(ths) / (2.0 * (1.0 - np.cos(ths)))) / ths ** 2)
    return np.eye(3) - 0.5 * S + b[..., None, None] * (S @ S)


def log_se3(T: np.ndarray) -> np.ndarray:
    """Twists (psi, rho) of rigid transforms (..., 4, 4) or (..., 3, 4): Exp(psi, rho) = [Exp(psi), Jl(psi) rho]."""
    T = np.asarray(T, dtype=np.float64)
    psi = log_so3(T[..., :3, :3])
    rho = np.einsum('...ij,...j->...i', _left_jacobian_inv(psi), T[..., :3, 3])
    return np.concatenate([psi, rho], -1)


def _inv(T: np.ndarray) -> np.ndarray:
    Ti = np.zeros(T.shape[:-2] + (4, 4))
    R = T[..., :3, :3]
    Ti[..., :3, :3] = np.swapaxes(R, -1, -2)
    Ti[..., :3, 3] = -np.einsum('...ji,...j->...i', R, T[..., :3, 3])
    Ti[..., 3, 3] = 1.0
    return Ti


def pose_errors(X_est: np.ndarray, X_true: np.ndarray) -> dict:
    """Per pose: 'xi' = ||Log(X_true^-1 X_est)||, 'rot' = its rotation angle
    (rad), 'trans' = the translation norm of X_true^-1 X_est (m)."""
    Xe = np.asarray(X_est, dtype=np.float64)
    Xt = np.asarray(X_true, dtype=np.float64)
    E = _inv(Xt) @ np.concatenate([Xe[..., :3, :], np.broadcast_to([0, 0, 0, 1.0], Xe.shape[:-2] + (1, 4))], -2)
    xi = log_se3(E)
    return {"xi": np.linalg.norm(xi, axis=-1), "rot": np.linalg.norm(xi[..., :3], axis=-1),
            "trans": np.linalg.norm(E[..., :3, 3], axis=-1)}


def depth_errors(alpha_est: np.ndarray, alpha_true: np.ndarray, used: np.ndarray) -> np.ndarray:
    """|alpha_est - alpha_true| / alpha_true over the used landmark slots."""
    a = np.asarray(alpha_est, dtype=np.float64)[used]
    t = np.asarray(alpha_true, dtype=np.float64)[used]
    return np.abs(a - t) / t


def summary(X_est, X_true, alpha_est=None, alpha_true=None, used=None) -> dict:
    """Mean and max of the pose and inverse-depth errors."""
    e = pose_errors(X_est, X_true)
    out = {"pose_xi_mean": float(e["xi"].mean()), "pose_xi_max": float(e["xi"].max()),
           "rot_deg_mean": float(np.degrees(e["rot"]).mean()), "trans_m_mean": float(e["trans"].mean())}
    if alpha_est is not None:
        d = depth_errors(alpha_est, alpha_true, used)
        out.update({"inv_depth_rel_mean": float(d.mean()), "inv_depth_rel_max": float(d.max())})
    return out
