"""The reference's CPU path, restated in float64 numpy: one call of
``Optimizer.step`` (PythonUtils/Optimizer.py:135-288).

TEST AND BASELINE INFRASTRUCTURE ONLY.  Only ``tests/`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
does.  It is timed as the reference CPU path the GPU step replaces, so it keeps
that path's algorithm and costs:

* every landmark x camera x frame from its anchor to the last frame gives two
  residual rows (Optimizer.py:170-229), weight 0 when the observed pixel lies
  outside the image (Optimizer.py:183-188) -- here also when the frame holds no
  observation of the landmark, which the caller marks the same way;
* a Cauchy weight with c = 1 on the norm of the homogeneous residual
  (Optimizer.py:195-199), no near-frame boost;
* a dense pose Jacobian J_T (rows x 6n) whose 2x6 block for every state i in
  [anchor, frame) is built by the chain of Optimizer.py:669-742, which
  recomputes all global poses for every block (Optimizer.py:709);
* the anchor-left inverse-depth Jacobian set to zero (Optimizer.py:221-222);
* dense products A, B, C, g (Optimizer.py:231-237), C + 1e-3 I twice
  (Optimizer.py:244, 251), H = A - B C^-1 B^T + I (Optimizer.py:256-263),
  np.linalg.solve (Optimizer.py:264);
* the adaptive step size on both poses and depths (Optimizer.py:267-281) and
  the inverse-depth clamp [1/20, 1/0.01] (Optimizer.py:32-35, 662-667);
* X <- X Exp(delta) with Exp(psi, rho) = [Rodrigues(psi), Jl(psi) rho]
  (LieUtils.py:22-25, 101-109).

The per-block arithmetic follows the reference file:line given at each
function; the code is this build's own.
"""
from __future__ import annotations

import numpy as np

from .ba_oracle import exp_se3, invert_pose, skew


def map_value_to_index(v: int, x: int, n: int) -> int:
    """Optimizer.py:12-18."""
    vc = max(0, min(v, x))
    return min(int((vc / x) * n), n - 1)


def _proj_jac(b: np.ndarray) -> np.ndarray:
    """d(b / b_z)/db for a 3-vector (Optimizer.py:776-792)."""
    return np.array([[1.0 / b[2], 0.0, -b[0] / (b[2] * b[2])],
                     [0.0, 1.0 / b[2], -b[1] / (b[2] * b[2])],
                     [0.0, 0.0, 0.0]])


class CpuPath:
    """State and one LM-like step of the reference CPU path.

    n incremental poses (n + 1 frames), m landmarks; ``anchor[l]`` is the host
    frame of landmark l; ``obs[(frame, l, right)]`` the homogeneous pixel
    (3,) of landmark l in that camera, absent where there is none.
    """

    def __init__(self, n, m, Kl, Kr, T_r_l, width, height, anchor, obs, poses=None, alpha=None):
        self.n, self.m = int(n), int(m)
        self.Kl, self.Kr = np.asarray(Kl, float), np.asarray(Kr, float)
        self.Kl_inv = np.linalg.inv(self.Kl)
        self.T_r_l = np.asarray(T_r_l, float)
        self.width, self.height = width, height
        self.anchor = [int(a) for a in anchor]
        self.obs = obs
        self.poses = [np.eye(4) for _ in range(self.n)] if poses is None else [np.array(P, float) for P in poses]
        self.alpha = np.ones(self.m) if alpha is None else np.array(alpha, float)
        self.min_alpha, self.max_alpha = 1.0 / 20.0, 1.0 / 0.01     # Optimizer.py:32-35
        self.step_size = 1.0
        self.prev_norm = 1e10

    # Optimizer.py:120-132
    def global_poses(self):
        G = [np.eye(4)]
        for X in self.poses:
            G.append(G[-1] @ np.linalg.inv(X))
        return G

    # StereoSetup.py:36-73
    def project(self, T_cam_in_global, p_global, right):
        T = T_cam_in_global @ self.T_r_l if right else T_cam_in_global
        pc = (np.linalg.inv(T) @ np.append(p_global, 1.0))[:3]
        b = (self.Kr if right else self.Kl) @ pc
        return b / b[2]

    def _observation(self, f, l, right):
        o = self.obs.get((f, l, right))
        return np.array([-1.0, -1.0, 1.0]) if o is None else np.asarray(o, float)

    # Optimizer.py:669-742: d(pixel)/d(xi_i), global poses recomputed per block
    def d_pixel_d_xi(self, pa, alpha, a, f, i, right):
        if not (a <= i < f):
            return np.zeros((3, 6))
        G = self.global_poses()
        Gn = G[f] @ self.T_r_l if right else G[f]
        T_a_i = np.linalg.inv(G[i]) @ G[a]
        T_a_n = np.linalg.inv(Gn) @ G[a]
        T_i_n = np.linalg.inv(Gn) @ G[i]
        pa_c = np.append(self.Kl_inv @ pa / alpha, 1.0)
        p_i = (T_a_i @ pa_c)[:3]
        p_n = (T_a_n @ pa_c)[:3]
        K = self.Kr if right else self.Kl
        D = np.zeros((3, 6))
        D[:, :3] = -T_i_n[:3, :3] @ skew(p_i)
        D[:, 3:] = T_i_n[:3, :3]
        return _proj_jac(K @ p_n) @ K @ D

    # Optimizer.py:745-774
    def d_pixel_d_alpha(self, pa, alpha, G, a, f, right):
        Gn = G[f] @ self.T_r_l if right else G[f]
        T = np.linalg.inv(Gn) @ G[a]
        pa_c = self.Kl_inv @ pa / alpha
        pn = T[:3, :3] @ pa_c + T[:3, 3]
        K = self.Kr if right else self.Kl
        return K @ _proj_jac(pn) @ T[:3, :3] @ (-pa_c / alpha)

    def step(self):
        """One call of Optimizer.step (Optimizer.py:135-288); returns the
        weighted residual norm."""
        n, m = self.n, self.m
        rows = m * (2 * n + 2) * 2
        J_T = np.zeros((rows, 6 * n))
        J_a = np.zeros((rows, m))
        r = np.zeros(rows)
        G = self.global_poses()
        row = 0
        for l in range(m):
            a = self.anchor[l]
            pa = self._observation(a, l, False)
            al = self.alpha[l]
            p_glob = (G[a] @ np.append(self.Kl_inv @ pa / al, 1.0))[:3]
            for right in (False, True):
                for f in range(a, n + 1):
                    o = self._observation(f, l, right)
                    res = o - self.project(G[f], p_glob, right)
                    inside = not (o[0] < 0 or o[1] < 0 or o[0] > self.width or o[1] > self.height)
                    w = 1.0 if inside else 0.0
                    sw = np.sqrt(1.0 / (1.0 + np.linalg.norm(res) ** 2))     # Cauchy, c = 1
                    for i in range(a, f):
                        J_T[row:row + 2, 6 * i:6 * i + 6] = w * sw * self.d_pixel_d_xi(pa, al, a, f, i, right)[:2]
                    da = self.d_pixel_d_alpha(pa, al, G, a, f, right)
                    if not right and a == f:
                        da = np.zeros(3)
                    J_a[row:row + 2, l] = w * sw * da[:2]
                    r[row:row + 2] = w * sw * res[:2]
                    row += 2
        A = J_T.T @ J_T
        B = J_T.T @ J_a
        C = J_a.T @ J_a
        g_T = J_T.T @ r
        g_a = J_a.T @ r
        C = C + 1e-3 * np.eye(m)
        C = C + 1e-3 * np.eye(m)
        C_inv = np.diag(1.0 / np.diagonal(C))
        BCi = B @ C_inv
        H = A - BCi @ B.T + np.eye(6 * n)
        d_pose = np.linalg.solve(H, g_T - BCi @ g_a)
        d_alpha = C_inv @ (g_a - B.T @ d_pose)
        r_norm = float(np.linalg.norm(r))
        if self.prev_norm > r_norm:
            self.step_size *= 1.01
        elif self.prev_norm < r_norm:
            self.step_size *= 0.9
        self.step_size = min(max(self.step_size, 0.02), 10.0)
        self.prev_norm = r_norm
        d_pose = self.step_size * d_pose
        d_alpha = self.step_size * d_alpha
        for k in range(n):
            self.poses[k] = self.poses[k] @ exp_se3(d_pose[6 * k:6 * k + 6])
        self.alpha = np.clip(self.alpha + d_alpha, self.min_alpha, self.max_alpha)
        self._last = (J_T, J_a, r, H, d_pose / self.step_size)
        return r_norm


def from_problem(p, poses=None, alpha=None) -> CpuPath:
    """The reference CPU path on a ``spgo.problems.Problem``: the problem's
    calibration, anchors and stereo pixels (frames without an observation of a
    landmark count as outside the image), its initial state by default."""
    intr = np.asarray(p.intrinsics, float)
    Kl = np.array([[intr[0], 0, intr[2]], [0, intr[1], intr[3]], [0, 0, 1.0]])
    Kr = np.array([[intr[4], 0, intr[6]], [0, intr[5], intr[7]], [0, 0, 1.0]])
    T_r_l = np.asarray(p.T_r_to_l, float).reshape(4, 4)
    m = int(p.n_landmarks)
    anchor = np.zeros(m, int)
    anchor[p.meas_l] = p.meas_a
    obs = {}
    uv = np.asarray(p.uv, float)
    for a, t, l, q in zip(p.meas_a, p.meas_t, p.meas_l, uv):
        obs[(int(t), int(l), False)] = np.array([q[0], q[1], 1.0])
        obs[(int(t), int(l), True)] = np.array([q[2], q[3], 1.0])
    X = p.X0 if poses is None else poses
    al = p.alpha0[:m] if alpha is None else alpha
    return CpuPath(p.n_poses, m, Kl, Kr, T_r_l, 640, 480, anchor, obs, poses=[np.asarray(x, float) for x in X],
                   alpha=np.asarray(al, float))


__all__ = ["CpuPath", "from_problem", "map_value_to_index", "invert_pose"]
