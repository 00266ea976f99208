"""Pins the restatement of the reference CPU path (oracle/cpu_path.py,
PythonUtils/Optimizer.py:135-288) that bench.py times as the CPU baseline:
its Jacobian blocks against central finite differences of its own projection,
zero residual and zero step at the true state without noise, the step against
a dense solve of the full damped normal equations (the Schur identity), and a
falling residual over repeated steps."""
import numpy as np
import pytest

from oracle import ba_oracle as O
from oracle import cpu_path as CP
from spgo.problems import make_problem


def _small(noise=1.0, seed=5):
    return make_problem(3, 12, 3, seed=seed, noise_px=noise)


def test_pose_and_depth_jacobians_match_finite_differences():
    p = _small()
    c = CP.from_problem(p)
    h = 1e-6
    checked = 0
    for l in range(c.m):
        a = c.anchor[l]
        pa = c._observation(a, l, False)
        al = c.alpha[l]
        for right in (False, True):
            for f in range(a + 1, c.n + 1):
                def pixel(poses, alpha):
                    saved = c.poses
                    c.poses = poses
                    G = c.global_poses()
                    c.poses = saved
                    pg = (G[a] @ np.append(c.Kl_inv @ pa / alpha, 1.0))[:3]
                    return c.project(G[f], pg, right)
                for i in range(a, f):
                    J = c.d_pixel_d_xi(pa, al, a, f, i, right)
                    for j in range(6):
                        e = np.zeros(6)
                        e[j] = h
                        Pp = [X.copy() for X in c.poses]
                        Pm = [X.copy() for X in c.poses]
                        Pp[i] = Pp[i] @ O.exp_se3(e)
                        Pm[i] = Pm[i] @ O.exp_se3(-e)
                        fd = (pixel(Pp, al) - pixel(Pm, al)) / (2 * h)
                        assert np.max(np.abs(fd - J[:, j])) < 1e-5 * max(1.0, np.max(np.abs(J)))
                        checked += 1
                da = c.d_pixel_d_alpha(pa, al, c.global_poses(), a, f, right)
                fd = (pixel(c.poses, al + h) - pixel(c.poses, al - h)) / (2 * h)
                assert np.max(np.abs(fd - da)) < 1e-5 * max(1.0, np.max(np.abs(da)))
    assert checked > 100


def test_zero_residual_and_zero_step_at_the_true_state():
    p = _small(noise=0.0)
    c = CP.from_problem(p, poses=p.X_true, alpha=p.alpha_true[:p.n_landmarks])
    X0 = [X.copy() for X in c.poses]
    rn = c.step()
    # the pixels are stored in float32: residuals of ~1e-5 px remain
    assert rn < 1e-3
    assert max(np.max(np.abs(X - Y)) for X, Y in zip(c.poses, X0)) < 1e-6


def test_step_solves_the_full_damped_normal_equations():
    p = _small(seed=6)
    c = CP.from_problem(p)
    c.step()
    J_T, J_a, r, H, d_pose = c._last
    # the full system with the reference's damping: A + I, C + 2e-3 I
    n6, m = J_T.shape[1], J_a.shape[1]
    J = np.concatenate([J_T, J_a], 1)
    K = J.T @ J
    K[:n6, :n6] += np.eye(n6)
    K[n6:, n6:] += 2e-3 * np.eye(m)
    full = np.linalg.solve(K, J.T @ r)
    assert np.allclose(full[:n6], d_pose, rtol=1e-8, atol=1e-10 * np.max(np.abs(full)))


def test_residual_falls_over_steps():
    p = make_problem(4, 30, 3, seed=8, noise_px=0.0)
    c = CP.from_problem(p)
    norms = [c.step() for _ in range(6)]
    assert norms[-1] < 0.5 * norms[0]
    assert c.step_size == pytest.approx(1.01 ** 6, rel=1e-12) or c.step_size < 1.01 ** 6
