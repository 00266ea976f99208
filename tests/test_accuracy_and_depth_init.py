"""Host-side tools: the accuracy harness (CUDAmanager.py:100-109) and the
stereo-disparity depth initialisation (PythonUtils/Optimizer.py:87-114)."""
import numpy as np

from oracle import ba_oracle as O
from spgo import accuracy
from spgo.problems import make_problem
from spgo.solver import Solver


def test_log_se3_inverts_exp():
    rng = np.random.default_rng(3)
    xi = rng.normal(size=(50, 6)) * np.array([0.8, 0.8, 0.8, 2.0, 2.0, 2.0])
    xi[0] = 0.0
    xi[1, :3] = [1e-8, 0, 0]
    T = np.stack([O.exp_se3(x) for x in xi])
    assert np.max(np.abs(accuracy.log_se3(T) - xi)) < 1e-9


def test_pose_errors_of_a_known_perturbation():
    rng = np.random.default_rng(4)
    p = make_problem(6, 40, 3, seed=2)
    xi = rng.normal(size=(6, 6)) * 0.05
    X_est = np.stack([p.X_true[k] @ O.exp_se3(xi[k]) for k in range(6)])
    e = accuracy.pose_errors(X_est, p.X_true)
    assert np.allclose(e["xi"], np.linalg.norm(xi, axis=1), atol=1e-9)
    assert np.allclose(e["rot"], np.linalg.norm(xi[:, :3], axis=1), atol=1e-9)
    assert np.all(accuracy.pose_errors(p.X_true, p.X_true)["xi"] < 1e-12)


def test_disparity_depths_recover_true_depths_without_noise():
    p = make_problem(8, 300, 3, seed=21, noise_px=0.0)
    s = Solver(p.n_poses, p.m_per_frame)
    s.load_calibration(p.intrinsics, p.T_r_to_l)
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    assert s.init_inverse_depths_from_disparity() == p.n_landmarks
    a = s.get_inverse_depths()[:p.n_landmarks].astype(np.float64)
    # float32 pixels: about 1e-5 relative
    assert np.max(np.abs(a - p.alpha_true[:p.n_landmarks]) / p.alpha_true[:p.n_landmarks]) < 1e-4


def test_disparity_depths_match_the_formula_with_noise():
    p = make_problem(8, 300, 3, seed=22, noise_px=2.0)
    s = Solver(p.n_poses, p.m_per_frame)
    s.load_calibration(p.intrinsics, p.T_r_to_l)
    s.write_observations(p.meas_a, p.meas_t, p.meas_l, p.uv)
    before = s.get_inverse_depths()
    n = s.init_inverse_depths_from_disparity()
    host = p.meas_t == p.meas_a
    l = p.meas_l[host]
    uv = p.uv[host].astype(np.float64)
    ones = np.ones(l.size)
    d = O.depth_from_disparity(p.intrinsics, p.T_r_to_l, np.stack([uv[:, 0], uv[:, 1], ones], 1),
                               np.stack([uv[:, 2], uv[:, 3], ones], 1))
    a = s.get_inverse_depths().astype(np.float64)
    ok = d > 0
    assert n == int(ok.sum())
    assert np.max(np.abs(a[l[ok]] - 1.0 / d[ok]) * d[ok]) < 1e-5
    # landmarks with a non-positive depth and unused slots keep their values
    untouched = np.setdiff1d(np.arange(a.size), l[ok])
    assert np.array_equal(a[untouched], before[untouched].astype(np.float64))


def test_oracle_disparity_formula_is_exact_without_noise():
    # pins the oracle's restatement: noise-free stereo pixels give the true depth
    p = make_problem(8, 200, 3, seed=23, noise_px=0.0)
    host = p.meas_t == p.meas_a
    uv = p.uv[host].astype(np.float64)
    ones = np.ones(uv.shape[0])
    d = O.depth_from_disparity(p.intrinsics, p.T_r_to_l, np.stack([uv[:, 0], uv[:, 1], ones], 1),
                               np.stack([uv[:, 2], uv[:, 3], ones], 1))
    assert np.max(np.abs(d * p.alpha_true[p.meas_l[host]] - 1.0)) < 1e-4
