"""Pins the oracle's Lie-group functions against independent library routines
and the invariants the reference's own unit tests check
(tests/testExpSO3.cpp:27-83, tests/testJacobiansSO3.cpp:95-195,
tests/testSE3.cpp:175-232)."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.spatial.transform import Rotation

from oracle import ba_oracle as O


def _rand_rotvecs(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    ang = rng.uniform(0, np.pi, size=(n, 1))
    return v * ang


def test_exp_so3_matches_scipy_rotvec():
    rng = np.random.default_rng(0)
    psi = _rand_rotvecs(rng, 500)
    # special cases of tests/testExpSO3.cpp:41-54: norm pi, > pi, 1
    extra = np.array([[np.pi, 0, 0], [0, 0, np.pi], [2.0, 2.0, 2.0], [1.0, 0, 0], [0.3, -0.4, 0.5]])
    psi = np.concatenate([psi, extra])
    R = O.exp_so3(psi)
    Rs = Rotation.from_rotvec(psi).as_matrix()
    assert np.max(np.abs(R - Rs)) < 1e-12


def test_exp_so3_small_angle_branch():
    # below 1e-4 the reference switches to I + S + S^2/2 (SO3.cu:21-25)
    for mag in [0.0, 1e-10, 1e-9, 1e-8, 5e-5, 9.9e-5]:
        psi = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8]) * mag
        R = O.exp_so3(psi)
        S = O.skew(psi)
        assert np.allclose(R, np.eye(3) + S + 0.5 * S @ S, atol=0, rtol=0) or mag == 0.0
        # truncation error O(theta^3) against the exact rotation
        assert np.max(np.abs(R - Rotation.from_rotvec(psi).as_matrix())) < max(mag ** 3, 1e-15)


def test_rotation_is_orthonormal():
    rng = np.random.default_rng(1)
    R = O.exp_so3(_rand_rotvecs(rng, 200) * 3.0)
    I = np.swapaxes(R, -1, -2) @ R
    assert np.max(np.abs(I - np.eye(3))) < 1e-13
    assert np.max(np.abs(np.linalg.det(R) - 1.0)) < 1e-13


def test_exp_se3_matches_matrix_exponential():
    rng = np.random.default_rng(2)
    for _ in range(100):
        xi = np.concatenate([_rand_rotvecs(rng, 1)[0], rng.normal(size=3)])
        W = np.zeros((4, 4))
        W[:3, :3] = O.skew(xi[:3])
        W[:3, 3] = xi[3:]
        assert np.max(np.abs(O.exp_se3(xi) - sla.expm(W))) < 1e-12


def test_left_jacobian_finite_difference():
    # Exp(psi + h e_k) Exp(psi)^T = Exp(Jl(psi) h e_k) + O(h^2); the reference
    # checks analytic Jl against FD with eps 1e-4 at 5e-5 (testJacobiansSO3.cpp:8,157)
    rng = np.random.default_rng(3)
    h = 1e-6
    for psi in list(_rand_rotvecs(rng, 30)) + [np.array([5e-5, 0, 0])]:
        Jl = O.left_jacobian_so3(psi)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            Rp = O.exp_so3(psi + e) @ O.exp_so3(psi).T
            Rm = O.exp_so3(psi - e) @ O.exp_so3(psi).T
            col = (Rotation.from_matrix(Rp).as_rotvec() - Rotation.from_matrix(Rm).as_rotvec()) / (2 * h)
            assert np.max(np.abs(col - Jl[:, k])) < 1e-7


def test_invert_pose_and_exp_log_roundtrip():
    rng = np.random.default_rng(4)
    xi = np.concatenate([_rand_rotvecs(rng, 50), rng.normal(size=(50, 3))], axis=1)
    T = O.exp_se3(xi)
    assert np.max(np.abs(T @ O.invert_pose(T) - np.eye(4))) < 1e-13
    # Exp -> Log -> Exp reconstruction (tests/testSE3.cpp:105 uses 1e-3)
    psi = Rotation.from_matrix(T[:, :3, :3]).as_rotvec()
    rho = np.linalg.solve(O.left_jacobian_so3(psi), T[:, :3, 3, None])[..., 0]
    T2 = O.exp_se3(np.concatenate([psi, rho], axis=1))
    assert np.max(np.abs(T2 - T)) < 1e-12


def test_global_pose_chain():
    rng = np.random.default_rng(5)
    X = O.exp_se3(np.concatenate([_rand_rotvecs(rng, 7) * 0.3, rng.normal(size=(7, 3))], axis=1))
    T_rl = O.exp_se3(np.array([0.01, 0.02, 0.0, 0.12, 0.0, 0.0]))
    G, GR = O.global_poses(X, T_rl)
    assert np.allclose(G[0], np.eye(4))
    for k in range(7):
        # X_k maps frame k into frame k+1: G_{k+1}^-1 G_k = X_k
        assert np.max(np.abs(np.linalg.inv(G[k + 1]) @ G[k] - X[k])) < 1e-12
        assert np.max(np.abs(GR[k] - G[k] @ T_rl)) < 1e-14


@pytest.mark.parametrize("kind,param", [("cauchy", 9.0), ("huber", 2.0), ("none", 0.0)])
def test_robust_derivative(kind, param):
    opts = O.Options(robust=kind, robust_param=param)
    s = np.linspace(0.01, 40.0, 400)
    rho, drho = O.robust(s, opts)
    h = 1e-6
    fd = (O.robust(s + h, opts)[0] - O.robust(s - h, opts)[0]) / (2 * h)
    mask = np.abs(np.sqrt(s) - param) > 1e-3 if kind == "huber" else np.ones_like(s, bool)
    assert np.max(np.abs(fd[mask] - drho[mask])) < 1e-7
    if kind == "cauchy":
        # the reference weight sqrt(1/(1+s/9)) (Solver_cuda.cu:222) is sqrt(rho')
        assert np.allclose(np.sqrt(drho), np.sqrt(1.0 / (1.0 + s / 9.0)), rtol=1e-15)
