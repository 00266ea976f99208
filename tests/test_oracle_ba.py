"""Pins the oracle's bundle-adjustment step: Jacobians against central finite
differences (the idea of PythonUtils/DepthOptimization.py:97-150), the gradient
against a finite difference of the robust cost, the Schur solve against a
dense solve of the full normal equations, the reference's deterministic scene
(PythonUtils/utils.py:83-129, PythonUtils/Simulator.py:94-177) as a
known-answer test, and noise-free convergence to ground truth."""
import numpy as np
import pytest

from oracle import ba_oracle as O
from spgo.problems import make_problem


def tiny_problem(seed=7, noise=1.0, N=4, L=12, K=3, outliers=0.0):
    return make_problem(N, L, K, seed=seed, noise_px=noise, outlier_frac=outliers)


def _perturb_pose(X, i, j, h):
    Xp = X.copy()
    e = np.zeros(6)
    e[j] = h
    Xp[i] = X[i] @ O.exp_se3(e)
    return Xp


@pytest.mark.parametrize("robust", ["cauchy", "huber", "none"])
def test_jacobians_match_finite_differences(robust):
    p = tiny_problem()
    pa = O.from_problem(p)
    opts = O.Options(robust=robust, robust_param=9.0 if robust == "cauchy" else 2.0)
    X = p.X0.astype(np.float64)
    al = p.alpha0.astype(np.float64)
    cost, r, J_T, J_a = O.jacobians(pa, X, al, opts)
    e, w, _ = O.residuals(pa, X, al, opts)
    wrow = np.repeat(w, 2, axis=1).reshape(-1)          # weight per row 4m + 2c + xy
    JT = J_T.toarray() / wrow[:, None]
    Ja = J_a.toarray() / wrow[:, None]
    h = 1e-6
    # dz/dxi = -de/dxi, rows 4m + 2c + xy
    for i in range(p.n_poses):
        for j in range(6):
            ep = O.residuals(pa, _perturb_pose(X, i, j, h), al, opts)[0].reshape(-1)
            em = O.residuals(pa, _perturb_pose(X, i, j, -h), al, opts)[0].reshape(-1)
            fd = -(ep - em) / (2 * h)
            col = JT[:, 6 * i + j]
            assert np.max(np.abs(fd - col)) < 1e-5 * max(1.0, np.max(np.abs(col)))
    for l in np.unique(pa.l):
        ap, am = al.copy(), al.copy()
        ap[l] += h
        am[l] -= h
        fd = -(O.residuals(pa, X, ap, opts)[0] - O.residuals(pa, X, am, opts)[0]).reshape(-1) / (2 * h)
        col = Ja[:, l]
        assert np.max(np.abs(fd - col)) < 1e-5 * max(1.0, np.max(np.abs(col)))


def test_pose_jacobian_zero_outside_anchor_target_interval():
    # states i outside [a, t) get no Jacobian block (Solver_cuda.cu:248)
    p = tiny_problem()
    pa = O.from_problem(p)
    _, _, J_T, _ = O.jacobians(pa, p.X0.astype(np.float64), p.alpha0.astype(np.float64), O.Options())
    JT = J_T.toarray()
    for m in range(pa.a.size):
        for i in range(p.n_poses):
            blk = JT[4 * m: 4 * m + 4, 6 * i: 6 * i + 6]
            if not (pa.a[m] <= i < pa.t[m]):
                assert not blk.any()


@pytest.mark.parametrize("robust", ["cauchy", "huber"])
def test_gradient_is_minus_cost_gradient(robust):
    # g = J^T W r is minus the gradient of the robust cost 1/2 sum kappa rho(|e|^2):
    # checks residual sign, weights and Jacobians together.
    p = tiny_problem(seed=11, outliers=0.3)
    pa = O.from_problem(p)
    opts = O.Options(robust=robust, robust_param=9.0 if robust == "cauchy" else 2.0)
    X = p.X0.astype(np.float64)
    al = p.alpha0.astype(np.float64)
    lin = O.linearize(pa, X, al, opts)
    h = 1e-6
    for i in range(p.n_poses):
        for j in range(6):
            fd = (O.cost_only(pa, _perturb_pose(X, i, j, h), al, opts)
                  - O.cost_only(pa, _perturb_pose(X, i, j, -h), al, opts)) / (2 * h)
            assert abs(fd + lin.g_T[6 * i + j]) < 1e-5 * max(1.0, abs(fd))
    for l in np.unique(pa.l):
        ap, am = al.copy(), al.copy()
        ap[l] += h
        am[l] -= h
        fd = (O.cost_only(pa, X, ap, opts) - O.cost_only(pa, X, am, opts)) / (2 * h)
        assert abs(fd + lin.g_a[l]) < 1e-5 * max(1.0, abs(fd))


def test_schur_solve_equals_full_normal_equations():
    p = tiny_problem(seed=3, N=5, L=20, K=4)
    pa = O.from_problem(p)
    opts = O.Options()
    lin = O.linearize(pa, p.X0.astype(np.float64), p.alpha0.astype(np.float64), opts)
    lam = 0.37
    H, g = O.reduced_system(lin, lam, opts.eps)
    delta, ok = O.solve_pose_step(H, g)
    assert ok
    dal = O.back_substitute(lin, delta, opts.eps)
    used = np.unique(pa.l)
    P = 6 * p.n_poses
    Bd = lin.B.toarray()[:, used]
    full = np.block([[lin.A + lam * np.eye(P), Bd],
                     [Bd.T, np.diag(lin.C[used] + opts.eps)]])
    rhs = np.concatenate([lin.g_T, lin.g_a[used]])
    sol = np.linalg.solve(full, rhs)
    assert np.max(np.abs(sol[:P] - delta)) < 1e-9 * max(1.0, np.max(np.abs(delta)))
    assert np.max(np.abs(sol[P:] - dal[used])) < 1e-9 * max(1.0, np.max(np.abs(dal)))
    # the dense J route (the reference's SGEMMs) gives the same blocks
    JT = lin.J_T.toarray()
    Ja = lin.J_a.toarray()
    assert np.allclose(JT.T @ JT, lin.A, rtol=1e-12, atol=1e-9)
    assert np.allclose(JT.T @ Ja, lin.B.toarray(), rtol=1e-12, atol=1e-9)
    assert np.allclose(np.diag(Ja.T @ Ja), lin.C, rtol=1e-12, atol=1e-12)


def _deterministic_scene():
    """Closed-form poses of SamplePoses.create_samples_deterministic
    (PythonUtils/utils.py:83-129) and the single landmark of
    DeterministicSimulator (Simulator.py:94-125): left pixel (320, 100) at depth 5
    in frame 0, with the reference stereo rig (StereoSetup.py:9-33)."""
    poses = []
    for r, theta, phi in [(2.5, 0.0, 0.0), (2.5, 0.0, np.pi / 2), (3.5, 0.0, -np.pi / 2)]:
        pos = np.array([r * np.sin(phi) * np.cos(theta), r * np.sin(phi) * np.sin(theta), r * np.cos(phi)])
        fwd = -pos / np.linalg.norm(pos)
        up = np.array([0.0, 0.0, 1.0])
        if abs(fwd @ up) > 0.99:
            up = np.array([0.0, 1.0, 0.0])
        right = np.cross(up, fwd)
        right /= np.linalg.norm(right)     # the reference adds 1e-8 inside the norm; dropped so R is orthonormal
        upv = np.cross(fwd, right)
        T = np.eye(4)
        T[:3, :3] = np.stack([right, upv, fwd], axis=1)
        T[:3, 3] = pos
        poses.append(T)
    K = np.array([[320.0, 0, 320.0], [0, 320.0, 240.0], [0, 0, 1]])
    T_lr = np.eye(4)
    T_lr[0, 3] = -0.2
    T_rl = np.linalg.inv(T_lr)
    p_a = np.array([320.0, 100.0, 1.0])
    t_feat_ca = np.linalg.inv(K) @ p_a * 5.0
    Pw = poses[0][:3, :3] @ t_feat_ca + poses[0][:3, 3]
    return poses, K, T_rl, Pw


def test_deterministic_scene_known_answer():
    poses, K, T_rl, Pw = _deterministic_scene()
    # frame 2 sees the point in front of both cameras; frame 1 has it behind
    obs = {}
    for f, G in enumerate(poses):
        uv = []
        for T in (G, G @ T_rl):
            Pc = np.linalg.inv(T) @ np.append(Pw, 1.0)
            uv.append((K @ (Pc[:3] / Pc[2]))[:2])
            uv.append(Pc[2])
        obs[f] = uv
    assert obs[0][1] > 0 and obs[2][1] > 0 and obs[2][3] > 0
    assert np.allclose(obs[0][0], [320.0, 100.0], atol=1e-12)       # the anchor pixel itself
    frames = [0, 2]
    a = np.zeros(2, np.int32)
    t = np.array(frames, np.int32)
    l = np.zeros(2, np.int32)
    uv = np.array([[*obs[f][0], *obs[f][2]] for f in frames], np.float64)
    X_true = np.stack([np.linalg.inv(poses[k + 1]) @ poses[k] for k in range(2)])
    pa = O.problem_arrays(2, 1, [320, 320, 320, 240, 320, 320, 320, 240], T_rl.reshape(16), a, t, l, uv)
    pa.z[:] = uv.reshape(-1, 2, 2)          # keep the exact float64 pixels
    pa.ray[:] = uv[0, 0:2]
    pa.T_rl[:] = T_rl                      # and the exact baseline (0.2 is not a float32)
    e, w, cost = O.residuals(pa, X_true, np.array([0.2]), O.Options())
    assert np.max(np.abs(e)) < 1e-9
    assert cost.sum() < 1e-15
    # right camera of the anchor frame: disparity fx * b / z = 320 * 0.2 / 5 = 12.8 px
    assert abs((uv[0, 0] - uv[0, 2]) - 12.8) < 1e-9
    # at the optimum the gradient vanishes
    lin = O.linearize(pa, X_true, np.array([0.2]), O.Options())
    assert np.max(np.abs(lin.g_T)) < 1e-8 and np.max(np.abs(lin.g_a)) < 1e-8


def test_noise_free_convergence_to_ground_truth():
    p = make_problem(6, 120, 4, seed=21, noise_px=0.0, rot_perturb=0.02, trans_perturb=0.03,
                     alpha_perturb=0.05)
    pa = O.from_problem(p)
    # float32 rounding of the pixels is the only noise left
    X, al, lam, c0, recs = O.lm_solve(pa, p.X0.astype(np.float64), p.alpha0.astype(np.float64),
                                      15, O.Options(robust="none"))
    err = O.pose_errors(p.X_true, X)
    assert recs[-1].cost < 1e-6 * c0
    assert err[:, 0].max() < 1e-5 and err[:, 1].max() < 1e-4
    L = p.n_landmarks
    assert np.max(np.abs(al[:L] - p.alpha_true[:L]) / p.alpha_true[:L]) < 1e-4


def test_lm_costs_monotone_and_rejects_raise_lambda():
    p = tiny_problem(seed=5, N=6, L=40, K=4, outliers=0.1)
    pa = O.from_problem(p)
    opts = O.Options(lambda0=1e-4)
    X, al, lam, c0, recs = O.lm_solve(pa, p.X0.astype(np.float64), p.alpha0.astype(np.float64), 12, opts)
    prev = c0
    lam_prev = None
    for r in recs:
        assert r.cost <= prev
        if r.accepted:
            assert r.cost_candidate == r.cost and r.cost < prev
        else:
            assert r.cost == prev and r.cost_candidate >= prev
        prev = r.cost
    for r0, r1 in zip(recs, recs[1:]):
        expect = r0.lam * (opts.lam_down if r0.accepted else opts.lam_up)
        expect = min(max(expect, opts.lam_min), opts.lam_max)
        assert r1.lam == pytest.approx(expect, rel=1e-15)


def test_observation_table_semantics():
    # later writes to a (frame, landmark) cell win; a missing anchor cell gives a
    # zero ray pixel (Solver.cpp:40-56, Solver_cuda.cu:147)
    intr = [500, 500, 320, 240, 500, 500, 320, 240]
    T = np.eye(4).reshape(16)
    a = np.array([0, 0, 1], np.int32)
    t = np.array([1, 1, 2], np.int32)
    l = np.array([0, 0, 1], np.int32)
    uv = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12]], np.float32)
    pa = O.problem_arrays(2, 2, intr, T, a, t, l, uv)
    assert np.array_equal(pa.z[0], [[5, 6], [7, 8]]) and np.array_equal(pa.z[1], [[5, 6], [7, 8]])
    assert np.array_equal(pa.ray, np.zeros((3, 2)))


# ---------------------------------------------------------------------------
# Known answers for three constants of the step that finite differences cannot
# pin: the near-frame weight, the Huber cost above its kink and the step size.
# ---------------------------------------------------------------------------
def _identity_scene(targets, shift_px, robust="cauchy", param=9.0, near_boost=True):
    """One landmark anchored in frame 0 with identity incremental poses: every
    frame is the identity, so the predicted left pixel equals the anchor ray's
    pixel and each target's residual is the given pixel shift exactly."""
    N = max(targets) + 1
    ray = np.array([300.0, 200.0])
    intr = np.array([500.0, 500.0, 320.0, 240.0, 500.0, 500.0, 320.0, 240.0], np.float32)
    T = np.eye(4, dtype=np.float32).reshape(16)        # right camera at the left one
    a = [0] + [0] * len(targets)
    t = [0] + list(targets)
    uv = [[ray[0], ray[1], ray[0], ray[1]]]
    for _ in targets:
        uv.append([ray[0] + shift_px[0], ray[1] + shift_px[1], ray[0], ray[1]])
    pa = O.problem_arrays(N, 1, intr, T, a, t, [0] * len(a), np.array(uv, np.float32))
    opts = O.Options(robust=robust, robust_param=param, near_boost=near_boost)
    X = np.tile(np.eye(4), (N, 1, 1))
    return pa, X, np.array([0.5]), opts


def test_near_frame_weight_known_answer():
    # Solver_cuda.cu:222-228: w = sqrt(1 / (1 + |e|^2 / 9)), times 4 when |t - a| < 2
    pa, X, al, opts = _identity_scene([1, 2], (3.0, 4.0))
    e, w, cost = O.residuals(pa, X, al, opts)
    assert np.allclose(e[1, 0], [3.0, 4.0], atol=1e-9) and np.allclose(e[2, 0], [3.0, 4.0], atol=1e-9)
    assert w[1, 0] == pytest.approx(4.0 / np.sqrt(1.0 + 25.0 / 9.0), rel=1e-12)   # |t - a| = 1
    assert w[2, 0] == pytest.approx(1.0 / np.sqrt(1.0 + 25.0 / 9.0), rel=1e-12)   # |t - a| = 2
    assert w[1, 1] == pytest.approx(4.0, rel=1e-12) and w[2, 1] == pytest.approx(1.0, rel=1e-12)
    # cost 1/2 kappa c^2 log(1 + s / c^2) with kappa = 16 and 1
    assert cost[1, 0] == pytest.approx(0.5 * 16.0 * 9.0 * np.log(34.0 / 9.0), rel=1e-12)
    assert cost[2, 0] == pytest.approx(0.5 * 9.0 * np.log(34.0 / 9.0), rel=1e-12)
    # without the boost both weights are the plain Cauchy weight
    pa, X, al, opts = _identity_scene([1, 2], (3.0, 4.0), near_boost=False)
    _, w, _ = O.residuals(pa, X, al, opts)
    assert w[1, 0] == pytest.approx(w[2, 0], rel=1e-15)


def test_huber_cost_known_answer_and_continuity():
    # delta = 2 px: rho(s) = s inside, 2 delta sqrt(s) - delta^2 outside
    pa, X, al, opts = _identity_scene([2], (3.0, 4.0), robust="huber", param=2.0)
    _, w, cost = O.residuals(pa, X, al, opts)
    assert cost[1, 0] == pytest.approx(0.5 * (2.0 * 2.0 * 5.0 - 4.0), rel=1e-12)    # |e| = 5 > delta
    assert w[1, 0] == pytest.approx(np.sqrt(2.0 / 5.0), rel=1e-12)
    # continuous at the kink |e| = delta, from both sides
    for h in (1e-7, -1e-7):
        pa, X, al, opts = _identity_scene([2], (0.0, 2.0 + h), robust="huber", param=2.0)
        c = O.residuals(pa, X, al, opts)[2][1, 0]
        assert c == pytest.approx(0.5 * 4.0, abs=3e-7)
    s = np.array([4.0 - 1e-12, 4.0, 4.0 + 1e-12])
    rho, _ = O.robust(s, O.Options(robust="huber", robust_param=2.0))
    assert np.max(np.abs(rho - 4.0)) < 1e-11


def test_step_size_scales_the_pose_update_only():
    # Solver_cuda.cu:545-551: the alpha update carries no step size (551 overwrites 549)
    rng = np.random.default_rng(9)
    X = np.stack([O.exp_se3(x) for x in rng.normal(size=(3, 6)) * 0.1])
    al = np.array([0.5, 0.25, 1.0])
    d = rng.normal(size=18) * 0.05
    dal = np.array([0.01, -0.02, 0.03])
    X1, a1 = O.retract(X, al, d, dal, O.Options(step_size=1.0))
    Xh, ah = O.retract(X, al, d, dal, O.Options(step_size=0.25))
    assert np.array_equal(a1, ah)
    assert np.allclose(ah, al + dal, rtol=0, atol=1e-15)
    for k in range(3):
        assert np.allclose(Xh[k], X[k] @ O.exp_se3(0.25 * d[6 * k:6 * k + 6]), atol=1e-14)
        assert np.allclose(X1[k], X[k] @ O.exp_se3(d[6 * k:6 * k + 6]), atol=1e-14)
