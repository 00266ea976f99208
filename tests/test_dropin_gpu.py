"""The drop-in driven the way the reference's own harness drives its solver,
and the tools on either side of the step, on the GPU.

* ``build_ip.Solver.CudaSolver`` fed exactly as CUDAmanager.py:26-93,126 feeds
  ``Solver.CudaSolver``: one ``writeObservations`` call per measurement in the
  harness's enumeration, identity ``writeIncrementalPose`` for every frame,
  the constructor's inverse depths (1.0), then ``step(2)`` twice.
* A step started from the stereo-disparity inverse depths
  (``spgo_init_inverse_depths_from_disparity``; PythonUtils/Optimizer.py:87-114).
* The accuracy harness (CUDAmanager.py:100-109) on a GPU result.

Bar: 1e-9 relative on costs, poses and depths for the float64 state; the
float32 getters of the drop-in return that state rounded to float32.
"""
import numpy as np
import pytest

from build_ip import Solver as DropIn
from oracle import ba_oracle as O
from spgo import accuracy
from spgo.problems import make_manager_problem, make_problem
from spgo.solver import Solver, make_options, solver_from_problem

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _feed_like_cudamanager(p):
    """CUDAmanager.py:29, 69-93: constructor, calibration, one call per
    measurement, identity incremental poses."""
    s = DropIn.CudaSolver(p.n_poses, p.m_per_frame)
    s.loadCalibration(np.asarray(p.intrinsics, np.float32).reshape(8),
                      np.asarray(p.T_r_to_l, np.float32).reshape(16))
    for a, t, l, uv in zip(p.meas_a, p.meas_t, p.meas_l, p.uv):
        left = np.array([uv[0], uv[1], 1.0], np.float32)
        right = np.array([uv[2], uv[3], 1.0], np.float32)
        s.writeObservations(int(a), int(t), int(l), left, right)
    for k in range(p.n_poses):
        s.writeIncrementalPose(k, np.eye(4).reshape(16).astype(np.float32))
    return s


@pytest.mark.parametrize("n,m", [(8, 40), (12, 30)])
def test_cudasolver_driven_like_the_reference_harness(n, m):
    p = make_manager_problem(n, m, seed=100 + n)
    print(f"[dropin] n={n} m={m} measurements={p.n_obs}", flush=True)
    s = _feed_like_cudamanager(p)
    pa = O.from_problem(p)
    oo = O.Options(mode="reference")
    X = np.tile(np.eye(4), (n, 1, 1))
    al = np.ones(n * m)
    for call in range(2):
        X, al, _, c0, recs = O.lm_solve(pa, X, al, 2, oo)
        s.step(2)                                   # CUDAmanager.py:126
        g0, grecs = s.solver.stats()
        assert abs(g0 - c0) <= RTOL * c0, (call, g0, c0)
        for r, q in zip(recs, grecs):
            assert q["accepted"] and q["chol_ok"] == r.chol_ok
            assert abs(q["cost"] - r.cost) <= RTOL * r.cost, (call, q, r)
        gX, gal = s.solver.get_state()
        assert np.max(np.abs(gX[:, :3, :] - X[:, :3, :])) <= RTOL * max(1.0, np.max(np.abs(X)))
        used = np.unique(pa.l)
        assert np.max(np.abs(gal[used] - al[used]) / al[used]) <= RTOL
        # the reference getters: float32 copies of that state
        for k in range(n):
            T = s.getIncrementalPose(k)
            assert T.dtype == np.float32 and T.shape == (4, 4)
            assert np.array_equal(T[:3], gX[k, :3].astype(np.float32))
        d = s.getInverseDepths()
        assert d.dtype == np.float32 and np.array_equal(d, gal.astype(np.float32))
    # the harness's error measure (CUDAmanager.py:100-109) falls during the solve
    err = accuracy.pose_errors(np.stack([s.getIncrementalPose(k).astype(np.float64) for k in range(n)]),
                               p.X_true)["xi"]
    err0 = accuracy.pose_errors(np.tile(np.eye(4), (n, 1, 1)), p.X_true)["xi"]
    assert np.mean(err) < 0.2 * np.mean(err0)


def test_step_from_disparity_depths_matches_oracle():
    p = make_problem(10, 600, 4, seed=55, alpha_perturb=0.5)
    s = solver_from_problem(p, make_options(mode="lm"))
    n_set = s.init_inverse_depths_from_disparity()
    X0, a0 = s.get_state()
    # the library's depths are the oracle's triangulation of the host pixels
    host = p.meas_t == p.meas_a
    l = p.meas_l[host]
    uv = p.uv[host].astype(np.float64)
    one = np.ones(l.size)
    d = O.depth_from_disparity(p.intrinsics, p.T_r_to_l, np.stack([uv[:, 0], uv[:, 1], one], 1),
                               np.stack([uv[:, 2], uv[:, 3], one], 1))
    ok = d > 0
    assert n_set == int(ok.sum()) and n_set > 0.9 * p.n_landmarks
    assert np.max(np.abs(a0[l[ok]] * d[ok] - 1.0)) < 1e-12
    pa = O.from_problem(p)
    X, al, _, c0, recs = O.lm_solve(pa, X0, a0, 6, O.Options())
    s.step(6)
    g0, grecs = s.stats()
    print(f"[disparity] set={n_set} c0={c0:.6g} final={recs[-1].cost:.6g}", flush=True)
    assert abs(g0 - c0) <= RTOL * c0
    for r, q in zip(recs, grecs):
        assert q["accepted"] == r.accepted
        assert abs(q["cost"] - r.cost) <= RTOL * r.cost
    gX, gal = s.get_state()
    assert np.max(np.abs(gX[:, :3] - X[:, :3])) <= RTOL * max(1.0, np.max(np.abs(X)))
    used = np.unique(p.meas_l)
    assert np.max(np.abs(gal[used] - al[used]) / al[used]) <= RTOL
    # triangulated depths start closer to the truth than the perturbed ones
    e_tri = np.abs(a0[used] - p.alpha_true[used]) / p.alpha_true[used]
    e_per = np.abs(p.alpha0[used] - p.alpha_true[used]) / p.alpha_true[used]
    assert np.mean(e_tri) < np.mean(e_per)


def test_accuracy_harness_on_a_gpu_result():
    p = make_problem(12, 900, 5, seed=66)
    s = solver_from_problem(p, make_options(mode="lm"))
    s.step(8)
    X, al = s.get_state()
    e = accuracy.pose_errors(X, p.X_true)
    eo = O.pose_errors(p.X_true, X)                 # rotation angle, translation norm
    assert np.max(np.abs(e["rot"] - eo[:, 0])) < 1e-9
    assert np.max(np.abs(e["trans"] - eo[:, 1])) < 1e-9
    # the solve moved the poses towards the truth
    e0 = accuracy.pose_errors(O.project_rotations(p.X0.astype(np.float64)), p.X_true)
    assert np.mean(e["xi"]) < 0.5 * np.mean(e0["xi"])
    summ = accuracy.summary(X, p.X_true, al, p.alpha_true, np.unique(p.meas_l))
    assert summ["pose_xi_max"] == pytest.approx(float(np.max(e["xi"])), rel=1e-15)
