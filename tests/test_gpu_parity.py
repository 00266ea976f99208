"""Parity of the gfx950 FP64 path with the FP64 CPU oracle, through the C ABI.

Bar: float64 work, so per-iteration costs and the final state agree to 1e-9
relative (the tolerance of the FP64 plain path); accept/reject decisions and
Cholesky outcomes are identical.  Sizes are chosen so the oracle finishes in
seconds; the full 64-keyframe workload is checked through properties and the
oracle's first iterations.
"""
import numpy as np
import pytest

from oracle import ba_oracle as O
from spgo.problems import make_config, make_problem
from spgo.solver import Solver, make_options, solver_from_problem

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _opts_pair(mode="lm", robust="cauchy", near_boost=True, **kw):
    rp = 9.0 if robust == "cauchy" else (2.0 if robust == "huber" else 0.0)
    go = make_options(mode=mode, robust=robust, robust_param=rp if robust != "none" else None,
                      near_boost=near_boost, **kw)
    oo = O.Options(mode=mode, robust=robust, robust_param=rp, near_boost=near_boost)
    return go, oo


def _check_run(p_or_pa, s, iters, oo, X0, a0, tag):
    pa = p_or_pa
    X, al, lam, c0, recs = O.lm_solve(pa, X0, a0, iters, oo)
    s.step(iters)
    g0, grecs = s.stats()
    print(f"[{tag}] oracle c0={c0:.12g} gpu c0={g0:.12g}", flush=True)
    assert abs(g0 - c0) <= RTOL * abs(c0)
    assert len(grecs) == iters
    for i, (r, q) in enumerate(zip(recs, grecs)):
        assert q["accepted"] == r.accepted, (tag, i)
        assert q["chol_ok"] == r.chol_ok, (tag, i)
        assert abs(q["cost_candidate"] - r.cost_candidate) <= RTOL * abs(r.cost_candidate), (tag, i)
        assert abs(q["cost"] - r.cost) <= RTOL * abs(r.cost), (tag, i)
        assert q["lam"] == pytest.approx(r.lam, rel=1e-12, abs=0), (tag, i)
    gX, gal = s.get_state()
    assert np.max(np.abs(gX[:, :3, :] - X[:, :3, :])) <= RTOL * max(1.0, np.max(np.abs(X)))
    used = np.unique(pa.l)
    assert np.max(np.abs(gal[used] - al[used]) / np.abs(al[used])) <= RTOL
    return X, al, recs


@pytest.mark.parametrize("mode,factor", [("lm", 0), ("reference", 0), ("lm", 1), ("reference", 1),
                                         ("lm", 3), ("reference", 3)])
def test_c1_parity(mode, factor):
    # factor 0: automatic (the one-workgroup envelope Cholesky here), 1: dense
    # blocked Cholesky, 3: envelope Cholesky split at a separator
    p = make_config("c1_small")
    pa = O.from_problem(p)
    go, oo = _opts_pair(mode=mode, factorization=factor)
    s = solver_from_problem(p, go)
    _check_run(pa, s, 10, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"c1-{mode}")


def test_reduced_system_matches_oracle():
    p = make_problem(12, 900, 5, seed=77)
    pa = O.from_problem(p)
    go, oo = _opts_pair()
    s = solver_from_problem(p, go)
    lam = 0.25
    H, g, cost = s.linearize(lam)
    # the library projects uploaded rotations onto SO(3); so does the oracle's lm_solve
    X0 = O.project_rotations(p.X0.astype(np.float64))
    lin = O.linearize(pa, X0, p.alpha0.astype(np.float64), oo)
    Ho, go_ = O.reduced_system(lin, lam, oo.eps)
    print(f"[reduced] |H|={np.max(np.abs(Ho)):.4g} err={np.max(np.abs(H - Ho)):.3g} "
          f"|g|={np.max(np.abs(go_)):.4g} err={np.max(np.abs(g - go_)):.3g}", flush=True)
    assert abs(cost - lin.cost) <= RTOL * lin.cost
    assert np.max(np.abs(H - Ho)) <= 1e-10 * np.max(np.abs(Ho))
    assert np.max(np.abs(g - go_)) <= 1e-10 * np.max(np.abs(go_))


@pytest.mark.parametrize("factor", [0, 1])
def test_medium_parity_huber_and_options(factor):
    p = make_problem(24, 3000, 6, seed=31, outlier_frac=0.05)
    pa = O.from_problem(p)
    go, oo = _opts_pair(robust="huber", near_boost=False, factorization=factor)
    s = solver_from_problem(p, go)
    _check_run(pa, s, 8, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), "medium-huber")


@pytest.mark.parametrize("n_poses,n_lm,obs,seed,factor", [(20, 1500, 12, 41, 2), (26, 2000, 16, 42, 2),
                                                         (20, 1500, 12, 41, 3), (26, 2000, 16, 42, 3)])
def test_deep_profile_envelope(n_poses, n_lm, obs, seed, factor):
    # landmarks seen over 12 and 16 frames: the reduced system's profile is 4 to
    # 6 blocks of 16 deep, so the envelope factorisation (forced here, in one
    # workgroup or split at a separator) runs the panel on two wavefronts
    p = make_problem(n_poses, n_lm, obs, seed=seed)
    pa = O.from_problem(p)
    go, oo = _opts_pair(factorization=factor)
    s = solver_from_problem(p, go)
    print(f"[deep-profile] poses={n_poses} span={s.layout_info()['max_span']}", flush=True)
    assert s.layout_info()["factorization"] == factor
    _check_run(pa, s, 5, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"deep-{obs}-{factor}")


def test_long_trajectory_split_factorisation():
    # 48 keyframes: 18 blocks of 16 rows with a profile 3 blocks deep, so the
    # automatic choice is the split envelope factorisation; the one-workgroup
    # envelope and the oracle must agree with it
    p = make_problem(48, 3000, 8, seed=43)
    pa = O.from_problem(p)
    X0, a0 = p.X0.astype(np.float64), p.alpha0.astype(np.float64)
    for factor in (0, 2):
        go, oo = _opts_pair(factorization=factor)
        s = solver_from_problem(p, go)
        assert s.layout_info()["factorization"] == (3 if factor == 0 else 2)
        _check_run(pa, s, 4, oo, X0, a0, f"long-{factor}")


def test_reference_mode_step_size_and_split_steps():
    p = make_problem(8, 400, 4, seed=5)
    pa = O.from_problem(p)
    go, oo = _opts_pair(mode="reference")
    oo.step_size = 0.5
    s = solver_from_problem(p, go)
    s.set_step_size(0.5)
    X0, a0 = p.X0.astype(np.float64), p.alpha0.astype(np.float64)
    # step(3) then step(2) continues from the same state (reference semantics)
    X, al, _, _, _ = O.lm_solve(pa, X0, a0, 5, oo)
    s.step(3)
    s.step(2)
    gX, gal = s.get_state()
    assert np.max(np.abs(gX[:, :3] - X[:, :3])) <= RTOL * max(1.0, np.max(np.abs(X)))
    used = np.unique(pa.l)
    assert np.max(np.abs(gal[used] - al[used]) / al[used]) <= RTOL


def test_lm_lambda_persists_across_steps():
    p = make_problem(8, 400, 4, seed=9, outlier_frac=0.1)
    pa = O.from_problem(p)
    go, oo = _opts_pair(mode="lm")
    s = solver_from_problem(p, go)
    X0, a0 = p.X0.astype(np.float64), p.alpha0.astype(np.float64)
    X, al, lam, _, recs = O.lm_solve(pa, X0, a0, 6, oo)
    s.step(4)
    s.step(2)
    assert s.get_lambda() == pytest.approx(lam, rel=1e-12)
    gX, gal = s.get_state()
    assert np.max(np.abs(gX[:, :3] - X[:, :3])) <= RTOL * max(1.0, np.max(np.abs(X)))


def _edge_case_problem():
    """Measurements the reference accepts but the generator never makes:
    a target before the anchor, a cell written twice, a landmark with no
    anchor-cell pixel (zero ray, Solver_cuda.cu:147), a landmark seen only in
    its host frame, and an unused landmark slot."""
    p = make_problem(6, 60, 3, seed=12)
    a, t, l, uv = [list(x) for x in (p.meas_a, p.meas_t, p.meas_l, p.uv)]
    L = p.n_landmarks
    # a landmark anchored at frame >= 2 gets a measurement at t = a - 1
    lid = int(next(i for i in range(L) if p.meas_a[3 * i] >= 2))
    aa = int(p.meas_a[3 * lid])
    a.append(aa), t.append(aa - 1), l.append(lid), uv.append(np.array(uv[3 * lid]) + 3.0)
    # duplicate of an existing measurement (same cell)
    a.append(int(p.meas_a[5])), t.append(int(p.meas_t[5])), l.append(int(p.meas_l[5])), uv.append(uv[5])
    # landmark in a fresh slot with measurements only at a+1, a+2 (zero ray pixel)
    fresh = L
    a += [1, 1]
    t += [2, 3]
    l += [fresh, fresh]
    uv += [np.array([300.0, 200.0, 290.0, 200.0]), np.array([310.0, 205.0, 300.0, 205.0])]
    # landmark seen only in its host frame
    a.append(2), t.append(2), l.append(fresh + 1), uv.append(np.array([330.0, 250.0, 321.0, 250.0]))
    A = np.array(a, np.int32)
    Tt = np.array(t, np.int32)
    Ll = np.array(l, np.int32)
    UV = np.array(uv, np.float32)
    M = p.m_per_frame + 2
    alpha0 = np.ones(p.n_poses * M, np.float32)
    alpha0[:L] = p.alpha0[:L]
    alpha0[fresh] = 0.4
    alpha0[fresh + 1] = 0.3
    return p, A, Tt, Ll, UV, M, alpha0


@pytest.mark.parametrize("mode", ["lm", "reference"])
def test_edge_case_measurements(mode):
    p, A, Tt, Ll, UV, M, alpha0 = _edge_case_problem()
    N = p.n_poses
    pa = O.problem_arrays(N, N * M, p.intrinsics, p.T_r_to_l, A, Tt, Ll, UV)
    go, oo = _opts_pair(mode=mode)
    s = Solver(N, M, go)
    s.load_calibration(p.intrinsics, p.T_r_to_l)
    # half through the single-measurement entry point, half in bulk
    h = len(A) // 2
    for i in range(h):
        s.write_observation(A[i], Tt[i], Ll[i], [UV[i, 0], UV[i, 1], 1.0], [UV[i, 2], UV[i, 3], 1.0])
    s.write_observations(A[h:], Tt[h:], Ll[h:], UV[h:])
    for k in range(N):
        s.write_incremental_pose(k, p.X0[k].reshape(16))
    s.load_inverse_depths(alpha0)
    _check_run(pa, s, 6, oo, p.X0.astype(np.float64), alpha0.astype(np.float64), f"edge-{mode}")


def test_singular_system_reported_and_rejected():
    # pose 5 (frames 5 -> 6) is observed by nothing: with lambda = 0 the reduced
    # system is singular; the reference ignores the Cholesky info
    # (linear_solvers.cu:205), here it is reported and the state is kept.
    p = make_problem(6, 80, 2, seed=4)
    keep = p.meas_t <= 5
    a, t, l, uv = p.meas_a[keep], p.meas_t[keep], p.meas_l[keep], p.uv[keep]
    pa = O.problem_arrays(6, p.n_slots, p.intrinsics, p.T_r_to_l, a, t, l, uv)
    for mode in ["reference", "lm"]:
        go, oo = _opts_pair(mode=mode)
        s = Solver(6, p.m_per_frame, go)
        s.load_calibration(p.intrinsics, p.T_r_to_l)
        s.write_observations(a, t, l, uv)
        for k in range(6):
            s.write_incremental_pose(k, p.X0[k].reshape(16))
        s.load_inverse_depths(p.alpha0)
        _check_run(pa, s, 3, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"singular-{mode}")
        _, recs = s.stats()
        if mode == "reference":
            assert not any(r["chol_ok"] for r in recs)
        else:
            assert all(r["chol_ok"] for r in recs)


def test_graph_and_direct_launch_identical():
    # one graph carries a fixed number of iterations and a second one the
    # remainder: calls of 5, 4, 37, 70 and 3 iterations capture, reuse and
    # re-capture both; 70 records outgrow the initial record buffer
    p = make_problem(10, 800, 4, seed=8)
    res = []
    for ug in (1, 0):
        s = solver_from_problem(p, make_options(use_graph=ug))
        calls = []
        for n in (5, 4, 37, 70, 3):
            s.step(n)
            calls.append((s.get_state(), s.stats()))
        res.append(calls)
    for ((X1, a1), (c1, r1)), ((X2, a2), (c2, r2)), n in zip(res[0], res[1], (5, 4, 37, 70, 3)):
        assert np.array_equal(X1, X2) and np.array_equal(a1, a2)
        assert len(r1) == n
        assert [r["cost"] for r in r1] == [r["cost"] for r in r2]
        assert [r["lam"] for r in r1] == [r["lam"] for r in r2]


def test_tile_shapes_do_not_change_results_beyond_rounding():
    p = make_problem(10, 800, 4, seed=8)
    base = None
    for cl, gc in [(64, 1), (32, 2), (16, 4), (64, 8)]:
        s = solver_from_problem(p, make_options(chunk_landmarks=cl, group_chunks=gc))
        s.step(4)
        X, al = s.get_state()
        if base is None:
            base = (X, al)
        else:
            assert np.max(np.abs(X - base[0])) < 1e-11
            assert np.max(np.abs(al - base[1]) / np.abs(base[1])) < 1e-11


def test_c2_full_size_properties_and_first_iterations():
    p = make_config("c2_fp64")
    pa = O.from_problem(p)
    go, oo = _opts_pair(mode="lm")
    s = solver_from_problem(p, go)
    X0, a0 = p.X0.astype(np.float64), p.alpha0.astype(np.float64)
    # oracle for the first two iterations
    _, _, _, c0, recs = O.lm_solve(pa, X0, a0, 2, oo)
    s.step(2)
    g0, grecs = s.stats()
    assert abs(g0 - c0) <= RTOL * c0
    for r, q in zip(recs, grecs):
        assert abs(q["cost"] - r.cost) <= RTOL * r.cost
    s.step(18)
    _, grecs = s.stats()
    costs = [q["cost"] for q in grecs]
    assert all(b <= a for a, b in zip(costs, costs[1:]))
    assert costs[-1] < 0.2 * c0
    X, al = s.get_state()
    err = O.pose_errors(p.X_true, X)
    err0 = O.pose_errors(p.X_true, X0)
    print(f"[c2] cost {c0:.6g} -> {costs[-1]:.6g}; pose err {err0.mean(0)} -> {err.mean(0)}", flush=True)
    assert err[:, 0].mean() < 0.1 * err0[:, 0].mean() and err[:, 1].mean() < 0.1 * err0[:, 1].mean()


def test_mixed_precision_reaches_fp64_final_cost():
    # FP32 linearisation and landmark Schur, FP64 reduced system and Cholesky:
    # after the same number of LM iterations the cost agrees with the FP64 path
    # to 1e-6 relative (the FP32 cost's rounding makes accept/reject decisions
    # near convergence differ, so the iterates are not compared)
    p = make_problem(40, 4000, 10, seed=77, outlier_frac=0.05)
    finals = {}
    for prec in ("fp64", "mixed"):
        s = solver_from_problem(p, make_options(mode="lm", robust="huber", robust_param=2.0, precision=prec))
        s.step(30)
        c0, recs = s.stats()
        finals[prec] = recs[-1]["cost"]
        print(f"[mixed] {prec}: c0={c0:.10g} final={finals[prec]:.12g}", flush=True)
        assert recs[-1]["cost"] < 0.1 * c0
    assert abs(finals["mixed"] - finals["fp64"]) <= 1e-6 * finals["fp64"]


def test_mixed_reduced_system_tracks_fp64_along_the_trajectory():
    # The mixed path is compared where its iterates cannot drift apart: at the
    # states of an FP64 run, its reduced system, gradient and cost against the
    # FP64 path's, and the damped step each system gives.  Bars (FP32
    # linearisation, FP64 sums of the blocks; about ten times the errors seen):
    # cost 1e-7 relative; H 2e-6 of its largest entry; g 1e-5 of the largest
    # entry of the first state's gradient
    # (near convergence g is a small difference of sums whose FP32 rounding does
    # not shrink with it); the step 5e-5 of the first state's step in norm (the
    # same reason).
    p = make_problem(48, 6000, 10, seed=78, outlier_frac=0.05)
    o = dict(mode="lm", robust="huber", robust_param=2.0)
    s64 = solver_from_problem(p, make_options(precision="fp64", **o))
    s32 = solver_from_problem(p, make_options(precision="mixed", **o))
    lam = 1e-3
    gscale = dscale = None
    for it in range(4):
        X, al = s64.get_state()
        s32.set_state(X, al)
        H64, g64, c64 = s64.linearize(lam)
        H32, g32, c32 = s32.linearize(lam)
        eH = np.max(np.abs(H32 - H64)) / np.max(np.abs(H64))
        gscale = np.max(np.abs(g64)) if gscale is None else gscale
        eg = np.max(np.abs(g32 - g64)) / gscale
        d64 = np.linalg.solve(H64, g64)
        d32 = np.linalg.solve(H32, g32)
        dscale = np.linalg.norm(d64) if dscale is None else dscale
        ed = np.linalg.norm(d32 - d64) / dscale
        print(f"[mixed-track {it}] cost {abs(c32 - c64) / c64:.3g} H {eH:.3g} g {eg:.3g} step {ed:.3g}", flush=True)
        assert abs(c32 - c64) <= 1e-7 * c64
        assert eH <= 2e-6 and eg <= 1e-5
        assert ed <= 5e-5
        s64.step(3)
