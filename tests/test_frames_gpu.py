"""Parity of the frame-coordinate path (SPGO_COORD_FRAMES) with the FP64 CPU
oracle, through the C ABI.

The oracle builds the reference's incremental-pose system with the explicit
Jacobian over every state in [anchor, target) (Solver_cuda.cu:248-309); the
library solves the same step in frame coordinates.  Bar: float64 work, so
per-iteration costs and the final state agree to 1e-9 relative, accept/reject
decisions and Cholesky outcomes are identical; the reduced system mapped back
to incremental coordinates agrees to 1e-10.  Covered: short spans forced into
frame coordinates (both modes, both factorisations), landmarks spanning up to
34 poses with loop revisits, and the reference harness's enumeration at 30
keyframes (landmarks seen in up to 30 frames).
"""
import numpy as np
import pytest

from build_ip import Solver as DropIn
from oracle import ba_oracle as O
from spgo import _lib
from spgo.problems import make_config, make_manager_problem, make_problem, make_revisit_problem
from spgo.solver import make_options, solver_from_problem

pytestmark = pytest.mark.gpu

RTOL = 1e-9
FRAMES = _lib.SPGO_COORD_FRAMES


def _check(p, go, oo, iters, tag, X0=None, a0=None):
    pa = O.from_problem(p)
    X0 = p.X0.astype(np.float64) if X0 is None else X0
    a0 = p.alpha0.astype(np.float64) if a0 is None else a0
    s = solver_from_problem(p, go)
    info = s.layout_info()
    print(f"[{tag}] coordinates={info['coordinates']} factorization={info['factorization']} "
          f"max_span={info['max_span']} groups={info['n_groups']}", flush=True)
    assert info["coordinates"] == FRAMES
    X, al, lam, c0, recs = O.lm_solve(pa, X0, a0, iters, oo)
    s.step(iters)
    g0, grecs = s.stats()
    print(f"[{tag}] oracle c0={c0:.12g} gpu c0={g0:.12g} final {recs[-1].cost:.12g} {grecs[-1]['cost']:.12g}",
          flush=True)
    assert abs(g0 - c0) <= RTOL * abs(c0)
    for i, (r, q) in enumerate(zip(recs, grecs)):
        assert q["accepted"] == r.accepted, (tag, i)
        assert q["chol_ok"] == r.chol_ok, (tag, i)
        assert abs(q["cost_candidate"] - r.cost_candidate) <= RTOL * abs(r.cost_candidate), (tag, i)
        assert abs(q["cost"] - r.cost) <= RTOL * abs(r.cost), (tag, i)
    gX, gal = s.get_state()
    assert np.max(np.abs(gX[:, :3, :] - X[:, :3, :])) <= RTOL * max(1.0, np.max(np.abs(X)))
    used = np.unique(pa.l)
    assert np.max(np.abs(gal[used] - al[used]) / np.abs(al[used])) <= RTOL
    return s


@pytest.mark.parametrize("mode,factor", [("lm", 0), ("reference", 0), ("lm", 1), ("reference", 1)])
def test_short_spans_in_frame_coordinates(mode, factor):
    p = make_config("c1_small")
    go = make_options(mode=mode, factorization=factor, coordinates=FRAMES)
    _check(p, go, O.Options(mode=mode), 10, f"c1-frames-{mode}-{factor}")


def test_frame_system_maps_to_the_oracle_system():
    p = make_revisit_problem(24, 700, 5, seed=71, revisit_frac=0.3)
    pa = O.from_problem(p)
    s = solver_from_problem(p, make_options())
    lam = 0.25
    H, g, cost = s.linearize(lam)
    X0 = O.project_rotations(p.X0.astype(np.float64))
    lin = O.linearize(pa, X0, p.alpha0.astype(np.float64), O.Options())
    Ho, go_ = O.reduced_system(lin, lam, 1e-4)
    print(f"[frames-system] |H|={np.max(np.abs(Ho)):.4g} err={np.max(np.abs(H - Ho)):.3g} "
          f"|g|={np.max(np.abs(go_)):.4g} err={np.max(np.abs(g - go_)):.3g}", flush=True)
    assert s.layout_info()["coordinates"] == FRAMES
    assert abs(cost - lin.cost) <= RTOL * lin.cost
    assert np.max(np.abs(H - Ho)) <= 1e-10 * np.max(np.abs(Ho))
    assert np.max(np.abs(g - go_)) <= 1e-10 * np.max(np.abs(go_))


@pytest.mark.parametrize("mode", ["lm", "reference"])
def test_loop_revisits_match_oracle(mode):
    # 40 keyframes, the last quarter revisiting the first: spans up to 34 poses
    p = make_revisit_problem(40, 1500, 6, seed=7, revisit_frac=0.25)
    assert int(np.max(p.meas_t - p.meas_a)) > 30
    _check(p, make_options(mode=mode), O.Options(mode=mode), 8, f"revisit-{mode}")


def test_huber_with_outliers_in_frame_coordinates():
    p = make_problem(16, 900, 6, seed=17, outlier_frac=0.05)
    go = make_options(mode="lm", robust="huber", robust_param=2.0, coordinates=FRAMES)
    _check(p, go, O.Options(robust="huber", robust_param=2.0), 8, "huber-frames")


def test_reference_harness_at_30_keyframes():
    # CUDAmanager.py:69-93 enumeration: landmarks written for every frame that
    # sees them from their host on -- up to 30 frames, spans up to 30 poses
    p = make_manager_problem(30, 12, seed=130)
    s = DropIn.CudaSolver(p.n_poses, p.m_per_frame)
    s.loadCalibration(p.intrinsics, p.T_r_to_l)
    for a, t, l, uv in zip(p.meas_a, p.meas_t, p.meas_l, p.uv):
        s.writeObservations(int(a), int(t), int(l), np.array([uv[0], uv[1], 1.0], np.float32),
                            np.array([uv[2], uv[3], 1.0], np.float32))
    for k in range(p.n_poses):
        s.writeIncrementalPose(k, np.eye(4).reshape(16).astype(np.float32))
    info = s.solver.layout_info()
    print(f"[harness-30] measurements={info['n_measurements']} max_span={info['max_span']} "
          f"coordinates={info['coordinates']}", flush=True)
    assert info["coordinates"] == FRAMES and info["max_span"] == 30
    pa = O.from_problem(p)
    X = np.tile(np.eye(4), (30, 1, 1))
    al = np.ones(p.n_slots)
    for call in range(2):
        X, al, _, c0, recs = O.lm_solve(pa, X, al, 2, O.Options(mode="reference"))
        s.step(2)
        g0, grecs = s.solver.stats()
        assert abs(g0 - c0) <= RTOL * c0
        for r, q in zip(recs, grecs):
            assert abs(q["cost"] - r.cost) <= RTOL * r.cost
        gX, gal = s.solver.get_state()
        assert np.max(np.abs(gX[:, :3] - X[:, :3])) <= RTOL * max(1.0, np.max(np.abs(X)))
        used = np.unique(pa.l)
        assert np.max(np.abs(gal[used] - al[used]) / al[used]) <= RTOL


def test_mixed_precision_frames_reaches_fp64_final_cost():
    p = make_revisit_problem(48, 3000, 8, seed=9, revisit_frac=0.2)
    finals = {}
    for prec in ("fp64", "mixed"):
        s = solver_from_problem(p, make_options(precision=prec))
        assert s.layout_info()["coordinates"] == FRAMES
        s.step(15)
        c0, recs = s.stats()
        finals[prec] = recs[-1]["cost"]
        print(f"[frames-{prec}] c0={c0:.10g} final={finals[prec]:.10g}", flush=True)
    assert abs(finals["mixed"] - finals["fp64"]) <= 1e-6 * finals["fp64"]
