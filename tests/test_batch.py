"""Batches of independent windows (spgo_batch_*): every problem of a batch takes
exactly the steps it takes alone.

The batch runs the same kernels as a lone solver on the same per-problem data
(same layout, envelope Cholesky, summation order), so the FP64 records and
states of a batch member equal those of the lone solver with the envelope
factorisation bit for bit; the tests allow 1e-12 relative.  One window is also
checked directly against the FP64 oracle at the parity bar (1e-9).
"""
import numpy as np
import pytest

from oracle import ba_oracle as O
from spgo.problems import make_problem, make_window_batch
from spgo.solver import Batch, Solver, make_options, solver_from_problem

TIGHT = 1e-12


def _windows(n, seed=11, **kw):
    return make_window_batch(n, n_range=(8, 20), l_range=(150, 500), k_range=(3, 8), seed=seed, **kw)


# ---------------------------------------------------------------- host side (CPU)
def test_batch_create_rejects_bad_input():
    ws = _windows(2)
    a = solver_from_problem(ws[0])
    b = solver_from_problem(ws[1])
    with pytest.raises(RuntimeError, match="repeats"):
        Batch([a, a])
    intr = ws[1].intrinsics.copy()
    intr[0] += 1.0
    b.load_calibration(intr, ws[1].T_r_to_l)
    with pytest.raises(RuntimeError, match="calibration"):
        Batch([a, b])
    with pytest.raises(RuntimeError, match="at least one"):
        Batch([])


def test_batch_state_before_first_step_is_the_solvers():
    ws = _windows(3)
    ss = [solver_from_problem(w) for w in ws]
    bt = Batch(ss)
    info = bt.info()
    assert info["n_problems"] == 3 and info["n_poses"] == 0
    for p, s in enumerate(ss):
        X, al = bt.get_state(p)
        Xs, als = s.get_state()
        assert np.array_equal(X, Xs) and np.array_equal(al, als)
    bt.store()   # nothing stepped: no-op


# ---------------------------------------------------------------- GPU
def _lone(w, opts, iters):
    s = solver_from_problem(w, opts)
    s.step(iters)
    return s.stats(), s.get_state()


def _compare(bt, p, lone, tag, tol=TIGHT):
    (c0, recs), (X, al) = lone
    bc0, brecs = bt.stats(p)
    assert len(brecs) == len(recs), tag
    assert abs(bc0 - c0) <= tol * abs(c0), (tag, bc0, c0)
    for i, (r, q) in enumerate(zip(recs, brecs)):
        assert q["accepted"] == r["accepted"] and q["chol_ok"] == r["chol_ok"], (tag, i)
        assert abs(q["cost"] - r["cost"]) <= tol * abs(r["cost"]), (tag, i, q["cost"], r["cost"])
        assert abs(q["cost_candidate"] - r["cost_candidate"]) <= tol * abs(r["cost_candidate"]), (tag, i)
        assert q["lam"] == pytest.approx(r["lam"], rel=1e-14, abs=0), (tag, i)
    bX, bal = bt.get_state(p)
    assert np.max(np.abs(bX - X)) <= tol * max(1.0, np.max(np.abs(X))), tag
    used = al != 0
    assert np.max(np.abs(bal[used] - al[used]) / np.abs(al[used])) <= tol, tag


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["lm", "reference"])
def test_batch_members_match_lone_solvers_fp64(mode):
    ws = _windows(12, seed=21)
    opts = make_options(mode=mode, factorization=2)
    bt = Batch([solver_from_problem(w, opts) for w in ws], opts)
    bt.step(8)
    info = bt.info()
    print(f"[batch-{mode}] {info}", flush=True)
    assert info["n_poses"] == sum(w.n_poses for w in ws)
    assert info["n_measurements"] == sum(w.n_obs for w in ws)
    for p, w in enumerate(ws):
        _compare(bt, p, _lone(w, opts, 8), f"{mode}-{p}")


@pytest.mark.gpu
def test_batch_window_matches_oracle():
    ws = _windows(4, seed=5)
    opts = make_options(mode="lm")
    bt = Batch([solver_from_problem(w, opts) for w in ws], opts)
    bt.step(6)
    p = 2
    w = ws[p]
    pa = O.from_problem(w)
    X, al, lam, c0, recs = O.lm_solve(pa, w.X0.astype(np.float64), w.alpha0.astype(np.float64), 6,
                                      O.Options(mode="lm", robust="cauchy", robust_param=9.0, near_boost=True))
    bc0, brecs = bt.stats(p)
    print(f"[batch-oracle] oracle c0={c0:.12g} batch c0={bc0:.12g} final {recs[-1].cost:.12g} "
          f"{brecs[-1]['cost']:.12g}", flush=True)
    assert abs(bc0 - c0) <= 1e-9 * c0
    for i, (r, q) in enumerate(zip(recs, brecs)):
        assert q["accepted"] == r.accepted and q["chol_ok"] == r.chol_ok, i
        assert abs(q["cost"] - r.cost) <= 1e-9 * r.cost, i
    bX, bal = bt.get_state(p)
    assert np.max(np.abs(bX[:, :3, :] - X[:, :3, :])) <= 1e-9 * max(1.0, np.max(np.abs(X)))
    used = np.unique(pa.l)
    assert np.max(np.abs(bal[used] - al[used]) / np.abs(al[used])) <= 1e-9


@pytest.mark.gpu
def test_batch_mixed_precision_members_match_lone_mixed():
    # FP32 assembly: the same kernels on the same data, so batch and lone agree
    # as tightly as in FP64
    ws = _windows(8, seed=33, outlier_frac=0.05)
    opts = make_options(mode="lm", robust="huber", robust_param=2.0, precision="mixed", factorization=2)
    bt = Batch([solver_from_problem(w, opts) for w in ws], opts)
    bt.step(10)
    for p, w in enumerate(ws):
        _compare(bt, p, _lone(w, opts, 10), f"mixed-{p}")


@pytest.mark.gpu
def test_batch_split_steps_graph_and_store():
    ws = _windows(6, seed=44)
    opts = make_options(mode="lm", factorization=2)
    go = make_options(mode="lm", factorization=2, use_graph=1)
    bt = Batch([solver_from_problem(w, opts) for w in ws], go)
    bt.step(3)
    bt.step(4)
    for p, w in enumerate(ws):
        s = solver_from_problem(w, opts)
        s.step(3)
        s.step(4)
        _compare(bt, p, (s.stats(), s.get_state()), f"split-{p}")
    last = [bt.stats(p)[1][-1]["cost"] for p in range(len(ws))]
    # store: the solvers now hold the batch's states
    bt.store()
    for p, s in enumerate(bt.solvers):
        X, al = bt.get_state(p)
        Xs, als = s.get_state()
        assert np.array_equal(X[:, :3, :], Xs[:, :3, :]) and np.array_equal(al, als)
    # reload: the next step starts again from the (stored) solver states, whose
    # cost is the last kept cost
    bt.reload()
    bt.step(1)
    for p in range(len(ws)):
        assert bt.stats(p)[0] == pytest.approx(last[p], rel=1e-10)


@pytest.mark.gpu
def test_batch_autotune_keeps_results():
    ws = _windows(10, seed=55)
    opts = make_options(mode="lm")
    bt = Batch([solver_from_problem(w, opts) for w in ws], opts)
    bt.step(5)
    ref = [bt.stats(p) for p in range(len(ws))]
    best = bt.autotune(2)
    print(f"[autotune] {best}", flush=True)
    assert best["chunk_landmarks"] in (16, 32, 48, 64) and best["group_chunks"] in (1, 2, 4)
    o = bt.get_options()
    assert (o.chunk_landmarks, o.group_chunks) == (best["chunk_landmarks"], best["group_chunks"])
    bt.step(5)   # reloads the unchanged solver states with the chosen shape
    for p in range(len(ws)):
        c0, recs = bt.stats(p)
        r0, rrecs = ref[p]
        assert c0 == pytest.approx(r0, rel=1e-12)
        # another tile shape changes only summation order
        assert recs[-1]["cost"] == pytest.approx(rrecs[-1]["cost"], rel=1e-9)


@pytest.mark.gpu
def test_batch_with_many_rows_matches_lone_solvers():
    # more than 4096 pose rows: the reduced systems are formed one pose per
    # workgroup, the lone solvers one row per workgroup
    ws = make_window_batch(52, n_range=(12, 20), l_range=(150, 300), k_range=(3, 8), seed=66)
    assert 6 * sum(w.n_poses for w in ws) > 4096
    opts = make_options(mode="lm", factorization=2)
    bt = Batch([solver_from_problem(w, opts) for w in ws], opts)
    bt.step(5)
    for p in (0, 17, 51):
        _compare(bt, p, _lone(ws[p], opts, 5), f"rows-{p}")
