"""The mixed-precision, batched and learned-CG workloads at their benchmark
sizes, through properties that do not depend on size: the initial cost
against the FP64 oracle, monotone accepted costs, and the final cost against
the FP64 path (within 1e-6 relative, the precision target of these modes).

256 keyframes / 100,000 landmarks / 1,000,000 stereo measurements with 5 %
outliers (Huber 2 px); 512 independent windows of 8-32 keyframes and
200-2,000 landmarks; 1,024 keyframes / 400,000 landmarks with 12 observations
each and loop revisits at 20 % of the frames.  The sizes the oracle solves in seconds are covered in
test_gpu_parity.py, test_batch.py and test_lpcg_gpu.py.
"""
import numpy as np
import pytest

from oracle import ba_oracle as O
from spgo.problems import WINDOW_BATCH, make_config, make_window_batch
from spgo.solver import Batch, make_options, solver_from_problem

RTOL = 1e-9        # FP64 path against the FP64 oracle
MIXED_C0 = 1e-5    # FP32 linearisation: the initial cost against the FP64 oracle
FINAL = 1e-6       # final cost of a mixed-precision or batched run against the FP64 path


@pytest.mark.gpu
def test_mixed_precision_full_size_against_oracle_and_fp64():
    p = make_config("c3_mixed")
    pa = O.from_problem(p)
    oo = O.Options(mode="lm", robust="huber", robust_param=2.0, near_boost=True)
    # the solver projects uploaded rotations onto SO(3); so does the oracle here
    X0 = O.project_rotations(p.X0.astype(np.float64))
    c_oracle = O.cost_only(pa, X0, p.alpha0.astype(np.float64), oo)
    finals = {}
    for prec in ("fp64", "mixed"):
        s = solver_from_problem(p, make_options(mode="lm", robust="huber", robust_param=2.0, near_boost=True,
                                                precision=prec))
        s.step(10)
        c0, recs = s.stats()
        costs = [r["cost"] for r in recs]
        print(f"[full-c3] {prec}: oracle c0 {c_oracle:.12g}, gpu c0 {c0:.12g}, final {costs[-1]:.12g}", flush=True)
        assert abs(c0 - c_oracle) <= (RTOL if prec == "fp64" else MIXED_C0) * c_oracle
        assert len(recs) == 10
        assert all(b <= a for a, b in zip(costs, costs[1:]))
        assert costs[-1] < 0.2 * c0
        finals[prec] = costs[-1]
    assert abs(finals["mixed"] - finals["fp64"]) <= FINAL * finals["fp64"]


@pytest.mark.gpu
def test_batch_full_size_against_lone_fp64_solvers():
    ws = make_window_batch(**WINDOW_BATCH)
    assert len(ws) == 512
    o = make_options(mode="lm", precision="mixed", use_graph=1, chunk_landmarks=64, group_chunks=1)
    bt = Batch([solver_from_problem(w, o) for w in ws], o)
    bt.step(10)
    fo = make_options(mode="lm")
    worst = 0.0
    for p in range(0, len(ws), 8):
        s = solver_from_problem(ws[p], fo)
        s.step(10)
        lone_c0, lone = s.stats()
        c0, recs = bt.stats(p)
        assert len(recs) == 10
        costs = [r["cost"] for r in recs]
        assert all(b <= a for a, b in zip(costs, costs[1:])), p
        assert abs(c0 - lone_c0) <= MIXED_C0 * lone_c0, p
        rel = abs(costs[-1] - lone[-1]["cost"]) / lone[-1]["cost"]
        worst = max(worst, rel)
        assert rel <= FINAL, (p, rel)
    print(f"[full-c4] 64 of 512 windows: worst final-cost difference to the lone FP64 solver {worst:.3g}",
          flush=True)


@pytest.mark.gpu
def test_revisit_trajectory_full_size_cholesky_and_cg():
    # 1,024 keyframes (a 6,144 x 6,144 reduced system), 400,000 landmarks, 12
    # observations each, the last 20 % of the frames revisiting the first:
    # landmarks observed again ~820 frames after their anchor, solved in frame
    # coordinates.  The FP64 Cholesky path and LM with CG (the network trained
    # on revisit systems of 128-256 keyframes, the Cholesky computing the step
    # when CG does not reach 1e-8) end at the same cost; costs fall monotonically.
    from spgo import _lib
    from spgo.lpcg import LearnedPCG, solver_systems
    from spgo.problems import make_revisit_problem

    K = 12
    train = [make_revisit_problem(n, 390 * n, K, seed=5100 + n) for n in (128, 192, 256)]
    net = LearnedPCG(K - 1, seed=11)
    net.load(solver_systems(train, K - 1, states_per_problem=3, iters_between=1, restarts=4))
    net.train(200, lr=1e-3)
    p = make_revisit_problem(1024, 400000, K, seed=5005)
    assert int(np.max(p.meas_t - p.meas_a)) > 800
    finals = {}
    for tag, kw in (("cholesky", {}), ("cg", {"factorization": 4, "pcg_max_iter": 100})):
        s = solver_from_problem(p, make_options(mode="lm", **kw))
        assert s.layout_info()["coordinates"] == _lib.SPGO_COORD_FRAMES
        if tag == "cg":
            s.set_preconditioner(net)
        s.step(12)
        c0, recs = s.stats()
        costs = [r["cost"] for r in recs]
        assert all(b <= a for a, b in zip(costs, costs[1:])), tag
        assert costs[-1] < 0.2 * c0, tag
        if tag == "cg":
            print(f"[full-c5] CG iterations {[r['pcg_iters'] for r in recs]}, "
                  f"fallbacks {sum(r['pcg_fallback'] for r in recs)}", flush=True)
        finals[tag] = costs[-1]
        s.close()
    net.close()
    rel = abs(finals["cg"] - finals["cholesky"]) / finals["cholesky"]
    print(f"[full-c5] CG: final cost {finals['cg']:.12g} against {finals['cholesky']:.12g}, "
          f"relative difference {rel:.3g}", flush=True)
    assert rel <= FINAL
