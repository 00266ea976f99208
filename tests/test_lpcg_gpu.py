"""Learned-preconditioner PCG on the GPU, through the C ABI: features, loss and
gradient against the CPU oracle (bfloat16 MFMA operands and float32
accumulation in the network, FP64 loss), PCG against dense
solves, and what training gains over block Jacobi on held-out systems that
the solver itself generated."""
import numpy as np
import pytest

from oracle import lpcg_oracle as LP
from spgo.lpcg import LearnedPCG, band_width, solver_systems
from spgo.problems import make_problem

pytestmark = pytest.mark.gpu

W = 6


def _problems(seeds, N=24, L=1200, K=6):
    return [make_problem(N, L, K, seed=s) for s in seeds]


def test_features_loss_and_gradient_match_the_oracle():
    systems = solver_systems(_problems([11, 12], N=20, L=900), W, states_per_problem=2)
    lp = LearnedPCG(W, seed=5)
    p = lp.get_params()
    rng = np.random.default_rng(3)
    p = p + rng.normal(0, 0.02, p.size).astype(np.float32)   # every path live, head away from zero
    lp.set_params(p)
    lp.load(systems)
    X, C = lp.features()
    Hbs = [s[0] for s in systems]
    Cc = np.concatenate([LP.features(Hb)[0] for Hb in Hbs])
    Xc = np.concatenate([LP.features(Hb)[3] for Hb in Hbs])
    assert np.allclose(C, Cc, rtol=1e-10, atol=1e-12 * np.abs(Cc).max())
    assert np.allclose(X, Xc, rtol=1e-5, atol=1e-6)
    Lg, gg = lp.loss_grad()
    # the GPU's layers take bfloat16 operands on MFMA: the oracle rounds the same operands
    Lc, gc = LP.loss_and_grad(LP.Params.from_flat(p.astype(np.float64)), Hbs, bf=True)
    print(f"[lpcg-grad] loss gpu {Lg:.9g} cpu {Lc:.9g}; |grad| {np.linalg.norm(gc):.4g} "
          f"err {np.linalg.norm(gg - gc):.3g}", flush=True)
    assert abs(Lg - Lc) <= 1e-6 + 1e-4 * abs(Lc)
    assert np.linalg.norm(gg - gc) <= 2e-3 * np.linalg.norm(gc)


def test_pcg_solves_and_block_jacobi_matches_the_oracle():
    systems = solver_systems(_problems([21, 22, 23], N=30, L=1500), W, states_per_problem=1)
    lp = LearnedPCG(W, seed=1)
    lp.load(systems)
    it, rel, x = lp.solve(learned=False, tol=1e-10)
    o = 0
    for k, (Hb, g) in enumerate(systems):
        n = Hb.shape[0]
        H = LP.dense_from_band(Hb)
        xs = x[o:o + n].reshape(-1)
        assert rel[k] <= 1e-10
        assert np.linalg.norm(H @ xs - g.reshape(-1)) <= 1.5e-10 * np.linalg.norm(g)
        C = LP.features(Hb)[0]
        _, it_c, _ = LP.pcg(Hb, g, C=C, tol=1e-10)
        print(f"[lpcg-pcg] system {k}: {it[k]} iterations (oracle {it_c})", flush=True)
        assert abs(int(it[k]) - it_c) <= 2
        o += n
    # the untrained network is block Jacobi: same iteration counts
    it2, rel2, _ = lp.solve(learned=True, tol=1e-10)
    assert np.all(np.abs(it2 - it) <= 1) and np.all(rel2 <= 1e-10)


def test_training_beats_block_jacobi_on_held_out_systems():
    train = solver_systems(_problems(range(100, 106), N=32, L=1500), W, states_per_problem=2)
    held = solver_systems(_problems(range(200, 203), N=40, L=2000), W, states_per_problem=2)
    assert max(band_width(LP.dense_from_band(Hb)) for Hb, _ in train + held) <= W
    lp = LearnedPCG(W, seed=7)
    lp.load(train)
    losses = lp.train(200, lr=1e-3)
    print(f"[lpcg-train] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[0] == pytest.approx(0.0, abs=1e-6)      # block Jacobi scores 0
    assert losses[-1] < -0.05
    lp.load(held)
    it_j, rel_j, _ = lp.solve(learned=False, tol=1e-8)
    it_l, rel_l, x = lp.solve(learned=True, tol=1e-8)
    print(f"[lpcg-train] held-out PCG iterations: block Jacobi {it_j.tolist()}, learned {it_l.tolist()}", flush=True)
    assert np.all(rel_l <= 1e-8) and np.all(rel_j <= 1e-8)
    assert it_l.mean() <= 0.8 * it_j.mean()
    # the exact solution judges the learned solve: dense residual of the system
    o = 0
    for Hb, g in held:
        n = Hb.shape[0]
        H = LP.dense_from_band(Hb)
        assert np.linalg.norm(H @ x[o:o + n].reshape(-1) - g.reshape(-1)) <= 1.5e-8 * np.linalg.norm(g)
        o += n


def _lm_run(p, iters, **kw):
    from spgo.solver import make_options, solver_from_problem
    net = kw.pop("net", None)
    s = solver_from_problem(p, make_options(mode="lm", **kw))
    if net is not None:
        s.set_preconditioner(net)
    s.step(iters)
    return s.stats(), s


def test_lm_step_with_cg_matches_the_cholesky_path():
    p = make_problem(40, 3000, 6, seed=61)
    (c0, ref), _ = _lm_run(p, 10)
    # block-Jacobi CG to 1e-11: the same LM iterates as the exact solve
    (c1, cg), s = _lm_run(p, 10, factorization=4, pcg_tol=1e-11, pcg_max_iter=2000)
    assert s.layout_info()["factorization"] == 4
    print(f"[lm-cg] CG iterations per step {[r['pcg_iters'] for r in cg]}", flush=True)
    assert c1 == pytest.approx(c0, rel=1e-12)
    for r, q in zip(ref, cg):
        assert q["accepted"] == r["accepted"] and not q["pcg_fallback"] and q["pcg_iters"] > 0
        assert q["cost"] == pytest.approx(r["cost"], rel=1e-8)
    # CG that cannot converge in 2 iterations: the Cholesky computes every step
    (c2, fb), _ = _lm_run(p, 10, factorization=4, pcg_max_iter=2)
    for r, q in zip(ref, fb):
        assert q["pcg_fallback"] and q["pcg_iters"] == 2
        assert q["cost"] == pytest.approx(r["cost"], rel=1e-12) and q["accepted"] == r["accepted"]


def test_lm_step_with_the_trained_network():
    train = solver_systems(_problems(range(300, 306), N=32, L=1500), W, states_per_problem=2)
    lp = LearnedPCG(W, seed=9)
    lp.load(train)
    lp.train(200, lr=1e-3)
    p = make_problem(48, 4000, 6, seed=77)
    (c0, ref), _ = _lm_run(p, 15)
    (cj, bj), _ = _lm_run(p, 15, factorization=4)
    (cl, le), s = _lm_run(p, 15, factorization=4, net=lp)
    it_j = [r["pcg_iters"] for r in bj]
    it_l = [r["pcg_iters"] for r in le]
    print(f"[lm-cg] CG iterations per LM step: block Jacobi {it_j}, learned {it_l}; final cost "
          f"Cholesky {ref[-1]['cost']:.10g} learned CG {le[-1]['cost']:.10g}", flush=True)
    assert not any(r["pcg_fallback"] for r in le)
    assert sum(it_l) <= 0.8 * sum(it_j)
    # PCG to 1e-8 reaches the exact-solve LM's final cost to 1e-6
    assert abs(le[-1]["cost"] - ref[-1]["cost"]) <= 1e-6 * ref[-1]["cost"]
