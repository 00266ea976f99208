"""Block cyclic reduction (SPGO_FACTOR_CR) against the FP64 CPU oracle.

The reduction is the Cholesky factorisation of the reduced system with its
super-rows (16 D rows, D the profile depth in blocks of 16) eliminated
odd-even level by level, so the FP64 bar is the plain path's: every
iteration's candidate and kept cost, the accept and Cholesky flags, the final
poses and depths within 1e-9 relative.  The cases cover profile depths 1 to 4
(super-rows of 16 to 64 rows), one to five levels, a last super-row that is
partly padding, and a singular system.
"""
import numpy as np
import pytest

from oracle import ba_oracle as O
from spgo.problems import make_config, make_problem
from spgo.solver import Solver, solver_from_problem
from test_gpu_parity import _check_run, _opts_pair

pytestmark = pytest.mark.gpu

CR = 5


def _levels(info):
    S = 16 * max(info["profile_depth"], 1)
    ns = -(-info["reduced_size"] // S)
    return S, ns, int(np.ceil(np.log2(ns))) if ns > 1 else 0


@pytest.mark.parametrize("mode", ["lm", "reference"])
def test_cr_c1_parity(mode):
    p = make_config("c1_small")
    pa = O.from_problem(p)
    go, oo = _opts_pair(mode=mode, factorization=CR)
    s = solver_from_problem(p, go)
    assert s.layout_info()["factorization"] == CR
    _check_run(pa, s, 10, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"cr-c1-{mode}")


@pytest.mark.parametrize("n_poses,n_lm,obs,seed,robust",
                         [(24, 3000, 6, 31, "huber"),     # depth 2: S = 32, 5 super-rows
                          (48, 3000, 8, 43, "cauchy"),    # depth 3: S = 48, 6 super-rows
                          (100, 5000, 6, 44, "cauchy"),   # depth 2: 19 super-rows, 5 levels
                          (40, 3000, 10, 46, "cauchy"),   # depth 4: S = 64, 4 super-rows
                          (64, 4000, 10, 45, "huber")])   # depth 4: 6 super-rows
def test_cr_parity_depths_and_levels(n_poses, n_lm, obs, seed, robust):
    p = make_problem(n_poses, n_lm, obs, seed=seed, outlier_frac=0.05 if robust == "huber" else 0.0)
    pa = O.from_problem(p)
    go, oo = _opts_pair(robust=robust, factorization=CR)
    s = solver_from_problem(p, go)
    info = s.layout_info()
    assert info["factorization"] == CR
    S, ns, L = _levels(info)
    print(f"[cr] poses={n_poses} depth={info['profile_depth']} S={S} super-rows={ns} levels={L}", flush=True)
    _check_run(pa, s, 5, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"cr-{n_poses}-{obs}")


def test_cr_singular_system_reported_and_rejected():
    # pose 5 is observed by nothing: with lambda = 0 the reduced system is
    # singular and some pivot of the reduction fails; the state is kept
    p = make_problem(6, 80, 2, seed=4)
    keep = p.meas_t <= 5
    a, t, l, uv = p.meas_a[keep], p.meas_t[keep], p.meas_l[keep], p.uv[keep]
    pa = O.problem_arrays(6, p.n_slots, p.intrinsics, p.T_r_to_l, a, t, l, uv)
    for mode in ["reference", "lm"]:
        go, oo = _opts_pair(mode=mode, factorization=CR)
        s = Solver(6, p.m_per_frame, go)
        s.load_calibration(p.intrinsics, p.T_r_to_l)
        s.write_observations(a, t, l, uv)
        for k in range(6):
            s.write_incremental_pose(k, p.X0[k].reshape(16))
        s.load_inverse_depths(p.alpha0)
        _check_run(pa, s, 3, oo, p.X0.astype(np.float64), p.alpha0.astype(np.float64), f"cr-singular-{mode}")
        _, recs = s.stats()
        if mode == "reference":
            assert not any(r["chol_ok"] for r in recs)
        else:
            assert all(r["chol_ok"] for r in recs)
