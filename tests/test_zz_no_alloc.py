"""spgo_reduced_band_noalloc (the band staged in the library's scratch array,
no device allocation) against spgo_reduced_band (a buffer taken per call):
identical bands, right-hand sides and costs on two calls in a row, the second
with a smaller problem and band; sizes past the scratch capacity and solvers
without device state are refused."""
import numpy as np
import pytest

from spgo import _lib
from spgo.problems import make_problem, make_revisit_problem
from spgo.solver import Solver, make_options, solver_from_problem


@pytest.mark.gpu
def test_noalloc_band_matches_the_allocating_call_twice():
    cases = [(make_revisit_problem(40, 1500, 6, seed=7, revisit_frac=0.25), 11, 1e-3),
             (make_problem(8, 300, 4, seed=3), 3, 0.5)]
    for i, (p, W, lam) in enumerate(cases):
        print(f"[noalloc] call {i + 1}: N={p.n_poses} W={W}", flush=True)
        s = solver_from_problem(p, make_options(mode="lm"))
        s.step(2)
        Hb0, g0, c0 = s.reduced_band(W, lam)
        Hb1, g1, c1 = s.reduced_band_noalloc(W, lam)
        assert np.array_equal(Hb0, Hb1)
        assert np.array_equal(g0, g1)
        assert c0 == c1
        assert np.max(np.abs(Hb1)) > 0
        s.close()


def test_noalloc_band_refusals_on_host():
    s = Solver(2048, 1)
    with pytest.raises(RuntimeError, match="scratch capacity"):
        s.reduced_band_noalloc(16, 0.0)
    s2 = Solver(8, 2)
    with pytest.raises(RuntimeError, match="device state"):
        s2.reduced_band_noalloc(3, 0.0)
