"""Pins of the learned-preconditioner oracle (CPU): finite-difference gradient,
normalisation identities, block-Jacobi limit, PCG against a dense solve."""
import numpy as np
import pytest

from oracle import ba_oracle as O
from oracle import lpcg_oracle as LP
from spgo.problems import make_problem


def _systems(count=2, N=9, L=300, K=4, W=4, seed=5):
    out = []
    for k in range(count):
        p = make_problem(N, L, K, seed + k)
        pa = O.from_problem(p)
        X0 = O.project_rotations(p.X0.astype(np.float64))
        lin = O.linearize(pa, X0, p.alpha0.astype(np.float64), O.Options())
        H, g = O.reduced_system(lin, 1e-4, 1e-4)
        # the band holds every nonzero block: spans of K observations couple
        # poses at most K - 2 apart
        assert np.all(H[6 * (K - 1):, :6] == 0) if N > K else True
        out.append(LP.make_system(H, g, W))
    return out


def test_band_round_trip_and_feature_identities():
    sy = _systems(1)[0]
    H = LP.dense_from_band(sy.Hb)
    assert np.allclose(LP.band_from_dense(H, sy.Hb.shape[1] - 1), sy.Hb)
    C, s, N1, x, wsum = LP.features(sy.Hb)
    for i in range(C.shape[0]):
        assert np.allclose(C[i] @ sy.Hb[i, 0] @ C[i].T, np.eye(6), atol=1e-10)
        assert np.allclose(np.triu(C[i], 1), 0)
    assert np.allclose(N1[1], C[1] @ sy.Hb[1, 1] @ C[0].T)
    assert s[3, 2] == pytest.approx(np.linalg.norm(C[3] @ sy.Hb[3, 2] @ C[1].T))
    assert np.all(x[:, 39] == 1.0) and np.all(wsum >= 1.0)
    # matvec against the dense matrix
    v = np.random.default_rng(0).normal(size=sy.g.shape)
    assert np.allclose(LP.band_matvec(sy.Hb, v).reshape(-1), H @ v.reshape(-1))


def test_untrained_network_is_block_jacobi_and_pcg_solves():
    sy = _systems(1)[0]
    P = LP.init_params(3)
    C, s, N1, x, wsum = LP.features(sy.Hb)
    out, _ = LP.network(P, x, s, wsum)
    assert np.all(out == 0)
    Gd, Go, A, B = LP.g_blocks(out, C)
    H = LP.dense_from_band(sy.Hb)
    # P = G^T G equals blockdiag(H_ii)^-1
    n = C.shape[0]
    Pm = np.zeros_like(H)
    for i in range(n):
        Pm[6 * i:6 * i + 6, 6 * i:6 * i + 6] = np.linalg.inv(sy.Hb[i, 0])
    r = np.random.default_rng(1).normal(size=sy.g.shape)
    assert np.allclose(LP.apply_GT(Gd, Go, LP.apply_G(Gd, Go, r)).reshape(-1), Pm @ r.reshape(-1))
    xs, it, rel = LP.pcg(sy.Hb, sy.g, Gd, Go, tol=1e-10)
    assert rel <= 1e-10
    assert np.allclose(xs, sy.x, rtol=1e-7, atol=1e-9 * np.max(np.abs(sy.x)))
    # the G-transpose apply is the transpose of the G apply
    u = np.random.default_rng(2).normal(size=sy.g.shape)
    Gm = np.zeros_like(H)
    for i in range(n):
        Gm[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Gd[i]
        if i:
            Gm[6 * i:6 * i + 6, 6 * i - 6:6 * i] = Go[i]
    assert np.allclose(LP.apply_G(Gd, Go, r).reshape(-1), Gm @ r.reshape(-1))
    assert np.allclose(LP.apply_GT(Gd, Go, u).reshape(-1), Gm.T @ u.reshape(-1))


def test_gradient_matches_central_differences():
    systems = _systems(2, N=7, L=200, K=3, W=3, seed=9)
    rng = np.random.default_rng(4)
    P = LP.init_params(7)
    v = P.flat()
    v += rng.normal(0, 0.05, v.size)          # a head away from zero: every path live
    P = LP.Params.from_flat(v)
    L, grad = LP.loss_and_grad(P, systems)
    assert L == pytest.approx(LP.forward_loss(P, systems), rel=1e-12)
    idx = rng.choice(v.size, 60, replace=False)
    # the last parameters are the head: make sure some are checked
    idx = np.concatenate([idx, np.arange(v.size - 5, v.size)])
    h = 1e-6
    for k in idx:
        vp = v.copy()
        vp[k] += h
        vm = v.copy()
        vm[k] -= h
        fd = (LP.forward_loss(LP.Params.from_flat(vp), systems) - LP.forward_loss(LP.Params.from_flat(vm), systems)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])


def _dense_G(Gd, Go):
    n = Gd.shape[0]
    G = np.zeros((6 * n, 6 * n))
    for i in range(n):
        G[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Gd[i]
        if i:
            G[6 * i:6 * i + 6, 6 * i - 6:6 * i] = Go[i]
    return G


def test_loss_matches_dense_kaporin_definition_and_fsai_minimises_it():
    sy = _systems(1, N=10, L=500, K=4, W=3, seed=31)[0]
    H = LP.dense_from_band(sy.Hb)
    n6 = H.shape[0]
    P = LP.init_params(2)
    v = P.flat()
    v[-(LP.NOUT * LP.HID + LP.NOUT * LP.NIN + LP.NOUT):] = np.random.default_rng(6).normal(
        0, 0.02, LP.NOUT * LP.HID + LP.NOUT * LP.NIN + LP.NOUT)
    P = LP.Params.from_flat(v)
    C, s, N1, x, wsum = LP.features(sy.Hb)
    out, _ = LP.network(P, x, s, wsum)
    Gd, Go, _, _ = LP.g_blocks(out, C)
    G = _dense_G(Gd, Go)
    # Kaporin: log(tr(P H)/n) - log det(P H)/n, minus its value at block Jacobi
    # (G = blockdiag C_i), which cancels log det H and the C_i
    M = G @ H @ G.T
    direct = np.log(np.trace(M) / n6) - np.linalg.slogdet(M)[1] / n6
    Gj = _dense_G(C, np.zeros_like(C))
    Mj = Gj @ H @ Gj.T
    direct -= np.log(np.trace(Mj) / n6) - np.linalg.slogdet(Mj)[1] / n6
    assert LP.system_loss(P, sy.Hb)[0] == pytest.approx(direct, rel=1e-9, abs=1e-12)
    # the closed-form FSAI(1) blocks score below block Jacobi (0) and below the
    # random head
    Fd, Fo = LP.fsai1(sy.Hb)
    Mf = _dense_G(Fd, Fo) @ H @ _dense_G(Fd, Fo).T
    kf = np.log(np.trace(Mf) / n6) - np.linalg.slogdet(Mf)[1] / n6 - (
        np.log(np.trace(Mj) / n6) - np.linalg.slogdet(Mj)[1] / n6)
    assert kf < 0.0 and kf < direct


def test_fsai_like_head_beats_block_jacobi():
    # the head can express the first-order factored approximate inverse:
    # A = I, B_i = -N_{i,1} through the skip weights W_k.  That preconditioner
    # needs fewer PCG iterations than block Jacobi on a solver-built system,
    # so the family the network searches contains real gains.
    sy = _systems(1, N=24, L=1500, K=6, W=6, seed=21)[0]
    P = LP.init_params(1)
    for k in range(36):
        P.W_k[21 + k, k] = -1.0
    C, s, N1, x, wsum = LP.features(sy.Hb)
    out, _ = LP.network(P, x, s, wsum)
    Gd, Go, _, _ = LP.g_blocks(out, C)
    _, it_l, _ = LP.pcg(sy.Hb, sy.g, Gd, Go)
    _, it_j, _ = LP.pcg(sy.Hb, sy.g, C=C)
    print(f"[fsai-head] block Jacobi {it_j} iterations, first-order factored inverse {it_l}")
    assert it_l < it_j


def test_bfloat16_rounding_known_answers():
    # round to nearest even on the 8-bit significand of bfloat16
    one = 1.0
    assert LP.bf16(one) == 1.0
    assert LP.bf16(1.0 + 2.0 ** -8) == 1.0                    # tie -> even (1.0)
    assert LP.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6     # tie -> even (1 + 2^-6)
    assert LP.bf16(1.0 + 2.0 ** -8 + 2.0 ** -12) == 1.0 + 2.0 ** -7
    assert LP.bf16(-3.0e-3) == -float(np.float32(LP.bf16(3.0e-3)))
    x = np.random.default_rng(0).normal(size=1000)
    assert np.max(np.abs(LP.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_bfloat16_network_stays_close_to_the_exact_one():
    rng = np.random.default_rng(4)
    P = LP.init_params(3)
    P.W_o = rng.normal(size=P.W_o.shape) * 0.05
    n = 30
    Hb = np.zeros((n, 3, 6, 6))
    for i in range(n):
        M = rng.normal(size=(6, 6))
        Hb[i, 0] = M @ M.T + 6 * np.eye(6)
        for d in (1, 2):
            if i - d >= 0:
                Hb[i, d] = 0.3 * rng.normal(size=(6, 6))
    C, s, N1, x, wsum = LP.features(Hb)
    out, _ = LP.network(P, x, s, wsum)
    outb, _ = LP.network(P, x, s, wsum, bf=True)
    rel = np.linalg.norm(outb - out) / np.linalg.norm(out)
    assert 1e-5 < rel < 2e-2
    L, g = LP.loss_and_grad(P, [Hb])
    Lb, gb = LP.loss_and_grad(P, [Hb], bf=True)
    assert abs(Lb - L) < 1e-2 * abs(L) + 1e-3
    assert np.linalg.norm(gb - g) < 3e-2 * np.linalg.norm(g)
