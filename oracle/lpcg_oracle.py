"""CPU oracle of the learned-preconditioner PCG (test infrastructure only).

Checker for the HIP implementation in s-pgo_amd/csrc/k_lpcg.hip: only tests,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU leg import it.  The
reference has no learned preconditioner (it factors the reduced system with
torch's Cholesky, linear_solvers.cu:195-218); this module restates the
build's own design in float64 numpy:

* systems: reduced pose systems H (n 6x6 blocks per side, block band of
  width W: H_{i,i-d} = 0 for d > W) with right-hand side g, as formed by the
  solver (Solver_cuda.cu:690-736 semantics), stored as ``Hb[i, d]`` =
  H_{i,i-d} (row-major 6x6) for d = 0..W;
* features: C_i, the inverse of the lower Cholesky factor of H_ii
  (C_i H_ii C_i^T = I); normalised off-diagonal blocks N_{i,d} = C_i H_{i,i-d}
  C_{i-d}^T with norms s_{i,d}; node input x_i = [N_{i,1} (36), s_{i,1},
  s_{i+1,1}, sum_j s_ij / W, 1];
* network (products with operands rounded to bfloat16 when ``bf`` is set, as
  the GPU's MFMA layers compute them): h0 = relu(W_in x + b_in); three message-passing layers
  h_{l+1} = relu(W_s h_l + W_m m_l + b_l) with m_i = sum_j s_ij h_j /
  (1 + sum_j s_ij) over the band neighbours j of i; head out = W_o h_3 +
  W_k x + b_o (57 values);
* preconditioner P = G^T G, G block lower bidiagonal: G_ii = (I + tril(out[0:21]))
  C_i, G_{i,i-1} = out[21:57] C_{i-1}.  With the head at zero P is the block
  Jacobi preconditioner diag(H_ii)^-1;
* loss per system: Kaporin's condition measure of P^1/2 H P^1/2 up to its
  constant, L = log(tr(G H G^T) / 6n) - log det(G^T G) / 6n with the
  constant parts of log det dropped.  In the normalised blocks
  tr(G H G^T) = sum_i tr(A_i A_i^T) + 2 tr(A_i N_{i,1} B_i^T) + tr(B_i B_i^T)
  and log det(G^T G) = 2 sum_i sum_k log |(A_i)_kk| + const, so the loss
  needs only per-node terms; its minimiser over this pattern is the factored
  sparse approximate inverse (Kolotilina-Yeremin) with one off-diagonal
  block, and block Jacobi scores 0.  The batch loss is the mean over systems.

Parity: ``loss_and_grad``'s gradient is pinned by central finite differences
of ``forward_loss`` (tests/test_lpcg_oracle.py); the features by the
normalisation identities (C_i H_ii C_i^T = I); PCG by ``numpy.linalg.solve``.
"""
from __future__ import annotations

import dataclasses

import numpy as np

HID = 64
NIN = 40
NOUT = 57
NLAYER = 3
TRIL = [(r, c) for r in range(6) for c in range(r + 1)]   # 21 lower-triangle entries, row-major


@dataclasses.dataclass
class Params:
    W_in: np.ndarray   # (HID, NIN)
    b_in: np.ndarray   # (HID,)
    W_s: list          # NLAYER x (HID, HID)
    W_m: list          # NLAYER x (HID, HID)
    b: list            # NLAYER x (HID,)
    W_o: np.ndarray    # (NOUT, HID)
    W_k: np.ndarray    # (NOUT, NIN)
    b_o: np.ndarray    # (NOUT,)

    def flat(self) -> np.ndarray:
        parts = [self.W_in, self.b_in]
        for l in range(NLAYER):
            parts += [self.W_s[l], self.W_m[l], self.b[l]]
        parts += [self.W_o, self.W_k, self.b_o]
        return np.concatenate([p.reshape(-1) for p in parts])

    @staticmethod
    def from_flat(v: np.ndarray) -> "Params":
        v = np.asarray(v, np.float64)
        o = 0

        def take(*shape):
            nonlocal o
            n = int(np.prod(shape))
            r = v[o:o + n].reshape(shape)
            o += n
            return r.copy()
        W_in = take(HID, NIN)
        b_in = take(HID)
        W_s, W_m, b = [], [], []
        for _ in range(NLAYER):
            W_s.append(take(HID, HID))
            W_m.append(take(HID, HID))
            b.append(take(HID))
        W_o = take(NOUT, HID)
        W_k = take(NOUT, NIN)
        b_o = take(NOUT)
        assert o == v.size
        return Params(W_in, b_in, W_s, W_m, b, W_o, W_k, b_o)


N_PARAMS = HID * NIN + HID + NLAYER * (2 * HID * HID + HID) + NOUT * HID + NOUT * NIN + NOUT


def band_from_dense(H: np.ndarray, W: int) -> np.ndarray:
    """Hb[i, d] = H_{i,i-d} (6x6) for d = 0..W (zero where i - d < 0)."""
    n = H.shape[0] // 6
    Hb = np.zeros((n, W + 1, 6, 6))
    for i in range(n):
        for d in range(min(W, i) + 1):
            Hb[i, d] = H[6 * i:6 * i + 6, 6 * (i - d):6 * (i - d) + 6]
    return Hb


def dense_from_band(Hb: np.ndarray) -> np.ndarray:
    n, W1 = Hb.shape[0], Hb.shape[1]
    H = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for d in range(min(W1 - 1, i) + 1):
            H[6 * i:6 * i + 6, 6 * (i - d):6 * (i - d) + 6] = Hb[i, d]
            H[6 * (i - d):6 * (i - d) + 6, 6 * i:6 * i + 6] = Hb[i, d].T
    return H


def features(Hb: np.ndarray):
    """C (n,6,6), s (n, W+1) with s[i, d] = ||N_{i,d}||_F (s[:, 0] = 0), N1 (n,6,6)
    = N_{i,1}, node inputs x (n, NIN), aggregation weights wsum (n,) = 1 + sum_j s_ij."""
    n, W1 = Hb.shape[0], Hb.shape[1]
    W = W1 - 1
    C = np.zeros((n, 6, 6))
    for i in range(n):
        Lc = np.linalg.cholesky(Hb[i, 0])
        C[i] = np.linalg.inv(Lc)
    s = np.zeros((n, W1))
    N1 = np.zeros((n, 6, 6))
    for i in range(n):
        for d in range(1, min(W, i) + 1):
            Nb = C[i] @ Hb[i, d] @ C[i - d].T
            s[i, d] = np.sqrt(np.sum(Nb * Nb))
            if d == 1:
                N1[i] = Nb
    tot = np.zeros(n)
    for i in range(n):
        for d in range(1, W + 1):
            if i - d >= 0:
                tot[i] += s[i, d]
            if i + d < n:
                tot[i] += s[i + d, d]
    x = np.zeros((n, NIN))
    x[:, :36] = N1.reshape(n, 36)
    x[:, 36] = s[:, 1] if W >= 1 else 0.0
    if W >= 1:
        x[:-1, 37] = s[1:, 1]
    x[:, 38] = tot / max(W, 1)
    x[:, 39] = 1.0
    return C, s, N1, x, 1.0 + tot


def aggregate(h: np.ndarray, s: np.ndarray, wsum: np.ndarray) -> np.ndarray:
    n, W1 = s.shape
    m = np.zeros_like(h)
    for i in range(n):
        for d in range(1, W1):
            if i - d >= 0:
                m[i] += s[i, d] * h[i - d]
            if i + d < n:
                m[i] += s[i + d, d] * h[i + d]
    return m / wsum[:, None]


def aggregate_T(dm: np.ndarray, s: np.ndarray, wsum: np.ndarray) -> np.ndarray:
    """Transpose of ``aggregate``: dh_j = sum_i s_ij / wsum_i dm_i."""
    n, W1 = s.shape
    q = dm / wsum[:, None]
    dh = np.zeros_like(dm)
    for j in range(n):
        for d in range(1, W1):
            if j + d < n:
                dh[j] += s[j + d, d] * q[j + d]
            if j - d >= 0:
                dh[j] += s[j, d] * q[j - d]
    return dh


def bf16(a) -> np.ndarray:
    """a rounded to bfloat16 (through float32, round to nearest even), as
    float64: the operand precision of the network's MFMA products on the GPU."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def network(P: Params, x: np.ndarray, s: np.ndarray, wsum: np.ndarray, bf: bool = False):
    """Forward pass; returns out (n, NOUT) and the cache for the backward pass.
    bf: the operands of every product rounded to bfloat16, as on the GPU."""
    r = bf16 if bf else _ident
    z0 = r(x) @ r(P.W_in).T + P.b_in
    h = [np.maximum(z0, 0.0)]
    zs, ms = [z0], []
    for l in range(NLAYER):
        m = aggregate(h[-1], s, wsum)
        z = r(h[-1]) @ r(P.W_s[l]).T + r(m) @ r(P.W_m[l]).T + P.b[l]
        ms.append(m)
        zs.append(z)
        h.append(np.maximum(z, 0.0))
    out = r(h[-1]) @ r(P.W_o).T + r(x) @ r(P.W_k).T + P.b_o
    return out, (h, zs, ms)


def g_blocks(out: np.ndarray, C: np.ndarray):
    """G_ii (n,6,6) and G_{i,i-1} (n,6,6; zero for i = 0) from the head output."""
    n = out.shape[0]
    A = np.tile(np.eye(6), (n, 1, 1))
    for k, (r, c) in enumerate(TRIL):
        A[:, r, c] += out[:, k]
    B = out[:, 21:57].reshape(n, 6, 6)
    Gd = A @ C
    Go = np.zeros((n, 6, 6))
    Go[1:] = B[1:] @ C[:-1]
    return Gd, Go, A, B


def apply_G(Gd, Go, r):
    """u = G r (r: (n,6))."""
    u = np.einsum('nij,nj->ni', Gd, r)
    u[1:] += np.einsum('nij,nj->ni', Go[1:], r[:-1])
    return u


def apply_GT(Gd, Go, u):
    """v = G^T u."""
    v = np.einsum('nji,nj->ni', Gd, u)
    v[:-1] += np.einsum('nji,nj->ni', Go[1:], u[1:])
    return v


def band_matvec(Hb, x):
    """y = H x for the block band Hb (x: (n,6))."""
    n, W1 = Hb.shape[0], Hb.shape[1]
    y = np.zeros_like(x)
    for i in range(n):
        y[i] += Hb[i, 0] @ x[i]
        for d in range(1, W1):
            if i - d >= 0:
                y[i] += Hb[i, d] @ x[i - d]
                y[i - d] += Hb[i, d].T @ x[i]
    return y


@dataclasses.dataclass
class System:
    Hb: np.ndarray   # (n, W+1, 6, 6)
    g: np.ndarray    # (n, 6)
    x: np.ndarray    # (n, 6) = H^-1 g


def make_system(H: np.ndarray, g: np.ndarray, W: int) -> System:
    Hb = band_from_dense(H, W)
    x = np.linalg.solve(dense_from_band(Hb), g.reshape(-1)).reshape(-1, 6)
    return System(Hb, g.reshape(-1, 6).copy(), x)


def system_loss(P: Params, Hb: np.ndarray, bf: bool = False):
    """Kaporin loss of one system and what its gradient needs."""
    C, s, N1, x, wsum = features(Hb)
    out, cache = network(P, x, s, wsum, bf)
    _, _, A, B = g_blocks(out, C)
    B = B.copy()
    B[0] = 0.0
    n6 = 6 * x.shape[0]
    t = (np.einsum('nij,nij->n', A, A) + 2.0 * np.einsum('nij,njk,nik->n', A, N1, B)
         + np.einsum('nij,nij->n', B, B))
    T = float(t.sum())
    dg = np.abs(np.diagonal(A, axis1=1, axis2=2))
    L = np.log(T / n6) - 2.0 / n6 * float(np.log(dg).sum())
    return L, (x, s, wsum, cache, A, B, N1, T, n6)


def forward_loss(P: Params, systems) -> float:
    """Mean loss over the systems (each a block band Hb, or an object with .Hb)."""
    return float(np.mean([system_loss(P, getattr(sy, "Hb", sy))[0] for sy in systems]))


def loss_and_grad(P: Params, systems, bf: bool = False):
    """Mean loss over the systems and its gradient (flat, like Params.flat()).
    bf: every product's operands rounded to bfloat16 (forward and backward),
    as the GPU's MFMA layers compute them."""
    rb = bf16 if bf else _ident
    S = len(systems)
    gW_in = np.zeros_like(P.W_in)
    gb_in = np.zeros_like(P.b_in)
    gW_s = [np.zeros_like(w) for w in P.W_s]
    gW_m = [np.zeros_like(w) for w in P.W_m]
    gb = [np.zeros_like(w) for w in P.b]
    gW_o = np.zeros_like(P.W_o)
    gW_k = np.zeros_like(P.W_k)
    gb_o = np.zeros_like(P.b_o)
    total = 0.0
    for sy in systems:
        L, (x, s, wsum, (h, zs, ms), A, B, N1, T, n6) = system_loss(P, getattr(sy, "Hb", sy), bf)
        total += L
        n = x.shape[0]
        # dt/dA = 2 A + 2 B N^T, dt/dB = 2 A N + 2 B; the log-det term on diag(A)
        dA = (2.0 * A + 2.0 * np.einsum('nij,nkj->nik', B, N1)) / T
        for k in range(6):
            dA[:, k, k] -= 2.0 / n6 / A[:, k, k]
        dB = (2.0 * np.einsum('nij,njk->nik', A, N1) + 2.0 * B) / T
        dB[0] = 0.0
        dout = np.zeros((n, NOUT))
        for k, (r, c) in enumerate(TRIL):
            dout[:, k] = dA[:, r, c]
        dout[:, 21:57] = dB.reshape(n, 36)
        dout /= S
        gW_o += rb(dout).T @ rb(h[-1])
        gW_k += rb(dout).T @ rb(x)
        gb_o += dout.sum(0)
        dh = rb(dout) @ rb(P.W_o)
        for l in reversed(range(NLAYER)):
            dz = dh * (zs[l + 1] > 0)
            gW_s[l] += rb(dz).T @ rb(h[l])
            gW_m[l] += rb(dz).T @ rb(ms[l])
            gb[l] += dz.sum(0)
            dm = rb(dz) @ rb(P.W_m[l])
            dh = rb(dz) @ rb(P.W_s[l]) + aggregate_T(dm, s, wsum)
        dz0 = dh * (zs[0] > 0)
        gW_in += rb(dz0).T @ rb(x)
        gb_in += dz0.sum(0)
    grad = Params(gW_in, gb_in, gW_s, gW_m, gb, gW_o, gW_k, gb_o).flat()
    return total / S, grad


def fsai1(Hb: np.ndarray):
    """The loss's minimiser over the pattern in closed form: G_ii, G_{i,i-1} of
    the factored sparse approximate inverse with one off-diagonal block."""
    n = Hb.shape[0]
    Gd = np.zeros((n, 6, 6))
    Go = np.zeros((n, 6, 6))
    for i in range(n):
        if i == 0:
            X = np.linalg.inv(Hb[0, 0])
            Gd[0] = np.linalg.solve(np.linalg.cholesky(X), X.T)
            continue
        M = np.block([[Hb[i - 1, 0], Hb[i, 1].T], [Hb[i, 1], Hb[i, 0]]])
        E = np.zeros((12, 6))
        E[6:] = np.eye(6)
        X = np.linalg.solve(M, E)
        Lc = np.linalg.cholesky(X[6:])
        R = np.linalg.solve(Lc, X.T)
        Go[i] = R[:, :6]
        Gd[i] = R[:, 6:]
    return Gd, Go


def init_params(seed: int) -> Params:
    """He-style initialisation of the hidden layers; the head starts at zero, so
    the untrained preconditioner is block Jacobi."""
    rng = np.random.default_rng(seed)
    W_in = rng.normal(0, np.sqrt(2.0 / NIN), (HID, NIN))
    W_s = [rng.normal(0, np.sqrt(1.0 / HID), (HID, HID)) for _ in range(NLAYER)]
    W_m = [rng.normal(0, np.sqrt(1.0 / HID), (HID, HID)) for _ in range(NLAYER)]
    return Params(W_in, np.zeros(HID), W_s, W_m, [np.zeros(HID) for _ in range(NLAYER)],
                  np.zeros((NOUT, HID)), np.zeros((NOUT, NIN)), np.zeros(NOUT))


def pcg(Hb, g, Gd=None, Go=None, C=None, tol=1e-8, maxit=5000):
    """Preconditioned CG on H x = g with P = G^T G (or block Jacobi from C when
    Gd is None).  Returns (x, iterations, relative residual)."""
    if Gd is None:
        Gd = C.copy()
        Go = np.zeros_like(C)

    def prec(r):
        return apply_GT(Gd, Go, apply_G(Gd, Go, r))
    x = np.zeros_like(g)
    r = g.copy()
    z = prec(r)
    p = z.copy()
    rz = float(np.sum(r * z))
    g0 = np.sqrt(float(np.sum(g * g)))
    for it in range(maxit):
        Hp = band_matvec(Hb, p)
        a = rz / float(np.sum(p * Hp))
        x += a * p
        r -= a * Hp
        rn = np.sqrt(float(np.sum(r * r)))
        if rn <= tol * g0:
            return x, it + 1, rn / g0
        z = prec(r)
        rz2 = float(np.sum(r * z))
        p = z + (rz2 / rz) * p
        rz = rz2
    return x, maxit, rn / g0
