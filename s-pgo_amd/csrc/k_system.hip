// Reduced pose system, its Cholesky factorisation and the triangular solves
// (the work of Solver_cuda.cu:690-736 and linear_solvers.cu:195-218).
//
// k_form_system sums, for every element of the lower triangle, the group
// blocks D (already in incremental-pose coordinates) of the landmark groups
// whose pose span covers it:
//     H_ij = sum_a Ad(G_a^-1 G_i)^T D_a(i, j) Ad(G_a^-1 G_j) + lambda [i = j]
// (lambda: the damping the reference defines but never launches,
// Solver_cuda.cu:348-359).  The factorisation is a blocked right-looking
// Cholesky with NB x NB blocks: one kernel factors the diagonal block in
// registers of one wavefront (and inverts it), the other blocks of the panel
// are then products with that inverse, and a second kernel applies the
// trailing update on MFMA.
#include "spgo_device.h"
#include "spgo_kernels.h"

namespace spgo {

namespace {
typedef double __attribute__((ext_vector_type(4))) d4;

__device__ __forceinline__ double readlane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
}  // namespace

// ---------------------------------------------------------------------------
// reduced system
// ---------------------------------------------------------------------------
constexpr int FS_MAXG = 512;   // groups whose span can reach one pose row

template <typename T>
__global__ __launch_bounds__(256) void k_form_system(Layout lay, StateBuf<T> B0, StateBuf<T> B1, Control* ctl,
                                                     double* H, double* gvec, int Pp) {
    __shared__ long long goff[FS_MAXG];
    __shared__ int ganc[FS_MAXG], gK[FS_MAXG], gdoff[FS_MAXG];
    const int row = blockIdx.x;
    const int tid = threadIdx.x;
    const int P = 6 * lay.N;
    if (row == 0 && tid == 0) ctl->chol_ok = 1;   // reset for this iteration's factorisation
    const StateBuf<T>& bs = ctl->cur ? B1 : B0;
    const int bi = row / 6;
    const int amin = max(0, bi - lay.kmax + 1);
    // groups of anchors amin..bi (contiguous, sorted by anchor); K = 0 marks
    // groups whose span does not reach pose bi.  The host keeps their number
    // within FS_MAXG.
    const int g0 = lay.anchor_grp[amin];
    const int n = min(lay.anchor_grp[bi + 1] - g0, FS_MAXG);
    for (int e = tid; e < n; e += 256) {
        const int g = g0 + e;
        const int a = lay.grp_anchor[g], K = lay.grp_K[g];
        ganc[e] = a;
        gK[e] = (bi - a < K) ? K : 0;
        goff[e] = lay.grp_D_off[g];
        gdoff[e] = lay.grp_d_off[g];
    }
    __syncthreads();
    // compact the covering groups in group order (anchors ascending), with the
    // base offset of row `row` inside each group's block
    __shared__ long long rowoff[FS_MAXG];
    __shared__ int canc[FS_MAXG], cK6[FS_MAXG];
    __shared__ int ncov;
    if (tid == 0) {
        int c = 0;
        for (int e = 0; e < n; ++e) {
            if (gK[e] == 0) continue;
            const int a = ganc[e], K6 = 6 * gK[e];
            canc[c] = a;
            cK6[c] = K6;
            rowoff[c] = goff[e] + (long long)(row - 6 * a) * K6 - 6 * a;
            gdoff[c] = gdoff[e] + row - 6 * a;
            ++c;
        }
        ncov = c;
    }
    __syncthreads();
    const int nc = ncov;
    for (int col = tid; col <= row; col += 256) {
        const int bj = col / 6;
        int ne = 0;
        while (ne < nc && canc[ne] <= bj) ++ne;   // covering groups with anchor <= bj: a prefix
        double h0 = 0.0, h1 = 0.0;
        int e = 0;
        for (; e + 1 < ne; e += 2) {
            h0 += (double)bs.D[rowoff[e] + col];
            h1 += (double)bs.D[rowoff[e + 1] + col];
        }
        if (e < ne) h0 += (double)bs.D[rowoff[e] + col];
        double h = h0 + h1;
        if (col == row) {
            h += ctl->lambda;
            double gs = 0.0;
            for (int q = 0; q < ne; ++q) gs += (double)bs.d[gdoff[q]];
            gvec[row] = gs;
        }
        H[(long long)row * Pp + col] = h;
    }
    (void)P;
}

template <typename T>
void launch_form_system(const Layout& lay, const StateBuf<T>* bufs, Control* ctl, double* H, double* g,
                        int Pp, hipStream_t s) {
    const int P = 6 * lay.N;
    hipLaunchKernelGGL(k_form_system<T>, dim3(P), dim3(256), 0, s, lay, bufs[0], bufs[1], ctl, H, g, Pp);
}
template void launch_form_system<double>(const Layout&, const StateBuf<double>*, Control*, double*, double*,
                                         int, hipStream_t);
template void launch_form_system<float>(const Layout&, const StateBuf<float>*, Control*, double*, double*,
                                        int, hipStream_t);

// ---------------------------------------------------------------------------
// blocked Cholesky, block column k
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chol_panel(double* H, double* Linv, int Pp, int k, Control* ctl) {
    __shared__ double Lkk[NB][NB + 1];
    __shared__ double Li[NB][NB + 1];
    __shared__ double Ab[NB][NB + 1];
    const int tid = threadIdx.x;
    const int k0 = k * NB;
    if (tid < 64) {
        // one wavefront: lane i holds row i of the diagonal block
        const int i = tid & (NB - 1);
        double a[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = H[(long long)(k0 + i) * Pp + k0 + j];
        bool bad = false;
#pragma clang loop unroll(full)
        for (int j = 0; j < NB; ++j) {
            double piv = readlane_d(a[j], j);
            if (!(piv > 0.0) || !isfinite(piv)) {
                bad = true;
                piv = 1.0;
            }
            const double ljj = sqrt(piv);
            const double inv = 1.0 / ljj;
            a[j] = (i == j) ? ljj : a[j] * inv;
#pragma clang loop unroll(full)
            for (int q = j + 1; q < NB; ++q) {
                const double lq = readlane_d(a[j], q);
                a[q] -= a[j] * lq;
            }
        }
        if (tid < NB) {
#pragma unroll
            for (int j = 0; j < NB; ++j) Lkk[i][j] = (j <= i) ? a[j] : 0.0;
        }
        if (bad && blockIdx.x == 0 && tid == 0) ctl->chol_ok = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (tid < NB) {
            // column tid of L^-1 by forward substitution
            const int cidx = tid;
            double x[NB];
#pragma clang loop unroll(full)
            for (int rr = 0; rr < NB; ++rr) {
                double s = (rr == cidx) ? 1.0 : 0.0;
#pragma clang loop unroll(full)
                for (int m = 0; m < rr; ++m) s -= Lkk[rr][m] * x[m];
                x[rr] = s / Lkk[rr][rr];
            }
#pragma unroll
            for (int rr = 0; rr < NB; ++rr) Li[rr][cidx] = x[rr];
        }
    }
    __syncthreads();
    if (blockIdx.x == 0) {
        for (int e = tid; e < NB * NB; e += 256) {
            const int rr = e / NB, cc = e % NB;
            if (cc <= rr) H[(long long)(k0 + rr) * Pp + k0 + cc] = Lkk[rr][cc];
            Linv[(long long)k * NB * NB + e] = Li[rr][cc];
        }
        return;
    }
    const int b0 = k0 + blockIdx.x * NB;
    for (int e = tid; e < NB * NB; e += 256) {
        const int rr = e / NB, cc = e % NB;
        Ab[rr][cc] = H[(long long)(b0 + rr) * Pp + k0 + cc];
    }
    __syncthreads();
    // L_bk = A_bk L_kk^-T
    for (int e = tid; e < NB * NB; e += 256) {
        const int rr = e / NB, cc = e % NB;
        double s = 0.0;
        for (int m = 0; m <= cc; ++m) s += Ab[rr][m] * Li[cc][m];
        H[(long long)(b0 + rr) * Pp + k0 + cc] = s;
    }
}

// Trailing update A_ij -= L_ik L_jk^T for k < j <= i, one NB x NB tile per
// workgroup, four 16x16 MFMA tiles per wavefront-quarter of the block.
__global__ __launch_bounds__(256) void k_chol_update(double* H, int Pp, int k, int nblk) {
    const int n = nblk - k - 1;
    const int t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while (ti * (ti + 1) / 2 > t) --ti;
    const int tj = t - ti * (ti + 1) / 2;
    if (ti >= n) return;
    const int bi = k + 1 + ti, bj = k + 1 + tj;
    __shared__ double Ai[NB][NB + 2];
    __shared__ double Aj[NB][NB + 2];
    const int tid = threadIdx.x;
    const int k0 = k * NB;
    for (int e = tid; e < NB * NB; e += 256) {
        const int rr = e / NB, cc = e % NB;
        Ai[rr][cc] = H[(long long)(bi * NB + rr) * Pp + k0 + cc];
        Aj[rr][cc] = H[(long long)(bj * NB + rr) * Pp + k0 + cc];
    }
    __syncthreads();
    // wave w computes the 16x16 sub-tile (w / 2, w % 2) of the 32x32 tile
    const int w = tid >> 6, lane = tid & 63;
    const int sr = (w >> 1) * 16, sc = (w & 1) * 16;
    d4 acc = 0.0;
#pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
        const double av = Ai[sr + (lane & 15)][kk + (lane >> 4)];
        const double bv = Aj[sc + (lane & 15)][kk + (lane >> 4)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int rr = sr + mfma16_row(0.0, lane, v), cc = sc + (lane & 15);
        H[(long long)(bi * NB + rr) * Pp + bj * NB + cc] -= acc[v];
    }
}

void launch_cholesky(double* H, double* Linv, int Pp, Control* ctl, hipStream_t s) {
    const int nblk = Pp / NB;
    for (int k = 0; k < nblk; ++k) {
        hipLaunchKernelGGL(k_chol_panel, dim3(nblk - k), dim3(256), 0, s, H, Linv, Pp, k, ctl);
        const int n = nblk - k - 1;
        if (n > 0) hipLaunchKernelGGL(k_chol_update, dim3(n * (n + 1) / 2), dim3(256), 0, s, H, Pp, k, nblk);
    }
}

// ---------------------------------------------------------------------------
// Envelope (profile) Cholesky and both triangular solves in one workgroup.
//
// H is block-banded with EB x EB blocks: block row i has nonzero blocks only
// in columns first[i] .. i (first[] nondecreasing, i - first[i] <= D), and
// the Cholesky factor keeps that envelope.  The blocks that are still to be
// updated live in a sliding LDS window of W = D + 3 block rows; block (i, j)
// sits in the slot of the unordered pair (i mod W, j mod W), which is unique
// among the resident rows.  Right-looking over block columns k, two barriers
// per column:
//   phase A  wavefront 0 (and 1, 2 for profiles deeper than 3 blocks) factors
//            the panel of column k from registers, one row per lane: lanes
//            0-15 the diagonal block, lanes 16-63 three blocks below it (a
//            second or third wavefront repeats the diagonal block for its own
//            three blocks, so no wavefront waits on another inside the column
//            loop).  The right-hand side rides along as one more column, so
//            the forward solve is done with the factorisation.  The panel is
//            written to LDS only.  Meanwhile the other wavefronts apply the
//            part of the previous column's trailing update that the next
//            panel does not need, and wavefront 3 moves the block row it
//            prefetched into the window, inverts the previous diagonal
//            factor, copies the previous panel to global memory (both for
//            the backward solve) and prefetches the next block row.
//   phase B  the trailing update of block column k + 1 (the next panel) on
//            MFMA, one 16x16 tile per wavefront.
// The backward solve
// x_k = L_kk^-T (y_k - sum_i L_(k+i),k^T x_(k+i)) then runs block by block from
// the bottom as products with the stored inverses, its loads one block ahead.
// ---------------------------------------------------------------------------
constexpr int EB = 16;            // block size: one MFMA tile
constexpr int EDMAX = 7;          // deepest profile: blocks below the diagonal
constexpr int EST = EB + 1;       // LDS row stride of a window block
constexpr int EBLK = EB * EST;    // doubles per window block

// 1/sqrt(x) in double: the hardware estimate refined by one Halley step
// (cubic convergence), a few FMAs instead of the library's long sequence.
__device__ __forceinline__ double rsqrt_fast(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

// keep the compiler from reordering the LDS accesses of one wavefront (the
// hardware executes a wavefront's LDS operations in order)
__device__ __forceinline__ void wave_lds_order() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }

// workgroup barrier that orders LDS only: it does not wait for outstanding
// global loads or stores (s_waitcnt lgkmcnt(0), vmcnt/expcnt left at their maximum)
__device__ __forceinline__ void lds_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__host__ __device__ constexpr int env_window_slots(int D) { return (D + 3) * (D + 4) / 2; }

// window slot of block (i, j) from the residues a = i mod W, b = j mod W
__device__ __forceinline__ double* env_slot(double* win, int a, int b) {
    const int hi = a > b ? a : b, lo = a > b ? b : a;
    return win + (hi * (hi + 1) / 2 + lo) * EBLK;
}

// C(bi, bj) -= L(bi, kp) L(bj, kp)^T on one 16x16 MFMA tile (one wavefront);
// arguments are window residues
__device__ __forceinline__ void env_tile(double* win, int ri, int rj, int rk, int lane) {
    const double* Ai = env_slot(win, ri, rk) + (lane & 15) * EST + (lane >> 4);
    const double* Aj = env_slot(win, rj, rk) + (lane & 15) * EST + (lane >> 4);
    double* C = env_slot(win, ri, rj) + (lane & 15);
    double av[4], bv[4], cv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        av[s] = Ai[4 * s];
        bv[s] = Aj[4 * s];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) cv[v] = C[mfma16_row(0.0, lane, v) * EST];
    d4 acc = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s], bv[s], acc, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) C[mfma16_row(0.0, lane, v) * EST] = cv[v] - acc[v];
}

// Inverse of the lower-triangular diagonal factor held column-major (stride
// EST) in LDS, with its inverse diagonal dinv: lane c < 16 forms column c of
// L^-1 by forward substitution and stores it row-major to dst (16 x 16).
__device__ __forceinline__ void env_invert(const double* Lc, const double* dinv, double* dst, int lane) {
    const int c = lane & 15;
    double s[EB];
#pragma unroll
    for (int r = 0; r < EB; ++r) s[r] = (r == c) ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < EB; ++q) {
        s[q] *= dinv[q];
#pragma unroll
        for (int r = q + 1; r < EB; ++r) s[r] -= Lc[q * EST + r] * s[q];
    }
    if (lane < EB) {
#pragma unroll
        for (int r = 0; r < EB; ++r) dst[r * EB + c] = s[r];
    }
}

template <int DT>
__global__ __launch_bounds__(256) void k_chol_env(double* H, int Pp, int D, const int* first, const double* g,
                                                  double* x, double* winv, Control* ctl) {
    constexpr int EPRE = 4 * (DT + 1);      // prefetched doubles per lane of one block row
    extern __shared__ double lds[];
    const int W = D + 3;
    const int nb = Pp / EB;
    double* win = lds;                                  // env_window_slots(D) blocks
    double* ring = win + env_window_slots(D) * EBLK;    // last two diagonal factors, column-major
    double* xs = ring + 2 * EBLK;                       // right-hand side, then solution (Pp)
    double* dv = xs + Pp;                               // 1 / L_ii (Pp)
    int* mk = (int*)(dv + Pp);                          // blocks below the diagonal per block column
    int* fst = mk + nb;                                 // first[] staged
    __shared__ int sbad;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    auto wrap = [W](int r) { return r >= W ? r - W : r; };

    for (int i = tid; i < Pp; i += 256) xs[i] = g[i];
    for (int i = tid; i < nb; i += 256) fst[i] = first[i];
    if (tid == 0) sbad = 0;
    __syncthreads();
    for (int k = tid; k < nb; k += 256) {
        int m = 0;
        while (m < D && k + 1 + m < nb && fst[k + 1 + m] <= k) ++m;
        mk[k] = m;
    }
    // rows 0 .. D of the window; every thread owns one element of each block
    {
        const int r = tid >> 4, c = tid & 15;
        double v[(DT + 1) * (DT + 2) / 2];
        int n = 0;
#pragma unroll
        for (int i = 0; i <= DT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const bool use = i <= D && i < nb && j >= fst[min(i, nb - 1)];
                v[n++] = use ? H[(long long)(EB * i + r) * Pp + EB * j + c] : 0.0;
            }
        n = 0;
#pragma unroll
        for (int i = 0; i <= DT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const bool use = i <= D && i < nb && j >= fst[min(i, nb - 1)];
                if (use) env_slot(win, i, j)[r * EST + c] = v[n];
                ++n;
            }
    }
    // wavefront 3 keeps the next block row in registers until it has a slot.
    // Every lane loads unconditionally (the block index is clamped to the row)
    // and only the stores to the window are predicated.
    double pre[EPRE];
    auto prefetch = [&](int i) {
        const int ic = min(i, nb - 1);
        const int f0 = fst[ic];
#pragma unroll
        for (int t = 0; t < EPRE; ++t) {
            const int e = lane + 64 * t, b = min(e >> 8, ic - f0), rr = (e >> 4) & 15, cc = e & 15;
            pre[t] = H[(long long)(EB * ic + rr) * Pp + EB * (f0 + b) + cc];
        }
    };
    if (w == 3) prefetch(D + 1);
    __syncthreads();

    int rk = 0;   // k mod W
    for (int k = 0; k < nb; ++k, rk = wrap(rk + 1)) {
        const int m = mk[k];
        const int nf = m > 3 ? (m + 2) / 3 : 1;         // factoring wavefronts
        // ---------------- phase A
        if (w < nf) {
            const int bi = (lane < EB) ? 0 : 1 + 3 * w + ((lane - EB) >> 4);
            const int rr = lane & 15;
            const bool valid = bi <= m;
            const int bic = valid ? bi : 0;
            double a[EB];
            {
                const double* src = env_slot(win, wrap(rk + bic), rk) + rr * EST;
#pragma unroll
                for (int c = 0; c < EB; ++c) a[c] = valid ? src[c] : 0.0;
            }
            double b = valid ? xs[EB * (k + bic) + rr] : 0.0;
            double yv = 0.0, iv = 0.0;
            bool bad = false;
            int rl = lane;
            asm volatile("" : "+v"(rl));
#pragma unroll
            for (int j = 0; j < EB; ++j) {
                // column j: u holds A(:, j) after the updates of columns < j.
                // For lane j, u * inv = sqrt(pivot) and sc = 1, so the update
                // clobbers only entries of rows <= j that are already final
                // (their y_j and 1/L_jj are kept in yv, iv).
                const double u = a[j];
                const double piv = readlane_d(u, j);
                bad |= !(piv > 0.0 && piv < __builtin_inf());
                const double inv = rsqrt_fast(piv);
                const double bj = readlane_d(b, j);
                const double l = u * inv;
                const double sc = l * inv;
                if (rl == j) {
                    yv = bj * inv;
                    iv = inv;
                }
                a[j] = l;
                b -= sc * bj;
#pragma unroll
                for (int q = j + 1; q < EB; ++q) a[q] -= sc * readlane_d(u, q);
            }
            // the panel stays in LDS (the diagonal factor column-major in the
            // ring); wavefront 3 copies it to global memory in the next phase A,
            // so this wavefront never waits on global stores
            if (valid) {
                if (bi == 0) {
                    if (w == 0) {
                        double* rg = ring + (k & 1) * EBLK + rr;
#pragma unroll
                        for (int c = 0; c < EB; ++c) rg[c * EST] = (c <= rr) ? a[c] : 0.0;
                        dv[EB * k + rr] = iv;
                        xs[EB * k + rr] = yv;
                    }
                } else {
                    double* wl = env_slot(win, wrap(rk + bi), rk) + rr * EST;
#pragma unroll
                    for (int c = 0; c < EB; ++c) wl[c] = a[c];
                    xs[EB * (k + bi) + rr] = b;
                }
            }
            if (bad && lane == 0) sbad = 1;
        } else {
            // wavefront 3: the prefetched block row k + D + 1 into its slot (row
            // k - 2's, free since the last phase A)
            const int inew = k + D + 1;
            if (w == 3 && inew < nb) {
                const int f0 = fst[inew], nbl = inew - f0 + 1, ri = wrap(rk + D + 1);
#pragma unroll
                for (int t = 0; t < EPRE; ++t) {
                    const int e = lane + 64 * t, bb = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
                    if (bb < nbl) env_slot(win, ri, wrap(rk + f0 + bb - k))[rr * EST + cc] = pre[t];
                }
            }
            if (k >= 1) {
                const int kp = k - 1, mp = mk[kp], rkp = rk == 0 ? W - 1 : rk - 1;
                // rest of column kp's trailing update: tiles (kp+i, kp+j), 2 <= j <= i <= mp
                const int nt = (mp - 1) * mp / 2;
                for (int t = w - nf; t < nt; t += 4 - nf) {
                    int i = 2;
                    while ((i - 1) * i / 2 <= t) ++i;
                    const int j = 2 + t - (i - 2) * (i - 1) / 2;
                    env_tile(win, wrap(rkp + i), wrap(rkp + j), rkp, lane);
                }
                if (w == 3) {
                    // panel kp to global memory for the backward solve: the
                    // inverse of its diagonal factor and the blocks below it
                    env_invert(ring + (kp & 1) * EBLK, dv + EB * kp, winv + (long long)kp * EB * EB, lane);
                    for (int i = 1; i <= mp; ++i) {
                        const double* src = env_slot(win, wrap(rkp + i), rkp);
                        double* dst = H + (long long)EB * (kp + i) * Pp + EB * kp;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const int e = lane + 64 * t, rr = e >> 4, cc = e & 15;
                            dst[(long long)rr * Pp + cc] = src[rr * EST + cc];
                        }
                    }
                }
            }
            // wavefront 3: prefetch block row k + D + 2 (global stores and
            // loads are issued last, so no wavefront waits on them in this phase)
            if (w == 3 && inew + 1 < nb) prefetch(inew + 1);
        }
        lds_barrier();
        // ---------------- phase B: tiles (k + i, k + 1), 1 <= i <= m
        for (int t = w; t < m; t += 4) env_tile(win, wrap(rk + 1 + t), wrap(rk + 1), rk, lane);
        lds_barrier();
    }
    if (w == 3) env_invert(ring + ((nb - 1) & 1) * EBLK, dv + EB * (nb - 1), winv + (long long)(nb - 1) * EB * EB, lane);
    __syncthreads();   // the factor and the inverses in global memory are complete

    // backward solve by wavefront 0.  Lane (c, g) sums rows g, g+4, g+8, g+12
    // of column c over the blocks below the diagonal; z = y_k - that sum goes
    // through LDS so each lane can form its share of x_k = L_kk^-T z from its
    // four entries of column c of L_kk^-1.  The next block's loads are issued
    // before the current block is solved.
    if (w == 0) {
        const int c = lane & 15, g4 = lane >> 4;
        double pl[2][4 * DT], wl[2][4];
        auto load = [&](int k, double* P, double* Wc) {
            const int kc = max(k, 0);
            const int m = mk[kc];
#pragma unroll
            for (int i = 1; i <= DT; ++i)
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    P[4 * (i - 1) + t] = H[(long long)(EB * (kc + min(i, m)) + g4 + 4 * t) * Pp + EB * kc + c];
#pragma unroll
            for (int t = 0; t < 4; ++t) Wc[t] = winv[(long long)kc * EB * EB + (g4 + 4 * t) * EB + c];
        };
        auto solve = [&](int k, const double* P, const double* Wc) {
            const int m = mk[k];
            double s = 0.0;
#pragma unroll
            for (int i = 1; i <= DT; ++i) {
                if (i <= m) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) s += P[4 * (i - 1) + t] * xs[EB * (k + i) + g4 + 4 * t];
                }
            }
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            const double z = xs[EB * k + c] - s;
            wave_lds_order();
            if (lane < EB) xs[EB * k + c] = z;
            wave_lds_order();
            double xv = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) xv += Wc[t] * xs[EB * k + g4 + 4 * t];
            xv += __shfl_xor(xv, 16, 64);
            xv += __shfl_xor(xv, 32, 64);
            wave_lds_order();
            if (lane < EB) xs[EB * k + c] = xv;
            wave_lds_order();
        };
        load(nb - 1, pl[0], wl[0]);
        int k = nb - 1;
        for (; k >= 1; k -= 2) {
            load(k - 1, pl[1], wl[1]);
            solve(k, pl[0], wl[0]);
            load(k - 2, pl[0], wl[0]);
            solve(k - 1, pl[1], wl[1]);
        }
        if (k == 0) solve(0, pl[0], wl[0]);
    }
    __syncthreads();
    for (int i = tid; i < Pp; i += 256) x[i] = xs[i];
    if (tid == 0 && sbad) ctl->chol_ok = 0;
}

size_t chol_band_lds(int Pp, int D) {
    const int nb = Pp / EB;
    return (size_t)(env_window_slots(D) + 2) * EBLK * sizeof(double) + (size_t)2 * Pp * sizeof(double) +
           (size_t)2 * nb * sizeof(int);
}

void launch_chol_band(double* H, int Pp, int D, const int* first, const double* g, double* x, double* winv,
                      Control* ctl, hipStream_t s) {
    if (D <= 3)
        hipLaunchKernelGGL(k_chol_env<3>, dim3(1), dim3(256), chol_band_lds(Pp, D), s, H, Pp, D, first, g, x, winv,
                           ctl);
    else
        hipLaunchKernelGGL(k_chol_env<EDMAX>, dim3(1), dim3(256), chol_band_lds(Pp, D), s, H, Pp, D, first, g, x,
                           winv, ctl);
}

int prepare_chol_band(int Pp, int D) {
    const int bytes = (int)chol_band_lds(Pp, D);
    const void* f = D <= 3 ? (const void*)k_chol_env<3> : (const void*)k_chol_env<EDMAX>;
    return (int)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

int chol_band_max_blocks() { return EDMAX; }
int chol_band_block() { return EB; }

// ---------------------------------------------------------------------------
// L L^T x = g, one workgroup, right-hand side in LDS
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_chol_solve(const double* H, const double* Linv, const double* g,
                                                     double* x, int Pp) {
    extern __shared__ double rhs[];
    __shared__ double y[NB];
    const int tid = threadIdx.x;
    const int nblk = Pp / NB;
    for (int i = tid; i < Pp; i += 1024) rhs[i] = g[i];
    __syncthreads();
    for (int k = 0; k < nblk; ++k) {
        const double* Li = Linv + (long long)k * NB * NB;
        if (tid < NB) {
            double s = 0.0;
            for (int m = 0; m <= tid; ++m) s += Li[tid * NB + m] * rhs[k * NB + m];
            y[tid] = s;
        }
        __syncthreads();
        if (tid < NB) rhs[k * NB + tid] = y[tid];
        for (int i = (k + 1) * NB + tid; i < Pp; i += 1024) {
            const double* row = H + (long long)i * Pp + k * NB;
            double s = 0.0;
#pragma unroll 8
            for (int m = 0; m < NB; ++m) s += row[m] * y[m];
            rhs[i] -= s;
        }
        __syncthreads();
    }
    for (int k = nblk - 1; k >= 0; --k) {
        const double* Li = Linv + (long long)k * NB * NB;
        if (tid < NB) {
            double s = 0.0;
            for (int m = tid; m < NB; ++m) s += Li[m * NB + tid] * rhs[k * NB + m];
            y[tid] = s;
        }
        __syncthreads();
        if (tid < NB) rhs[k * NB + tid] = y[tid];
        for (int j = tid; j < k * NB; j += 1024) {
            double s = 0.0;
#pragma unroll 8
            for (int m = 0; m < NB; ++m) s += H[(long long)(k * NB + m) * Pp + j] * y[m];
            rhs[j] -= s;
        }
        __syncthreads();
    }
    for (int i = tid; i < Pp; i += 1024) x[i] = rhs[i];
}

void launch_chol_solve(const double* H, const double* Linv, const double* g, double* delta, int Pp,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_chol_solve, dim3(1), dim3(1024), Pp * sizeof(double), s, H, Linv, g, delta, Pp);
}

}  // namespace spgo
