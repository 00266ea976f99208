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
// H is block-banded: block row i (BB x BB blocks) has nonzero blocks only in
// columns first[i] .. i (first[] nondecreasing), and the Cholesky factor keeps
// that envelope.  Right-looking over block columns k: the tall panel (the
// diagonal block and the m <= BMAXB blocks below it with first <= k) sits in
// registers, one wavefront per block and one row per lane; every column j is
// one LDS broadcast of the diagonal block's column and one barrier.  The
// right-hand side rides along as an extra column, so the forward solve is done
// with the factorisation.  The trailing update of the blocks below is on MFMA
// and the backward solve runs block by block from the stored factor.
// ---------------------------------------------------------------------------
constexpr int BB = 32;      // block size of the envelope factorisation
constexpr int BMAXB = 3;    // blocks below the diagonal (tall panel of at most 128 rows)
constexpr int PR = BB * (BMAXB + 1);
constexpr int FST = BB + 1; // LDS row stride of the staged panel

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
// global stores (s_waitcnt lgkmcnt(0), vmcnt/expcnt left at their maximum)
__device__ __forceinline__ void lds_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__global__ __launch_bounds__(256) void k_chol_band(double* H, int Pp, const int* first, const double* g, double* x,
                                                   double* dinv, Control* ctl) {
    // S: the tall panel of block column k (diagonal block and m <= BMAXB
    // blocks below), staged through LDS so global traffic is row-contiguous.
    // Wavefront 0 factors it, lane l holding panel rows l and l + 64, with no
    // barrier inside the column loop.  The right-hand side / solution xs and
    // the inverse diagonal dv stay in LDS for the whole factorisation.
    extern __shared__ double xs[];          // Pp
    double* dv = xs + Pp;                   // Pp
    __shared__ double S[PR * FST];
    __shared__ double part[BMAXB][BB];
    __shared__ int sbad;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nb = Pp / BB;
    for (int i = tid; i < Pp; i += 256) xs[i] = g[i];
    if (tid == 0) sbad = 0;
    __syncthreads();

    constexpr int PER = PR * BB / 256;      // staged elements per thread (16)
    for (int k = 0; k < nb; ++k) {
        int m = 0;
        while (m < BMAXB && k + 1 + m < nb && first[k + 1 + m] <= k) ++m;
        const int R = BB * (m + 1);
        const long long r0 = (long long)BB * k;
        {
            // stage the panel: all loads in flight, then into LDS
            double v[PER];
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int e = tid + 256 * i, r = e / BB, c = e % BB;
                v[i] = (r < R) ? H[(r0 + r) * Pp + r0 + c] : 0.0;
            }
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int e = tid + 256 * i, r = e / BB, c = e % BB;
                S[r * FST + c] = v[i];
            }
        }
        lds_barrier();
        if (w == 0) {
            // wavefront 0 factors the staged panel in place: lane l owns rows
            // l and l + 64 (zero rows beyond R are harmless).  Column j is
            // broadcast with v_readlane; the row updates are LDS
            // read-modify-writes, so no register arrays are live.
            double* rowA = S + lane * FST;
            double* rowB = S + (lane + 64) * FST;
            double bA = (lane < R) ? xs[r0 + lane] : 0.0;
            double bB = (lane + 64 < R) ? xs[r0 + 64 + lane] : 0.0;
            bool bad = false;
            double myinv = 0.0;
#pragma unroll 1
            for (int j = 0; j < BB; ++j) {
                const double u = rowA[j];     // unscaled column-j entry (diagonal rows are lanes 0..31)
                const double u2 = rowB[j];
                double piv = readlane_d(u, j);
                if (!(piv > 0.0) || !isfinite(piv)) {
                    bad = true;
                    piv = 1.0;
                }
                const double inv = rsqrt_fast(piv);
                const double ljj = piv * inv;
                const double yj = readlane_d(bA, j) * inv;
                const double lA = (lane == j) ? ljj : u * inv;
                const double lB = u2 * inv;
                rowA[j] = lA;
                rowB[j] = lB;
                if (lane == j) {
                    myinv = inv;
                    bA = yj;
                } else if (lane > j) {
                    bA -= lA * yj;
                }
                bB -= lB * yj;
                const double sA = lA * inv, sB = lB * inv;
#pragma unroll 4
                for (int q = j + 1; q < BB; ++q) {
                    const double cq = readlane_d(u, q);
                    rowA[q] -= sA * cq;
                    rowB[q] -= sB * cq;
                }
            }
            if (bad && lane == 0) sbad = 1;
            if (lane < BB) {
                // zero the strict upper triangle of the diagonal block
                for (int q = lane + 1; q < BB; ++q) rowA[q] = 0.0;
                dv[r0 + lane] = myinv;
            }
            if (lane < R) xs[r0 + lane] = bA;
            if (lane + 64 < R) xs[r0 + 64 + lane] = bB;
        }
        lds_barrier();
        // trailing update of the blocks (k+i, k+j), 1 <= j <= i <= m: four
        // 16x16 MFMA tiles each; the tile is prefetched before the products
        const int njobs = (m * (m + 1) / 2) * 4;
        for (int job = w; job < njobs; job += 4) {
            const int pr = job >> 2, sub = job & 3;
            int bi = 0;
            while ((bi + 1) * (bi + 2) / 2 <= pr) ++bi;
            const int bj = pr - bi * (bi + 1) / 2;     // 0-based panel indices, bj <= bi
            const int sr = (sub >> 1) * 16, sc = (sub & 1) * 16;
            double* Ht[4];
            double hv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long r = (long long)BB * (k + 1 + bi) + sr + mfma16_row(0.0, lane, v);
                const long long c = (long long)BB * (k + 1 + bj) + sc + (lane & 15);
                Ht[v] = H + r * Pp + c;
                hv[v] = *Ht[v];
            }
            const double* Fi = S + ((bi + 1) * BB + sr + (lane & 15)) * FST + (lane >> 4);
            const double* Fj = S + ((bj + 1) * BB + sc + (lane & 15)) * FST + (lane >> 4);
            d4 acc = 0.0;
#pragma unroll
            for (int kk = 0; kk < BB; kk += 4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Fi[kk], Fj[kk], acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) *Ht[v] = hv[v] - acc[v];
        }
        // the factor of block column k back to global memory
        {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int e = tid + 256 * i, r = e / BB, c = e % BB;
                if (r < R) H[(r0 + r) * Pp + r0 + c] = S[r * FST + c];
            }
        }
        __syncthreads();   // global: the updated blocks are staged by the next step
    }

    // backward solve L^T x = y from the bottom block row up: wavefronts 1..m
    // form the panel contributions, wavefront 0 the diagonal solve
    for (int k = nb - 1; k >= 0; --k) {
        int m = 0;
        while (m < BMAXB && k + 1 + m < nb && first[k + 1 + m] <= k) ++m;
        const int c0 = BB * k;
        if (w >= 1 && w <= m && lane < BB) {
            const int rb = BB * (k + w);
            double hv[BB];
#pragma unroll
            for (int r = 0; r < BB; ++r) hv[r] = H[(long long)(rb + r) * Pp + c0 + lane];
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int r = 0; r < BB; r += 2) {
                s0 += hv[r] * xs[rb + r];
                s1 += hv[r + 1] * xs[rb + r + 1];
            }
            part[w - 1][lane] = s0 + s1;
        }
        double cl[BB];   // column `lane` of the diagonal factor, loaded while the panel sums run
        if (w == 0 && lane < BB) {
#pragma unroll
            for (int r = 0; r < BB; ++r) cl[r] = H[(long long)(c0 + r) * Pp + c0 + lane];
        }
        lds_barrier();
        if (w == 0 && lane < BB) {
            double s = xs[c0 + lane];
            for (int i = 0; i < m; ++i) s -= part[i][lane];
            const double myinv = dv[c0 + lane];
            int rl = lane;
            asm volatile("" : "+v"(rl));
#pragma clang loop unroll(full)
            for (int r = BB - 1; r >= 0; --r) {
                const double xr = readlane_d(s, r) * readlane_d(myinv, r);
                if (rl == r) s = xr;
                else if (rl < r) s -= cl[r] * xr;
            }
            xs[c0 + lane] = s;
        }
        lds_barrier();
    }
    for (int i = tid; i < Pp; i += 256) {
        x[i] = xs[i];
        dinv[i] = dv[i];
    }
    if (tid == 0 && sbad) ctl->chol_ok = 0;
}

size_t chol_band_lds(int);

void launch_chol_band(double* H, int Pp, const int* first, const double* g, double* x, double* dinv, Control* ctl,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_chol_band, dim3(1), dim3(256), chol_band_lds(Pp), s, H, Pp, first, g, x, dinv, ctl);
}

int prepare_chol_band(int Pp) {
    return (int)hipFuncSetAttribute((const void*)k_chol_band, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)chol_band_lds(Pp));
}

int chol_band_max_blocks() { return BMAXB; }
int chol_band_block() { return BB; }
size_t chol_band_lds(int Pp) { return (size_t)2 * Pp * sizeof(double); }

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
