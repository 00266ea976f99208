// Reduced pose system, its Cholesky factorisation and the triangular solves
// (the work of Solver_cuda.cu:690-736 and linear_solvers.cu:195-218).
//
// k_form_system sums, for every 6x6 pose block (i, j), the anchor-frame blocks
// D of the landmark groups whose pose span covers it and maps them to the
// incremental-pose parameters with the adjoints of G_a^-1 G_i:
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
template <typename T>
__global__ __launch_bounds__(64) void k_form_system(Layout lay, StateBuf<T> B0, StateBuf<T> B1,
                                                    Control* ctl, double* H, double* gvec, int Pp) {
    const int p = blockIdx.x;
    const int lane = threadIdx.x;
    if (p == 0 && lane == 0) ctl->chol_ok = 1;   // reset for this iteration's factorisation
    // lower-triangle block pair (bi >= bj)
    int bi = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while ((bi + 1) * (bi + 2) / 2 <= p) ++bi;
    while (bi * (bi + 1) / 2 > p) --bi;
    const int bj = p - bi * (bi + 1) / 2;
    const StateBuf<T>& bs = ctl->cur ? B1 : B0;

    __shared__ double Ht[36], AdI[36], AdJ[36], M1[36], gt[6];
    const int r = lane / 6, c = lane % 6;
    double h = 0.0, gacc = 0.0;
    const int amin = max(0, bi - lay.kmax + 1);
    for (int a = amin; a <= bj; ++a) {
        const int gb = lay.anchor_grp[a], ge = lay.anchor_grp[a + 1];
        bool any = false;
        double hs = 0.0, gs = 0.0;
        for (int g = gb; g < ge; ++g) {
            const int K = lay.grp_K[g];
            if (bi - a >= K) continue;
            any = true;
            const int K6 = 6 * K;
            if (lane < 36)
                hs += (double)bs.D[lay.grp_D_off[g] + (long long)(6 * (bi - a) + r) * K6 + 6 * (bj - a) + c];
            if (bi == bj && lane < 6) gs += (double)bs.d[lay.grp_d_off[g] + 6 * (bi - a) + lane];
        }
        if (!any) continue;   // uniform across the wavefront
        if (lane < 36) {
            double Gai[12], Gi[12], E[12];
            pose_load(bs.GinvL + 12 * a, Gai);
            pose_load(bs.G + 12 * bi, Gi);
            pose_mul(Gai, Gi, E);
            AdI[lane] = adjoint_entry(E, r, c);
            pose_load(bs.G + 12 * bj, Gi);
            pose_mul(Gai, Gi, E);
            AdJ[lane] = adjoint_entry(E, r, c);
            Ht[lane] = hs;
        }
        if (bi == bj && lane < 6) gt[lane] = gs;
        __syncthreads();
        if (lane < 36) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += Ht[6 * r + k] * AdJ[6 * k + c];
            M1[lane] = s;
        }
        __syncthreads();
        if (lane < 36) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += AdI[6 * k + r] * M1[6 * k + c];
            h += s;
        }
        if (bi == bj && lane < 6) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += AdI[6 * k + lane] * gt[k];
            gacc += s;
        }
        __syncthreads();
    }
    if (lane < 36) {
        if (bi == bj && r == c) h += ctl->lambda;
        H[(long long)(6 * bi + r) * Pp + 6 * bj + c] = h;
    }
    if (bi == bj && lane < 6) gvec[6 * bi + lane] = gacc;
}

template <typename T>
void launch_form_system(const Layout& lay, const StateBuf<T>* bufs, Control* ctl, double* H, double* g,
                        int Pp, hipStream_t s) {
    const int nb = lay.N;
    const int npairs = nb * (nb + 1) / 2;
    hipLaunchKernelGGL(k_form_system<T>, dim3(npairs), dim3(64), 0, s, lay, bufs[0], bufs[1], ctl, H, g, Pp);
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
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double piv = readlane_d(a[j], j);
            if (!(piv > 0.0) || !isfinite(piv)) {
                bad = true;
                piv = 1.0;
            }
            const double ljj = sqrt(piv);
            const double inv = 1.0 / ljj;
            a[j] = (i == j) ? ljj : a[j] * inv;
#pragma unroll
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
#pragma unroll
            for (int rr = 0; rr < NB; ++rr) {
                double s = (rr == cidx) ? 1.0 : 0.0;
#pragma unroll
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
