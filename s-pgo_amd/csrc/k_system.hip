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
}  // namespace

// ---------------------------------------------------------------------------
// reduced system
// ---------------------------------------------------------------------------
constexpr int FS_MAXG = 512;   // groups whose span can reach one pose row

// R rows per workgroup (1, or the 6 rows of one pose, which share the covering
// list): one row per workgroup fills the device for small systems, one pose per
// workgroup stages each list once when there are many rows
template <typename T, int R>
__global__ __launch_bounds__(256) void k_form_system(Layout lay, StateBuf<T> B0, StateBuf<T> B1, Control* ctls,
                                                     double* H, double* gvec) {
    // the covering groups of this pose block (precomputed per layout, anchors
    // ascending); row r of a group's block starts at base + r K6
    __shared__ long long base[FS_MAXG];
    __shared__ int canc[FS_MAXG], k6s[FS_MAXG], gdo[FS_MAXG];
    const int row_g0 = blockIdx.x * R;            // first row, of all problems' poses
    const int tid = threadIdx.x;
    const int pg = row_g0 / 6;
    const ProbDesc pd = lay.probs[lay.pose_prob[pg]];
    Control* ctl = ctls + lay.pose_prob[pg];
    const int row0 = row_g0 - 6 * pd.pose_base;  // first row inside its problem
    const int cbase = 6 * pd.pose_base;
    if (row0 == 0 && tid == 0) {
        // reset for this iteration's factorisation: without a converged CG solve
        // the Cholesky computes the step
        ctl->chol_ok = 1;
        ctl->pcg_fail = 1;
        ctl->pcg_iters = 0;
    }
    const StateBuf<T> bs = ctl->cur ? B1 : B0;
    const int e0 = lay.fs_begin[pg];
    const int nc = min(lay.fs_begin[pg + 1] - e0, FS_MAXG);
    for (int e = tid; e < nc; e += 256) {
        const FsEntry f = lay.fs_list[e0 + e];
        canc[e] = f.anchor;
        k6s[e] = f.K6;
        base[e] = f.D_off - (long long)(6 * f.anchor) * f.K6 - 6 * f.anchor;
        gdo[e] = f.d_off - 6 * f.anchor;
    }
    __syncthreads();
    // items (r, col), c0 <= col <= row0 + r.  Left of c0, six times the first
    // covering group's anchor (the smallest), the rows are zero in H and stay
    // zero: H is filled with zeros when allocated, and no factorisation writes
    // outside the profile (the dense one keeps its exact zeros there).
    const int c0 = nc > 0 ? min(6 * canc[0] - cbase, row0) : row0;
    const int w0 = row0 - c0 + 1;   // items of the first row
    const int items = R * w0 + R * (R - 1) / 2;
    for (int t = tid; t < items; t += 256) {
        int r = 0;
        while (r + 1 < R && t >= (r + 1) * w0 + (r + 1) * r / 2) ++r;
        const int col = c0 + t - (r * w0 + r * (r - 1) / 2);
        const int row = row0 + r, row_g = row_g0 + r;
        const int colg = col + cbase, bj = colg / 6;
        // covering groups with anchor <= bj form a prefix (anchors ascending):
        // their blocks reach column col
        int lo = 0, hi = nc;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (canc[mid] <= bj) lo = mid + 1;
            else hi = mid;
        }
        const int ne = lo;
        double h0 = 0.0, h1 = 0.0;
        // loads in batches of 16 before their adds, so a long list is a few round
        // trips; even entries into h0, odd into h1.  Every load is issued (past the
        // list it repeats the last entry, weighted by zero): a conditional load
        // waits for its data before the next one starts.
        for (int eb = 0; eb < ne; eb += 16) {
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int e = min(eb + i, ne - 1);
                const T x = bs.D[base[e] + (long long)row_g * k6s[e] + colg];
                v[i] = (eb + i < ne) ? (double)x : 0.0;
            }
#pragma unroll
            for (int i = 0; i < 16; i += 2) {
                h0 += v[i];
                h1 += v[i + 1];
            }
        }
        double h = h0 + h1;
        if (col == row) {
            h += ctl->lambda;
            double gs = 0.0;
            for (int q = 0; q < nc; ++q) gs += (double)bs.d[gdo[q] + row_g];
            gvec[pd.vec_off + row] = gs;
        }
        H[pd.H_off + (long long)row * pd.Pp + col] = h;
    }
}

// Entry (e, j) of one block of the frame-coordinate system: the sum over the
// block's plan entries (ZCon), a block of a group's B (K6 > 0; a diagonal
// target block is stored lower, doff >= 0, so its upper entries are read
// transposed) or an anchor block R (K6 = -2: +R, K6 = -1: -R).  The plan
// entries and values are loaded in batches of 8 before their adds, so a list
// of hundreds of entries (a frame seen by many groups) costs a few dozen round
// trips rather than two per entry; past the list the last entry is loaded
// again and weighted by zero.  Each value enters as sign * value with the sign
// (+1, -1 or 0) chosen by selects: a value kept or dropped by a condition
// would let the compiler move its load under a branch and wait for each in
// turn.
__device__ __forceinline__ double frame_block_sum(const ZCon* __restrict__ con, int kb, int ke,
                                                  const double* __restrict__ Dw, const double* __restrict__ Rz,
                                                  int e, int j) {
    double h0 = 0.0, h1 = 0.0;
    for (int qb = kb; qb < ke; qb += 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            // the whole entry in one 16-byte load, the address by selects
            const int4 z = reinterpret_cast<const int4*>(con)[min(qb + i, ke - 1)];
            const long long off = (long long)(((unsigned long long)(unsigned)z.y << 32) | (unsigned)z.x);
            const bool dblk = z.z > 0;
            const bool tr = dblk && z.w >= 0 && j > e;
            const int ld = dblk ? z.z : 6;
            const long long idx = off + (long long)(tr ? j : e) * ld + (tr ? e : j);
            const uintptr_t base = dblk ? (uintptr_t)Dw : (uintptr_t)Rz;
            const double x = *reinterpret_cast<const __attribute__((address_space(1))) double*>(
                base + 8 * (uintptr_t)idx);
            const double w = qb + i < ke ? (z.z == -1 ? -1.0 : 1.0) : 0.0;
            v[i] = w * x;
        }
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            h0 += v[i];
            h1 += v[i + 1];
        }
    }
    return h0 + h1;
}

// Element x of a frame's gradient: the gradient parts (doff) of its diagonal
// block's plan entries, + for the anchor's own (K6 = -2), - for a target's,
// batched as frame_block_sum.
__device__ __forceinline__ double frame_grad_sum(const ZCon* __restrict__ con, int kb, int ke,
                                                 const double* __restrict__ Rz, int x) {
    double g0 = 0.0, g1 = 0.0;
    for (int qb = kb; qb < ke; qb += 8) {
        double v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int4 z = reinterpret_cast<const int4*>(con)[min(qb + i, ke - 1)];
            const double y = Rz[z.w + x];
            const double w = qb + i < ke ? (z.z == -2 ? 1.0 : -1.0) : 0.0;
            v[i] = w * y;
        }
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            g0 += v[i];
            g1 += v[i + 1];
        }
    }
    return g0 + g1;
}

// The reduced system in frame coordinates (one problem): workgroup r writes the
// six rows of frame r, every element left of the diagonal (zeros outside the
// pattern), as the sum of the group contributions the layout's plan lists for
// the block (r, c) (ZCon), plus the damping lambda T, T = D^T diag(F_i) D with
// (D z)_i = z_i - z_{i+1} and F_i = Ad(G_i^-1)^T Ad(G_i^-1): lambda ||delta||^2
// in these variables (T_rr = F_r + F_{r-1}, T_{r,r-1} = -F_{r-1}).
template <typename T>
__global__ __launch_bounds__(256) void k_form_frames(Layout lay, StateBuf<T> B0, StateBuf<T> B1, Control* ctl,
                                                     double* H, double* gvec) {
    constexpr int MAXN = 4096;   // frames up to which the row's block map is kept in LDS
    __shared__ int blk[MAXN];    // blk[c]: the plan index of block (r, c) in this row, -1 if none
    __shared__ double F[2][36];
    const int r = blockIdx.x;
    const int tid = threadIdx.x;
    const int Pp = lay.probs[0].Pp;
    if (r == 0 && tid == 0) {
        ctl->chol_ok = 1;
        ctl->pcg_fail = 1;
        ctl->pcg_iters = 0;
    }
    const StateBuf<T> bs = ctl->cur ? B1 : B0;
    const double lam = ctl->lambda;
    const int e0 = lay.zp_row[r], ne = lay.zp_row[r + 1] - e0;
    const bool mapped = r < MAXN;   // the same for the whole workgroup
    if (mapped) {
        for (int i = tid; i <= r; i += 256) blk[i] = -1;
        __syncthreads();
        for (int i = tid; i < ne; i += 256) blk[lay.zp_col[e0 + i]] = i;
    }
    if (tid < 72) {
        const int w = tid / 36, k = tid - 36 * w, f = r - w;
        double v = 0.0;
        if (f >= 0) {
            double Gi[12];
            pose_load(bs.GinvL + 12 * f, Gi);
            const int p = k / 6, q = k - 6 * p;
#pragma unroll
            for (int x = 0; x < 6; ++x) v += adjoint_entry(Gi, x, p) * adjoint_entry(Gi, x, q);
        }
        F[w][k] = v;
    }
    __syncthreads();
    const int items = 36 * r + 21;   // row 6r + e holds columns 0 .. 6r + e
    for (int t = tid; t < items; t += 256) {
        int e = 0;
        while (e < 5 && t >= (e + 1) * (6 * r + 1) + (e + 1) * e / 2) ++e;
        const int col = t - (e * (6 * r + 1) + e * (e - 1) / 2);
        const int c = col / 6, j = col - 6 * c;
        int lo;
        if (mapped) {
            lo = blk[c];
        } else {
            int hi = ne;
            lo = 0;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (lay.zp_col[e0 + mid] < c) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= ne || lay.zp_col[e0 + lo] != c) lo = -1;
        }
        double h = 0.0;
        if (lo >= 0) {
            const int k = e0 + lo;
            h = frame_block_sum(lay.zp_con, lay.zp_cbeg[k], lay.zp_cbeg[k + 1], bs.Dw, bs.Rz, e, j);
        }
        if (lam != 0.0) {
            if (c == r) h += lam * (F[0][6 * e + j] + F[1][6 * e + j]);
            else if (c == r - 1) h -= lam * F[1][6 * e + j];
        }
        H[(long long)(6 * r + e) * Pp + col] = h;
    }
    if (tid < 6) {
        // the gradient: the diagonal block's contributions (the row's last entry)
        double gs = 0.0;
        if (ne > 0 && lay.zp_col[e0 + ne - 1] == r) {
            const int k = e0 + ne - 1;
            gs = frame_grad_sum(lay.zp_con, lay.zp_cbeg[k], lay.zp_cbeg[k + 1], bs.Rz, tid);
        }
        gvec[6 * r + tid] = gs;
    }
}

template <typename T>
void launch_form_system(const Layout& lay, const StateBuf<T>* bufs, Control* ctls, double* H, double* g,
                        hipStream_t s) {
    if (lay.frame_mode) {
        hipLaunchKernelGGL((k_form_frames<T>), dim3(lay.N), dim3(256), 0, s, lay, bufs[0], bufs[1], ctls, H, g);
        return;
    }
    if (lay.form_rows == 6)
        hipLaunchKernelGGL((k_form_system<T, 6>), dim3(lay.N), dim3(256), 0, s, lay, bufs[0], bufs[1], ctls, H, g);
    else
        hipLaunchKernelGGL((k_form_system<T, 1>), dim3(6 * lay.N), dim3(256), 0, s, lay, bufs[0], bufs[1], ctls, H,
                           g);
}
template void launch_form_system<double>(const Layout&, const StateBuf<double>*, Control*, double*, double*,
                                         hipStream_t);
template void launch_form_system<float>(const Layout&, const StateBuf<float>*, Control*, double*, double*,
                                        hipStream_t);

// ---------------------------------------------------------------------------
// blocked Cholesky, block column k
// ---------------------------------------------------------------------------
// Blocks of step k: block 0 factors the diagonal block (and inverts it), block
// b > 0 the b-th active block row act[act_off[k] + b - 1] (the block rows below
// k whose profile reaches column block k; the others are zero there).  Every
// workgroup factors the diagonal block itself (one wavefront), so the panel
// blocks need no second launch; its other wavefronts load the panel block
// meanwhile.
//   factor: one row per lane, right-looking over the 32 columns.  Column j's
//     entry of row j + 1 reaches the lanes by a lane read; the rest of the
//     column goes through LDS, read back as broadcasts and applied at the start
//     of column j + 1.  Pivot j + 1 is t - (u_(j+1) / L_jj)^2, with t formed from
//     the diagonal lane's own registers, and 1/L_jj comes from the refined rsqrt.
//   inverse W = L^-1 by 16x16 blocks: both diagonal blocks by forward
//     substitution at once (lanes 0-15 and 16-31, one column of W each), then
//     W21 = -W22 (L21 W11) as two MFMA products.
//   panel: L_bk = A_bk W^T, one 16x16 MFMA tile per wavefront.
__global__ __launch_bounds__(256) void k_chol_panel(double* H, double* Linv, int Pp, int k, const int* act_off,
                                                    const int* act, Control* ctl) {
    __shared__ double Lkk[NB][NB + 1];   // the factor, row-major, zeros above the diagonal
    __shared__ double Li[NB][NB + 1];    // its inverse, row-major
    __shared__ double Ab[NB][NB + 1];    // the panel block of this workgroup
    __shared__ double colb[NB * NB];     // column broadcasts of the factor loop
    __shared__ double Tm[16][17];        // L21 W11
    __shared__ double dinv[NB];          // 1 / L_jj
    if (ctl->pcg_fail == 0) return;      // the CG solve of this iteration converged
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int k0 = k * NB;
    const int b0 = blockIdx.x > 0 ? act[act_off[k] + blockIdx.x - 1] * NB : 0;
    if (w == 0) {
        const int r = lane & (NB - 1);   // lanes 32-63 repeat lanes 0-31
        double a[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = H[(long long)(k0 + r) * Pp + k0 + j];
        double up[NB];
        double scp = 0.0, uprev = 0.0, iv = 0.0;
        double piv = readlane_d(a[0], 0);
        bool bad = false;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double t = 0.0;
            if (j + 1 < NB) t = readlane_d(j >= 1 ? fma(-scp, uprev, a[j + 1]) : a[j + 1], j + 1);
            if (j >= 1) {
#pragma unroll
                for (int q = j + 1; q < NB; ++q) {
                    a[q] = fma(-scp, up[q], a[q]);
                    asm volatile("" : "+v"(a[q]));
                }
            }
            // u = A(:, j) after the updates of columns < j; for lane j, u * inv is
            // L_jj and sc is 1, so the update only touches entries above the
            // diagonal of rows <= j, which are not kept
            const double u = a[j];
            if (j + 2 < NB) {
                if (lane < NB) colb[NB * j + lane] = u;
                wave_lds_order();
#pragma unroll
                for (int q = j + 2; q < NB; ++q) up[q] = colb[NB * j + q];
                wave_lds_order();
            }
            double uq = 0.0;
            if (j + 1 < NB) uq = readlane_d(u, j + 1);
            // a pivot outside (0, inf) fails the factorisation; 1 stands in for it
            const bool ok = piv > 0.0 && piv < __builtin_inf();
            bad = bad || !ok;
            const double inv = rsqrt_fast(ok ? piv : 1.0);
            const double l = u * inv;
            const double sc = l * inv;
            if (r == j) iv = inv;
            a[j] = l;
            if (j + 1 < NB) {
                a[j + 1] = fma(-sc, uq, a[j + 1]);
                const double x = uq * inv;
                piv = fma(-x, x, t);
            }
            scp = sc;
            uprev = u;
        }
        if (lane < NB) {
#pragma unroll
            for (int j = 0; j < NB; ++j) Lkk[r][j] = (j <= r) ? a[j] : 0.0;
            dinv[r] = iv;
        }
        if (bad && blockIdx.x == 0 && lane == 0) ctl->chol_ok = 0;
    } else if (blockIdx.x > 0) {
        for (int e = tid - 64; e < NB * NB; e += 192) {
            const int rr = e / NB, cc = e % NB;
            Ab[rr][cc] = H[(long long)(b0 + rr) * Pp + k0 + cc];
        }
    }
    __syncthreads();
    if (w == 0) {
        if (lane < NB) {
            // column c of the inverse of diagonal block h (rows and columns 16h ..)
            const int h = lane >> 4, c = lane & 15, o = 16 * h;
            double x[16];
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                double v = (rr == c) ? 1.0 : 0.0;
#pragma unroll
                for (int m = 0; m < rr; ++m) v = fma(-Lkk[o + rr][o + m], x[m], v);
                x[rr] = v * dinv[o + rr];
            }
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                Li[o + rr][o + c] = x[rr];
                Li[o + rr][16 - o + c] = 0.0;   // W12 = 0; W21 is overwritten below
            }
        }
    }
    __syncthreads();
    if (w == 0) {
        // T = L21 W11: A(m, kk) = L(16 + m, kk), B(kk, n) = W(kk, n)
        d4 acc = 0.0;
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Lkk[16 + (lane & 15)][4 * s4 + (lane >> 4)],
                                                       Li[4 * s4 + (lane >> 4)][lane & 15], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Tm[mfma16_row(0.0, lane, v)][lane & 15] = acc[v];
    }
    __syncthreads();
    if (w == 0) {
        // W21 = -W22 T
        d4 acc = 0.0;
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Li[16 + (lane & 15)][16 + 4 * s4 + (lane >> 4)],
                                                       Tm[4 * s4 + (lane >> 4)][lane & 15], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Li[16 + mfma16_row(0.0, lane, v)][lane & 15] = -acc[v];
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
    // L_bk = A_bk W^T: wavefront w forms the 16x16 tile (w / 2, w % 2);
    // A(m, kk) = A_bk(16 rt + m, kk), B(kk, n) = W(16 ct + n, kk)
    {
        const int rt = w >> 1, ct = w & 1;
        d4 acc = 0.0;
#pragma unroll
        for (int s8 = 0; s8 < 8; ++s8)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Ab[16 * rt + (lane & 15)][4 * s8 + (lane >> 4)],
                                                       Li[16 * ct + (lane & 15)][4 * s8 + (lane >> 4)], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v)
            H[(long long)(b0 + 16 * rt + mfma16_row(0.0, lane, v)) * Pp + k0 + 16 * ct + (lane & 15)] = acc[v];
    }
}

// Trailing update A_ij -= L_ik L_jk^T for the active block rows j <= i of step
// k, one NB x NB tile per workgroup, four 16x16 MFMA tiles per
// wavefront-quarter of the block.
__global__ __launch_bounds__(256) void k_chol_update(double* H, int Pp, int k, const int* act_off, const int* act,
                                                     const Control* ctl) {
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int a0 = act_off[k], n = act_off[k + 1] - a0;
    const int t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while (ti * (ti + 1) / 2 > t) --ti;
    const int tj = t - ti * (ti + 1) / 2;
    if (ti >= n) return;
    const int bi = act[a0 + ti], bj = act[a0 + tj];
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

void launch_cholesky(double* H, double* Linv, int Pp, const int* act_off, const int* act, const int* nact,
                     Control* ctl, hipStream_t s) {
    const int nblk = Pp / NB;
    for (int k = 0; k < nblk; ++k) {
        const int n = nact[k];
        hipLaunchKernelGGL(k_chol_panel, dim3(1 + n), dim3(256), 0, s, H, Linv, Pp, k, act_off, act, ctl);
        if (n > 0) hipLaunchKernelGGL(k_chol_update, dim3(n * (n + 1) / 2), dim3(256), 0, s, H, Pp, k, act_off, act, ctl);
    }
}

void dense_profile(int Pp, const std::vector<int>& fb16, std::vector<int>& first, std::vector<int>& act_off,
                   std::vector<int>& act, std::vector<int>& nact) {
    const int nblk = Pp / NB, r = NB / 16;
    first.assign(nblk, 0);
    for (int i = 0; i < nblk; ++i) {
        int f = i;
        for (int q = 0; q < r && r * i + q < (int)fb16.size(); ++q) f = std::min(f, fb16[r * i + q] / r);
        first[i] = f;
    }
    act_off.assign(nblk + 1, 0);
    act.clear();
    nact.assign(nblk, 0);
    for (int k = 0; k < nblk; ++k) {
        act_off[k] = (int)act.size();
        for (int i = k + 1; i < nblk; ++i)
            if (first[i] <= k) act.push_back(i);
        nact[k] = (int)act.size() - act_off[k];
    }
    act_off[nblk] = (int)act.size();
}

// ---------------------------------------------------------------------------
// Envelope (profile) Cholesky and both triangular solves.
//
// H is block-banded with EB x EB blocks: block row i has nonzero blocks only
// in columns first[i] .. i (first[] nondecreasing, i - first[i] <= D), and
// the Cholesky factor keeps that envelope.  One workgroup factors a block
// system right-looking over block columns k; the blocks still to be updated
// live in a sliding LDS window of W = D + 3 block rows, block (i, j) in the
// slot of the unordered pair (i mod W, j mod W), which is unique among the
// resident rows.  Two barriers per column:
//   phase A  wavefront 0 (and 1, 2 for profiles deeper than 3 blocks) factors
//            the panel of column k from registers, one row per lane: lanes
//            0-15 the diagonal block, lanes 16-63 three blocks below it (a
//            second or third wavefront repeats the diagonal block for its own
//            three blocks, so no wavefront waits on another inside the column
//            loop).  The right-hand side rides along as one more column, so
//            the forward solve is done with the factorisation.  A column's
//            entries reach the other lanes by a lane read (the one on the
//            pivot chain) and through LDS broadcasts (the rest).  The panel
//            is written to LDS only.  Meanwhile the other wavefronts apply
//            the part of the previous column's trailing update that the next
//            panel does not need, and wavefront 3 moves the block row it
//            prefetched into the window, inverts the previous diagonal
//            factor (W_k) and stores W_k and Q_(k,i) = L_(k+i,k) W_k, formed
//            on MFMA, for the backward solve, and prefetches the next block
//            row.
//   phase B  the trailing update of block column k + 1 (the next panel) on
//            MFMA, one 16x16 tile per wavefront.
// The backward solve x_k = W_k^T y_k - sum_i Q_(k,i)^T x_(k+i) runs block by
// block from the bottom, one wavefront, with only the x_(k+1) term on its
// sequential chain.  The kernels are compiled for profile depths 3, 4 and
// EDMAX (the depth sizes the prefetch registers and unrolled loops).
//
// Long profiles are split in two (a twisted factorisation): a separator S of
// D block rows in the middle leaves a top part T and a bottom part B that do
// not couple.  One kernel factors T top-down and B bottom-up at the same time
// in two workgroups (B as the reversed, transposed system J H J, read from
// H's lower triangle); each leaves its Schur update of S behind.  A second
// kernel sums the two updates, factors S, and runs the backward solves of T
// and B concurrently.  The sequential chain drops from nb to about nb/2 + D
// block columns.
// ---------------------------------------------------------------------------
constexpr int EB = 16;            // block size: one MFMA tile
typedef __attribute__((address_space(3))) double lds_double;
constexpr int EDMAX = 7;          // deepest profile: blocks below the diagonal
constexpr int EST = EB + 1;       // LDS row stride of a window block
constexpr int EBLK = EB * EST;    // doubles per window block



// workgroup barrier that orders LDS only: it does not wait for outstanding
// global loads or stores (s_waitcnt lgkmcnt(0), vmcnt/expcnt left at their maximum)
__device__ __forceinline__ void lds_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__host__ __device__ constexpr int env_window_slots(int D) { return (D + 3) * (D + 4) / 2; }

// wavefronts that factor one block column: the diagonal block and up to three
// blocks below it each
__host__ __device__ constexpr int env_factor_waves(int m) { return m > 3 ? (m + 2) / 3 : 1; }

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
// EST) in LDS, with its inverse diagonal dinv: every lane forms column
// c = lane & 15 of W = L^-1 by forward substitution, lanes 0..15 write W
// row-major over the factor, and every lane returns rows g + 4t
// (g = lane >> 4, t = 0..3) of its column, wv[t] = W(g + 4t, c).
__device__ __forceinline__ void env_invert(double* Lc, const double* dinv, double* wv, int lane) {
    const int c = lane & 15, g = lane >> 4;
    double s[EB];
#pragma unroll
    for (int r = 0; r < EB; ++r) s[r] = (r == c) ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < EB; ++q) {
        s[q] *= dinv[q];
#pragma unroll
        for (int r = q + 1; r < EB; ++r) s[r] -= Lc[q * EST + r] * s[q];
    }
    // W crosses lanes through LDS: the writes and the reads back are volatile,
    // so the compiler keeps them in order (and after the substitution's reads,
    // which may alias them), and LDS executes one wavefront's operations in
    // order.  Fences here would constrain the whole helper code (~150 more
    // registers).
    volatile lds_double* Lv = (lds_double*)Lc;
    if (lane < EB) {
#pragma unroll
        for (int r = 0; r < EB; ++r) Lv[r * EST + c] = s[r];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) wv[t] = Lv[(g + 4 * t) * EST + c];
}

// Backward-solve blocks in global memory are 16 x 16 and lane-major: lane
// (c, g) = (lane & 15, lane >> 4) holds entries (g + 4t, c), t = 0..3, as four
// consecutive doubles at 4 * lane.  That is the output layout of the f64
// 16x16x4 MFMA and the operand layout of the backward solve.
__device__ __forceinline__ int lm_index(int r, int c) { return 4 * (c + 16 * (r & 3)) + (r >> 2); }

// Where a local block system lives in global memory: local block row i is
// block row i of H (rev = 0), or block row nb - 1 - i with the system
// reversed and transposed (rev = 1, J H J read from H's lower triangle).
struct EnvSrc {
    double* H;    // lower triangle, row-major, row stride Pp
    int Pp, nb, rev;
};

__device__ __forceinline__ long long env_gaddr(const EnvSrc& s, int i, int j, int r, int c) {
    if (!s.rev) return (long long)(EB * i + r) * s.Pp + EB * j + c;
    return (long long)(EB * (s.nb - 1 - j) + (EB - 1 - c)) * s.Pp + EB * (s.nb - 1 - i) + (EB - 1 - r);
}

// One factorisation context in LDS, and the local system it works on.
struct EnvCtx {
    double* win;    // env_window_slots(D) blocks
    double* ring;   // the last two diagonal factors, column-major
    double* colb;   // per factoring wavefront: the diagonal rows' entries of each column (EB x EB)
    double* xs;     // right-hand side, then forward solution (EB * nloc)
    double* dv;     // 1 / L_ii (EB * nloc)
    int* mk;        // blocks below the diagonal per block column (nloc)
    int* fst;       // local profile (nloc), staged by the caller
    int nloc;       // block rows of the local system
    int steps;      // block columns factored (the rest stay in the window)
    int D;          // profile depth of the local system
    int zero_from;  // blocks (i, j) with i, j >= zero_from load as zero
    int wrev, wn;   // column k's panel and inverse diagonal block go to slot k, or wn - 1 - k
    int preloaded;  // the caller has placed every block of the (small) system in the window
};

// Panel store: for slot s (one per factored block column), the EDMAX blocks
// Q_(k,i) = L_(k+i,k) L_kk^-1 below the diagonal in the local system's
// orientation, lane-major 16 x 16.
__device__ __forceinline__ double* env_panel(double* Lp, int slot, int i) {
    return Lp + ((long long)slot * EDMAX + (i - 1)) * EB * EB;
}

__device__ __forceinline__ int env_wslot(const EnvCtx& c, int k) { return c.wrev ? c.wn - 1 - k : k; }

// blocks below the diagonal of every block column: the rows k+1.. with first <= k
__device__ __forceinline__ void env_counts(const int* fst, int nloc, int D, int* mk) {
    for (int k = threadIdx.x; k < nloc; k += 256) {
        int m = 0;
        while (m < D && k + 1 + m < nloc && fst[k + 1 + m] <= k) ++m;
        mk[k] = m;
    }
}

// Factor the first c.steps block columns of the local system and solve
// forward.  Called by all 256 threads after c.fst and c.xs are staged and a
// barrier.  On return (after a barrier) the window holds the blocks of rows
// and columns >= c.steps with every update applied and c.xs the forward
// solution of rows < EB * steps and the updated right-hand side of the rest.
// The panels and inverse diagonal blocks go to global memory; those of the
// last column may still be in progress on wavefront 3 (it also writes the
// ring), so callers synchronise the workgroup before the backward solve.
template <int DT>
__device__ void env_forward(const EnvSrc src, const EnvCtx c, double* winv, double* Lp, int* sbad) {
    constexpr int EPRE = 4 * (DT + 1);      // prefetched doubles per lane of one block row
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int D = c.D, W = D + 3, nloc = c.nloc;
    double* win = c.win;
    double* ring = c.ring;
    double* xs = c.xs;
    double* dv = c.dv;
    int* mk = c.mk;
    const int* fst = c.fst;
    auto wrap = [W](int r) { return r >= W ? r - W : r; };
    auto zero_blk = [&](int i, int j) { return i >= c.zero_from && j >= c.zero_from; };

    env_counts(fst, nloc, D, mk);
    // rows 0 .. D of the window; every thread owns one element of each block
    if (!c.preloaded) {
        const int r = tid >> 4, cc = tid & 15;
        double v[(DT + 1) * (DT + 2) / 2];
        int n = 0;
#pragma unroll
        for (int i = 0; i <= DT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const bool use = i <= D && i < nloc && j >= fst[min(i, nloc - 1)] && !zero_blk(i, j);
                v[n++] = use ? src.H[env_gaddr(src, i, j, r, cc)] : 0.0;
            }
        n = 0;
#pragma unroll
        for (int i = 0; i <= DT; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) {
                const bool use = i <= D && i < nloc && j >= fst[min(i, nloc - 1)];
                if (use) env_slot(win, i, j)[r * EST + cc] = v[n];
                ++n;
            }
    }
    // wavefront 3 keeps the next block row in registers until it has a slot.
    // Every lane loads unconditionally (the block index is clamped to the row)
    // and only the stores to the window are predicated.
    double pre[EPRE];
    auto prefetch = [&](int i) {
        const int ic = min(i, nloc - 1);
        const int f0 = fst[ic];
#pragma unroll
        for (int t = 0; t < EPRE; ++t) {
            const int e = lane + 64 * t, b = min(e >> 8, ic - f0), rr = (e >> 4) & 15, cc = e & 15;
            pre[t] = src.H[env_gaddr(src, ic, f0 + b, rr, cc)];
        }
    };
    if (w == 3 && D + 1 < nloc) prefetch(D + 1);
    __syncthreads();

    // panel kp to global memory for the backward solve (one wavefront): the
    // inverse W of its diagonal factor and Q_i = L_(kp+i,kp) W, formed on the
    // MFMA from the window blocks, both lane-major
    auto copy_out = [&](int kp, int rkp) {
        const int slot = env_wslot(c, kp);
        double wv[4];
        env_invert(ring + (kp & 1) * EBLK, dv + EB * kp, wv, lane);
        double* wd = winv + (long long)slot * EB * EB + 4 * lane;
#pragma unroll
        for (int t = 0; t < 4; ++t) wd[t] = wv[t];
        const int mp = mk[kp];
        for (int i = 1; i <= mp; ++i) {
            // A(m, k) = L(m, g + 4s) from the window, B(k, n) = W(g + 4s, n)
            const double* L = env_slot(win, wrap(rkp + i), rkp) + (lane & 15) * EST + (lane >> 4);
            d4 acc = 0.0;
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L[4 * s4], wv[s4], acc, 0, 0, 0);
            double* d = env_panel(Lp, slot, i) + 4 * lane;
#pragma unroll
            for (int v = 0; v < 4; ++v) d[v] = acc[v];
        }
        // past the profile, up to the depth the backward solve reads: zeros
        for (int i = mp + 1; i <= DT; ++i) {
            double* d = env_panel(Lp, slot, i) + 4 * lane;
#pragma unroll
            for (int v = 0; v < 4; ++v) d[v] = 0.0;
        }
    };
    // rest of column kp's trailing update: tiles (kp+i, kp+j), 2 <= j <= i <= mp,
    // on wavefronts w0 .. w0 + nw - 1
    auto rest_tiles = [&](int kp, int rkp, int w0, int nw) {
        const int mp = mk[kp];
        const int nt = (mp - 1) * mp / 2;
        if (w < w0 || w >= w0 + nw) return;
        for (int t = w - w0; t < nt; t += nw) {
            int i = 2;
            while ((i - 1) * i / 2 <= t) ++i;
            const int j = 2 + t - (i - 2) * (i - 1) / 2;
            env_tile(win, wrap(rkp + i), wrap(rkp + j), rkp, lane);
        }
    };

    // the same tiles over all four wavefronts, tile t on wavefront (nf + t) mod 4:
    // the helpers first, the factoring wavefronts (after their columns) last
    auto rest_share = [&](int kp, int rkp, int nf, int mp) {
        const int nt = (mp - 1) * mp / 2;
        for (int t = (w - nf + 4) & 3; t < nt; t += 4) {
            int i = 2;
            while ((i - 1) * i / 2 <= t) ++i;
            const int j = 2 + t - (i - 2) * (i - 1) / 2;
            env_tile(win, wrap(rkp + i), wrap(rkp + j), rkp, lane);
        }
    };

    int rk = 0;   // k mod W
    for (int k = 0; k < c.steps; ++k, rk = wrap(rk + 1)) {
        const int m = mk[k];
        const int nf = env_factor_waves(m);
        // ---------------- phase A
        if (w < nf) {
            // panel k - 1's depth, read now so that the check after the columns
            // costs the column chain nothing
            const int mprev = k >= 1 ? mk[k - 1] : 0;
            const int bi = (lane < EB) ? 0 : 1 + 3 * w + ((lane - EB) >> 4);
            const int rr = lane & 15;
            const bool valid = bi <= m;
            // lanes past the profile read the diagonal block instead: finite
            // values that are never broadcast or stored
            const int bic = valid ? bi : 0;
            double a[EB];
            {
                const double* s = env_slot(win, wrap(rk + bic), rk) + rr * EST;
#pragma unroll
                for (int cc = 0; cc < EB; ++cc) a[cc] = s[cc];
            }
            double b = xs[EB * (k + bic) + rr];
            // the column buffer as its own LDS base, so its accesses carry
            // immediate offsets
            lds_double* cb = (lds_double*)(c.colb + w * EB * EB);
            asm volatile("" : "+v"(cb));
            double yv = 0.0, iv = 0.0;
            int rl = lane;
            asm volatile("" : "+v"(rl));
            // Column j's entries u_q of the diagonal rows q > j reach every
            // lane two ways: u_(j+1) by a lane read, the rest through LDS, read
            // back as broadcasts and applied at the start of column j + 1, when
            // they have arrived.  The pivot chain is kept short: pivot j + 1 is
            // formed from t = A(j+1, j+1) before column j's update and u_(j+1)
            // as t - (u_(j+1) / L_jj)^2, two operations after 1/L_jj.
            double up[EB];
            double scp = 0.0;
            double piv = readlane_d(a[0], 0);
#pragma unroll
            for (int j = 0; j < EB; ++j) {
                if (j >= 1) {
                    // applied here, not sunk into a chain in front of a later pivot
#pragma unroll
                    for (int q = j + 1; q < EB; ++q) {
                        a[q] -= scp * up[q];
                        asm volatile("" : "+v"(a[q]));
                    }
                }
                // column j: u holds A(:, j) after the updates of columns < j.
                // For lane j, u * inv = sqrt(pivot) and sc = 1, so the update
                // clobbers only entries of rows <= j that are already final
                // (their y_j and 1/L_jj are kept in yv, iv; a select, not a
                // store by one lane, which would split the loop into blocks).
                const double u = a[j];
                if (j + 2 < EB) {
                    if (rl < EB) cb[EB * j + rl] = u;
                    wave_lds_order();
#pragma unroll
                    for (int q = j + 2; q < EB; ++q) up[q] = cb[EB * j + q];
                    wave_lds_order();
                }
                double t = 0.0, uq = 0.0;
                if (j + 1 < EB) {
                    t = readlane_d(a[j + 1], j + 1);
                    uq = readlane_d(u, j + 1);
                }
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
                if (j + 1 < EB) {
                    a[j + 1] -= sc * uq;
                    const double x = uq * inv;
                    piv = fma(-x, x, t);
                }
                scp = sc;
            }
            // a pivot outside (0, inf) leaves 1/L_jj outside it too
            const bool bad = rl < EB && !(iv > 0.0 && iv < __builtin_inf());
            // the panel stays in LDS (the diagonal factor column-major in the
            // ring); wavefront 3 copies it to global memory in the next phase A,
            // so this wavefront never waits on global stores
            if (valid && (bi != 0 || w == 0)) {
                const bool dg = bi == 0;
                double* dst = dg ? ring + (k & 1) * EBLK + rr : env_slot(win, wrap(rk + bi), rk) + rr * EST;
                const int st = dg ? EST : 1, zc = dg ? rr : EB;
#pragma unroll
                for (int cc = 0; cc < EB; ++cc) dst[cc * st] = (cc > zc) ? 0.0 : a[cc];
                if (dg) dv[EB * k + rr] = iv;
                xs[EB * (k + bi) + rr] = dg ? yv : b;
            }
            if (bad) *sbad = 1;
            // the tiles of panel k - 1 that fall to this wavefront (none unless
            // that panel has more than the helpers take)
            if (((w - nf + 4) & 3) < (mprev - 1) * mprev / 2)
                rest_share(k - 1, rk == 0 ? W - 1 : rk - 1, nf, mprev);
        } else {
            // Helper wavefronts nf .. 3.  Wavefront 3 moves the prefetched block
            // row k + D + 1 into its slot (row k - 2's, free since the last phase
            // A) and prefetches row k + D + 2; the first helper inverts and copies
            // out panel k - 1; then all helpers share the rest of panel k - 1's
            // trailing update.
            const int inew = k + D + 1;
            if (w == 3 && inew < nloc) {
                const int f0 = fst[inew], nbl = inew - f0 + 1, ri = wrap(rk + D + 1);
#pragma unroll
                for (int t = 0; t < EPRE; ++t) {
                    const int e = lane + 64 * t, bb = e >> 8, rr = (e >> 4) & 15, cc = e & 15;
                    if (bb < nbl)
                        env_slot(win, ri, wrap(rk + f0 + bb - k))[rr * EST + cc] =
                            zero_blk(inew, f0 + bb) ? 0.0 : pre[t];
                }
                if (inew + 1 < nloc) prefetch(inew + 1);
            }
            if (k >= 1) {
                const int rkp = rk == 0 ? W - 1 : rk - 1;
                if (w == nf) copy_out(k - 1, rkp);
                rest_share(k - 1, rkp, nf, mk[k - 1]);   // after the copy or the loader
            }
        }
        lds_barrier();
        // ---------------- phase B: tiles (k + i, k + 1), 1 <= i <= m
        for (int t = w; t < m; t += 4) env_tile(win, wrap(rk + 1 + t), wrap(rk + 1), rk, lane);
        lds_barrier();
    }
    if (c.steps > 0) {
        // the last factored column: its remaining trailing update, then its copy
        // by wavefront 3 while the others return
        const int kp = c.steps - 1, rkp = rk == 0 ? W - 1 : rk - 1;
        rest_tiles(kp, rkp, 0, 4);
        lds_barrier();
        if (w == 3) copy_out(kp, rkp);
    } else {
        lds_barrier();
    }
}

// Backward solve of the first `steps` block rows of a local system by one
// wavefront, from the bottom: xs holds the forward solution y of those rows and
// the final solution of the rows after them.  With W_k = L_kk^-1 and
// Q_(k,i) = L_(k+i,k) W_k,
//     x_k = W_k^T y_k - sum_(i>=1) Q_(k,i)^T x_(k+i).
// Lane (c, g) sums entries g, g+4, g+8, g+12 of column c of those blocks; a
// sum over the four rows of the wavefront gives x_k(c).  Only the x_(k+1) term
// is on the sequential chain: the rest of block k - 1's sum is formed while
// block k finishes.  Each block's loads are issued PD blocks ahead (PD
// buffers of 4 (DT + 1) doubles per lane).
template <int DT, int PD>
__device__ void env_backward(const EnvCtx c, const double* winv, const double* Lp, int lane) {
    const int cl = lane & 15, g4 = lane >> 4;
    const int* mk = c.mk;
    double* xs = c.xs;
    constexpr int NQ = 4 * (DT + 1);       // W_k, then Q_(k,1..DT)
    double qb[PD][NQ];
    auto load = [&](int k, double* Q) {
        const int slot = env_wslot(c, max(k, 0));
        const double* Wk = winv + (long long)slot * EB * EB + 4 * lane;
#pragma unroll
        for (int t = 0; t < 4; ++t) Q[t] = Wk[t];
        // blocks past the profile were stored as zeros
#pragma unroll
        for (int i = 1; i <= DT; ++i) {
            const double* Pk = env_panel((double*)Lp, slot, i) + 4 * lane;
#pragma unroll
            for (int t = 0; t < 4; ++t) Q[4 * i + t] = Pk[t];
        }
    };
    // entries g + 4t of row block r of xs, the row clamped to the local system
    // (its zero blocks meet finite values there)
    const int nl1 = c.nloc - 1;
    auto rows = [&](int r, double* v) {
        const double* xr = xs + EB * min(r, nl1) + g4;
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = xr[4 * t];
    };
    // block k's partial sum without the x_(k+1) term, four accumulators
    auto offchain = [&](int k, const double* Q) {
        double v[4], a[4];
        rows(k, v);
#pragma unroll
        for (int t = 0; t < 4; ++t) a[t] = Q[t] * v[t];
#pragma unroll
        for (int i = 2; i <= DT; ++i) {
            rows(k + i, v);
#pragma unroll
            for (int t = 0; t < 4; ++t) a[t] = fma(-Q[4 * i + t], v[t], a[t]);
        }
        return (a[0] + a[1]) + (a[2] + a[3]);
    };
    // x_k from its partial sum p, and the partial sum of block k - 1
    auto step = [&](int k, const double* Q, double p, const double* Qn) {
        double v[4];
        rows(k + 1, v);
        const double pn = k >= 1 ? offchain(k - 1, Qn) : 0.0;
        const double ch = fma(Q[4], v[0], Q[5] * v[1]) + fma(Q[6], v[2], Q[7] * v[3]);
        const double xk = row_sum4(p - ch);
        wave_lds_order();
        if (lane < EB) xs[EB * k + cl] = xk;
        wave_lds_order();
        return pn;
    };
    if (c.steps <= 0) return;
    // buffer q holds block k - q; after block k - q is solved it is refilled
    // with block k - q - PD
#pragma unroll
    for (int q = 0; q < PD; ++q) load(c.steps - 1 - q, qb[q]);
    int k = c.steps - 1;
    double p = offchain(k, qb[0]);
    for (; k >= PD - 1; k -= PD) {
#pragma unroll
        for (int q = 0; q < PD; ++q) {
            p = step(k - q, qb[q], p, qb[(q + 1) % PD]);
            load(k - q - PD, qb[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < PD; ++q)
        if (k - q >= 0) p = step(k - q, qb[q], p, qb[(q + 1) % PD]);
}

// LDS carve-up of one context (window for depth D, nloc block rows)
__device__ __forceinline__ EnvCtx env_ctx(double* base, int D, int nloc, double** end) {
    EnvCtx c;
    c.win = base;
    c.ring = c.win + env_window_slots(D) * EBLK;
    c.colb = c.ring + 2 * EBLK;
    c.xs = c.colb + env_factor_waves(D) * EB * EB;
    c.dv = c.xs + EB * nloc;
    c.mk = (int*)(c.dv + EB * nloc);
    c.fst = c.mk + nloc;
    *end = (double*)(c.fst + ((nloc + 1) & ~1));
    c.nloc = nloc;
    c.steps = nloc;
    c.D = D;
    c.zero_from = nloc;
    c.wrev = 0;
    c.wn = nloc;
    c.preloaded = 0;
    return c;
}

__host__ __device__ inline size_t env_ctx_bytes(int D, int nloc) {
    return (size_t)(env_window_slots(D) + 2) * EBLK * sizeof(double) +
           (size_t)env_factor_waves(D) * EB * EB * sizeof(double) + (size_t)2 * EB * nloc * sizeof(double) +
           (size_t)((nloc + 1) & ~1) * 2 * sizeof(int);
}

// ---- one workgroup per problem: its whole system ---------------------------
template <int DT>
__global__ __launch_bounds__(256, 2) void k_chol_env(const ProbDesc* probs, double* Hb, const int* firstb,
                                                  const double* gb, double* xb, double* winvb, double* Lpb,
                                                  Control* ctls) {
    extern __shared__ double lds[];
    __shared__ int sbad;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const ProbDesc pd = probs[blockIdx.x];
    const int Pp = pd.Pp, D = pd.D;
    double* H = Hb + pd.H_off;
    const int* first = firstb + pd.first_off;
    const double* g = gb + pd.vec_off;
    double* x = xb + pd.vec_off;
    double* winv = winvb + pd.winv_off;
    double* Lp = Lpb + pd.Lp_off;
    Control* ctl = ctls + blockIdx.x;
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int nb = Pp / EB;
    double* end;
    EnvCtx c = env_ctx(lds, D, nb, &end);
    for (int i = tid; i < Pp; i += 256) c.xs[i] = g[i];
    for (int i = tid; i < nb; i += 256) c.fst[i] = first[i];
    if (tid == 0) sbad = 0;
    __syncthreads();
    const EnvSrc src{H, Pp, nb, 0};
    env_forward<DT>(src, c, winv, Lp, &sbad);
    __syncthreads();   // the factor and the inverses in global memory are complete
    if (w == 0) env_backward<DT, 2>(c, winv, Lp, lane);
    __syncthreads();
    for (int i = tid; i < Pp; i += 256) x[i] = c.xs[i];
    if (tid == 0 && sbad) ctl->chol_ok = 0;
}

// ---- split factorisation, kernel 1: top part top-down (workgroup 0) and bottom
// part bottom-up (workgroup 1), each up to the separator [s0, s1) ------------
// Separator export (per half h): blocks (I, J), J <= I < nS, of the local
// system's rows steps + I in S + h nS^2 EB^2, and its right-hand side in
// S + 2 nS^2 EB^2 + h EB nS.
template <int DT>
__global__ __launch_bounds__(256) void k_chol_split1(double* H, int Pp, int D, int s0, int s1, const int* firstT,
                                                     const int* firstB, const double* g, double* x, double* winv,
                                                     double* Lp, double* S, Control* ctl) {
    extern __shared__ double lds[];
    __shared__ int sbad;
    const int tid = threadIdx.x;
    const int half = blockIdx.x;
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int nb = Pp / EB, nS = s1 - s0;
    const int nloc = half ? nb - s0 : s1;
    double* end;
    EnvCtx c = env_ctx(lds, D, nloc, &end);
    c.steps = half ? nb - s1 : s0;
    c.zero_from = half ? c.steps : nloc;   // the bottom half adds only its own update of S
    c.wrev = half;
    c.wn = nb;
    const int* fsrc = half ? firstB : firstT;
    for (int i = tid; i < nloc; i += 256) c.fst[i] = fsrc[i];
    for (int i = tid; i < EB * nloc; i += 256) {
        if (!half) c.xs[i] = g[i];
        else c.xs[i] = i < EB * c.steps ? g[Pp - 1 - i] : 0.0;
    }
    if (tid == 0) sbad = 0;
    __syncthreads();
    const EnvSrc src{H, Pp, nb, half};
    env_forward<DT>(src, c, winv, Lp, &sbad);
    // forward solution of the factored rows, and the separator's blocks and right-hand side
    for (int i = tid; i < EB * c.steps; i += 256) x[half ? Pp - 1 - i : i] = c.xs[i];
    const int W = D + 3;
    double* Sh = S + (long long)half * nS * nS * EB * EB;
    for (int e = tid; e < nS * nS * EB * EB; e += 256) {
        const int blk = e / (EB * EB), I = blk / nS, J = blk % nS, rc = e % (EB * EB);
        if (J <= I) {
            const int i = c.steps + I, j = c.steps + J;
            Sh[e] = env_slot(c.win, i % W, j % W)[(rc / EB) * EST + rc % EB];
        }
    }
    for (int t = tid; t < EB * nS; t += 256) S[(long long)2 * nS * nS * EB * EB + half * EB * nS + t] = c.xs[EB * c.steps + t];
    if (tid == 0 && sbad) ctl->chol_ok = 0;
}

// ---- split factorisation, kernel 2: the separator (sum of both updates,
// factored and solved), then the backward solves of both parts ---------------
template <int DT>
__global__ __launch_bounds__(256) void k_chol_split2(int Pp, int D, int s0, int s1, const int* firstT,
                                                     const int* firstB, double* x, const double* winv,
                                                     const double* Lp, const double* S, double* Sm, double* winvS,
                                                     double* LpS, Control* ctl) {
    extern __shared__ double lds[];
    __shared__ int sbad;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int nb = Pp / EB, nS = s1 - s0, PS = EB * nS;
    const int nT = s1, nB = nb - s0;   // local systems of the two parts, separator included
    double* end;
    EnvCtx cs = env_ctx(lds, nS - 1, nS, &end);
    EnvCtx ct = env_ctx(end, 0, nT, &end);      // only xs, mk, fst of these two are used
    EnvCtx cb = env_ctx(end, 0, nB, &end);
    ct.steps = s0;
    cb.steps = nb - s1;
    cb.wrev = 1;
    cb.wn = nb;
    // separator = top update + bottom update (reversed and transposed): its lower
    // blocks go straight into the window of its factorisation (all nS block rows
    // fit: the window has nS + 2)
    const double* St = S;
    const double* Sb = S + (long long)nS * nS * EB * EB;
    cs.preloaded = 1;
    for (int e0 = 0; e0 < PS * PS; e0 += 16 * 256) {
        // sixteen elements per thread with all loads in flight before the stores
        double va[16], vb[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = e0 + q * 256 + tid;
            const int row = e / PS, col = e % PS;
            va[q] = vb[q] = 0.0;
            if (e < PS * PS && col / EB <= row / EB) {
                const int I = row / EB, J = col / EB, r = row % EB, cc = col % EB;
                const int Ib = nS - 1 - J, Jb = nS - 1 - I;   // the bottom system's block (Ib >= Jb)
                va[q] = St[((long long)(I * nS + J) * EB + r) * EB + cc];
                vb[q] = Sb[((long long)(Ib * nS + Jb) * EB + (EB - 1 - cc)) * EB + (EB - 1 - r)];
            }
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = e0 + q * 256 + tid;
            const int row = e / PS, col = e % PS;
            if (e < PS * PS && col / EB <= row / EB)
                env_slot(cs.win, row / EB, col / EB)[(row % EB) * EST + col % EB] = va[q] + vb[q];
        }
    }
    (void)Sm;
    const double* Sr = S + (long long)2 * nS * nS * EB * EB;
    for (int t = tid; t < PS; t += 256) cs.xs[t] = Sr[t] + Sr[PS + PS - 1 - t];
    for (int i = tid; i < nS; i += 256) cs.fst[i] = 0;
    for (int i = tid; i < nT; i += 256) ct.fst[i] = firstT[i];
    for (int i = tid; i < nB; i += 256) cb.fst[i] = firstB[i];
    if (tid == 0) sbad = 0;
    __syncthreads();
    const EnvSrc ss{Sm, PS, nS, 0};   // not read: the window is preloaded
    env_forward<DT>(ss, cs, winvS, LpS, &sbad);
    env_counts(ct.fst, nT, D, ct.mk);
    env_counts(cb.fst, nB, D, cb.mk);
    __syncthreads();
    if (w == 0) env_backward<DT, DT <= 3 ? 4 : (DT <= 4 ? 3 : 2)>(cs, winvS, LpS, lane);
    __syncthreads();
    // the two parts: forward solutions of their rows, the separator's solution after them
    for (int i = tid; i < EB * s0; i += 256) ct.xs[i] = x[i];
    for (int i = tid; i < EB * (nb - s1); i += 256) cb.xs[i] = x[Pp - 1 - i];
    for (int t = tid; t < PS; t += 256) {
        ct.xs[EB * s0 + t] = cs.xs[t];
        cb.xs[EB * (nb - s1) + t] = cs.xs[PS - 1 - t];
    }
    __syncthreads();
    if (w == 0) env_backward<DT, DT <= 3 ? 4 : (DT <= 4 ? 3 : 2)>(ct, winv, Lp, lane);
    if (w == 1) env_backward<DT, DT <= 3 ? 4 : (DT <= 4 ? 3 : 2)>(cb, winv, Lp, lane);
    __syncthreads();
    for (int i = tid; i < EB * s0; i += 256) x[i] = ct.xs[i];
    for (int t = tid; t < PS; t += 256) x[EB * s0 + t] = cs.xs[t];
    for (int i = tid; i < EB * (nb - s1); i += 256) x[Pp - 1 - i] = cb.xs[i];
    if (tid == 0 && sbad) ctl->chol_ok = 0;
}

size_t chol_band_lds(int Pp, int D) { return env_ctx_bytes(D, Pp / EB); }

size_t chol_split_lds(int Pp, int D, int part) {
    const int nb = Pp / EB, nS = D > 1 ? D : 1;   // the separator has max(D, 1) block rows
    if (part == 1) return env_ctx_bytes(D, nb);   // either half has fewer than nb block rows
    return env_ctx_bytes(nS - 1, nS) + env_ctx_bytes(0, nb) * 2;
}

// compiled profile depths: 3 (C2-like), 4 (ten-frame spans) and EDMAX; the
// depth sizes the row prefetch, the unrolled window loads and the backward
// solve's buffers
template <int DT>
static void band_launch(const ProbDesc* probs, int n_probs, size_t lds, double* H, const int* first,
                        const double* g, double* x, double* winv, double* Lp, Control* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_chol_env<DT>, dim3(n_probs), dim3(256), lds, s, probs, H, first, g, x, winv, Lp, ctl);
}

void launch_chol_band(const ProbDesc* probs, int n_probs, int Dmax, size_t lds, double* H, const int* first,
                      const double* g, double* x, double* winv, double* Lp, Control* ctl, hipStream_t s) {
    if (Dmax <= 3) band_launch<3>(probs, n_probs, lds, H, first, g, x, winv, Lp, ctl, s);
    else if (Dmax == 4) band_launch<4>(probs, n_probs, lds, H, first, g, x, winv, Lp, ctl, s);
    else band_launch<EDMAX>(probs, n_probs, lds, H, first, g, x, winv, Lp, ctl, s);
}

template <int DT>
static void split_launch(double* H, int Pp, int D, int s0, int s1, const int* firstT, const int* firstB,
                         const double* g, double* x, double* winv, double* Lp, double* S, double* Sm, double* winvS,
                         double* LpS, Control* ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_chol_split1<DT>, dim3(2), dim3(256), chol_split_lds(Pp, D, 1), s, H, Pp, D, s0, s1, firstT,
                       firstB, g, x, winv, Lp, S, ctl);
    hipLaunchKernelGGL(k_chol_split2<DT>, dim3(1), dim3(256), chol_split_lds(Pp, D, 2), s, Pp, D, s0, s1, firstT,
                       firstB, x, winv, Lp, S, Sm, winvS, LpS, ctl);
}

void launch_chol_split(double* H, int Pp, int D, int s0, int s1, const int* firstT, const int* firstB,
                       const double* g, double* x, double* winv, double* Lp, double* S, double* Sm, double* winvS,
                       double* LpS, Control* ctl, hipStream_t s) {
    if (D <= 3) split_launch<3>(H, Pp, D, s0, s1, firstT, firstB, g, x, winv, Lp, S, Sm, winvS, LpS, ctl, s);
    else if (D == 4) split_launch<4>(H, Pp, D, s0, s1, firstT, firstB, g, x, winv, Lp, S, Sm, winvS, LpS, ctl, s);
    else split_launch<EDMAX>(H, Pp, D, s0, s1, firstT, firstB, g, x, winv, Lp, S, Sm, winvS, LpS, ctl, s);
}

int prepare_chol_band(int Dmax, size_t lds) {
    const void* f = Dmax <= 3    ? (const void*)k_chol_env<3>
                    : Dmax == 4 ? (const void*)k_chol_env<4>
                                : (const void*)k_chol_env<EDMAX>;
    return (int)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

int prepare_chol_split(int Pp, int D) {
    const void* f1 = D <= 3    ? (const void*)k_chol_split1<3>
                     : D == 4 ? (const void*)k_chol_split1<4>
                              : (const void*)k_chol_split1<EDMAX>;
    const void* f2 = D <= 3    ? (const void*)k_chol_split2<3>
                     : D == 4 ? (const void*)k_chol_split2<4>
                              : (const void*)k_chol_split2<EDMAX>;
    int rc = (int)hipFuncSetAttribute(f1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)chol_split_lds(Pp, D, 1));
    if (rc) return rc;
    return (int)hipFuncSetAttribute(f2, hipFuncAttributeMaxDynamicSharedMemorySize, (int)chol_split_lds(Pp, D, 2));
}

int chol_band_max_blocks() { return EDMAX; }
size_t chol_panel_doubles(int Pp) { return (size_t)(Pp / EB) * EDMAX * EB * EB; }
int chol_band_block() { return EB; }

// ---------------------------------------------------------------------------
// L L^T x = g after the dense blocked factorisation (L in H's lower triangle,
// the inverse diagonal blocks in Linv): one workgroup per 32-row block row,
// pipelined across the chip through per-row ready flags.
// ---------------------------------------------------------------------------
// Forward, L y = g: block row i needs y_j for the block columns first[i] ..
// i - 1 of its profile.  Backward, L^T x = y: block column i needs x_j for the
// block rows j > i whose profile reaches it (the column's active list).  A
// workgroup takes its block row from a ticket counter, so it only waits on
// rows whose workgroups started before it: the pipeline cannot deadlock,
// whatever the number of resident workgroups.  The far dependencies (ready
// first) are summed while the near ones finish; the factor blocks of the eight
// nearest are loaded into registers before any wait, and those eight are taken
// together once all their flags are up.  A row publishes its solution, then
// its flag; a reader waits for the flags (one thread per flag), meets the
// others at a barrier, and then reads the solutions (the exchange is described
// at xget).  The waits are bounded: past the bound the row is marked failed
// (the Cholesky flag) and taken as ready, so every workgroup exits.
// sync: [0] forward ticket, [1] backward ticket, [2, 2 + nb) forward flags,
// [2 + nb, 2 + 2 nb) backward flags; cleared by k_solve_reset before each use.
constexpr int SOLVE_SPIN_MAX = 1 << 20;
constexpr int SOLVE_NEAR = 8;    // nearest dependencies, preloaded and taken together
constexpr int SOLVE_FAR = 64;    // far dependencies whose flags are checked together
constexpr int SOLVE_GROUP = 4;   // far dependency blocks read together

// Everything one workgroup hands to another (the rows' solutions and flags)
// moves by relaxed atomic loads and stores at device scope.  Those are
// coherent across the XCDs' L2 caches by themselves, so no acquire or release
// (which would write back or invalidate a whole L2) is needed: a row waits
// until its own solution stores have completed before it raises its flag, and
// a reader loads the solution after it has seen the flag.
__device__ __forceinline__ double xget(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xput(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Once the Cholesky flag is down (a failed pivot, or a wait past its bound)
// the solution is discarded, and no row waits any longer.
__device__ __forceinline__ void solve_wait(const int* flag, Control* ctl) {
    // the Cholesky flag by device-coherent atomics (global accesses, and another
    // workgroup's write is seen)
    int* ok = &ctl->chol_ok;
    int n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (__hip_atomic_load(ok, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        if (++n > SOLVE_SPIN_MAX) {
            __hip_atomic_store(ok, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// wait for the flags of dependencies [0, n) (n <= 64), one thread each, then
// let the whole workgroup read what they published
template <typename F>
__device__ __forceinline__ void solve_wait_set(F flag_of, int n, int tid, Control* ctl) {
    if (tid < n) solve_wait(flag_of(tid), ctl);
    __syncthreads();
}

// wavefront 0 publishes the row it has stored: once all of its stores have
// completed (vmcnt 0 counts the whole wavefront's), lane 0 raises the flag
__device__ __forceinline__ void solve_publish(int* flag, int tid) {
    if (tid < 64) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (tid == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(64) void k_solve_reset(int* sync, int n, const Control* ctl) {
    if (ctl->pcg_fail == 0) return;
    for (int i = threadIdx.x; i < n; i += 64) sync[i] = 0;
}

__global__ __launch_bounds__(256) void k_solve_fwd(const double* H, const double* Linv, const double* g, double* y,
                                                   int Pp, const int* first, int* sync, Control* ctl) {
    __shared__ int srow;
    __shared__ double rr[NB];
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int tid = threadIdx.x;
    int* flag = sync + 2;
    if (tid == 0) srow = __hip_atomic_fetch_add(sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int i = srow;
    const int row = tid >> 3, c4 = (tid & 7) * 4;   // row of the block, four of its columns
    const double* Hr = H + (long long)(i * NB + row) * Pp + c4;
    const int jlo = first[i];
    const int nnear = min(SOLVE_NEAR, i - jlo), jn = i - nnear;   // near dependencies: jn .. i - 1
    double ln[SOLVE_NEAR][4];
#pragma unroll
    for (int u = 0; u < SOLVE_NEAR; ++u) {
        const int j = min(jn + u, max(i - 1, 0));
#pragma unroll
        for (int q = 0; q < 4; ++q) ln[u][q] = Hr[j * NB + q];
    }
    double acc = 0.0;
    for (int j0 = jlo; j0 < jn; j0 += SOLVE_FAR) {
        const int j1 = min(j0 + SOLVE_FAR, jn);
        solve_wait_set([&](int t) { return flag + j0 + t; }, j1 - j0, tid, ctl);
        for (int b0 = j0; b0 < j1; b0 += SOLVE_GROUP) {
            double lv[SOLVE_GROUP][4], yv[SOLVE_GROUP][4];
#pragma unroll
            for (int u = 0; u < SOLVE_GROUP; ++u) {
                const int j = min(b0 + u, j1 - 1);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    lv[u][q] = Hr[j * NB + q];
                    yv[u][q] = xget(y + j * NB + c4 + q);
                }
            }
#pragma unroll
            for (int u = 0; u < SOLVE_GROUP; ++u)
                if (b0 + u < j1)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc = fma(lv[u][q], yv[u][q], acc);
        }
    }
    solve_wait_set([&](int t) { return flag + jn + t; }, nnear, tid, ctl);
    {
        double yv[SOLVE_NEAR][4];
#pragma unroll
        for (int u = 0; u < SOLVE_NEAR; ++u) {
            const int j = min(jn + u, max(i - 1, 0));
#pragma unroll
            for (int q = 0; q < 4; ++q) yv[u][q] = xget(y + j * NB + c4 + q);
        }
#pragma unroll
        for (int u = 0; u < SOLVE_NEAR; ++u)
            if (u < nnear)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = fma(ln[u][q], yv[u][q], acc);
    }
    // the eight lanes of a row are consecutive
    acc += __shfl_xor(acc, 1, 8);
    acc += __shfl_xor(acc, 2, 8);
    acc += __shfl_xor(acc, 4, 8);
    if ((tid & 7) == 0) rr[row] = g[i * NB + row] - acc;
    __syncthreads();
    if (tid < NB) {
        const double* Li = Linv + (long long)i * NB * NB + tid * NB;
        double v = 0.0;
        for (int q = 0; q <= tid; ++q) v = fma(Li[q], rr[q], v);
        xput(y + i * NB + tid, v);
    }
    solve_publish(flag + i, tid);
}

__global__ __launch_bounds__(256) void k_solve_bwd(const double* H, const double* Linv, const double* y, double* x,
                                                   int Pp, const int* act_off, const int* act, int* sync,
                                                   Control* ctl) {
    __shared__ int srow;
    __shared__ double part[8][NB];
    __shared__ double rr[NB];
    if (ctl->pcg_fail == 0) return;   // the CG solve of this iteration converged
    const int tid = threadIdx.x, nb = Pp / NB;
    int* flag = sync + 2 + nb;
    if (tid == 0) srow = __hip_atomic_fetch_add(sync + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int i = nb - 1 - srow;
    const int c = tid & 31, q4 = (tid >> 5) * 4;   // column of the block, four of the dependency's rows
    const double* Hc = H + (long long)q4 * Pp + i * NB + c;
    // the column's active rows ascend: the first ones are the nearest
    const int a0 = act_off[i], a1 = act_off[i + 1];
    const int nnear = min(SOLVE_NEAR, a1 - a0), an = a0 + nnear;   // near dependencies: act[a0 .. an)
    double ln[SOLVE_NEAR][4];
#pragma unroll
    for (int u = 0; u < SOLVE_NEAR; ++u) {
        const int j = a1 > a0 ? act[min(a0 + u, a1 - 1)] : i;   // the last column has no rows below
#pragma unroll
        for (int q = 0; q < 4; ++q) ln[u][q] = Hc[(long long)(j * NB + q) * Pp];
    }
    double acc = 0.0;
    // the far rows, the farthest (ready first) first
    for (int e1 = a1; e1 > an; e1 -= SOLVE_FAR) {
        const int e0 = max(e1 - SOLVE_FAR, an);
        solve_wait_set([&](int t) { return flag + act[e0 + t]; }, e1 - e0, tid, ctl);
        for (int b1 = e1; b1 > e0; b1 -= SOLVE_GROUP) {
            double lv[SOLVE_GROUP][4], xv[SOLVE_GROUP][4];
#pragma unroll
            for (int u = 0; u < SOLVE_GROUP; ++u) {
                const int j = act[max(b1 - 1 - u, e0)];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    lv[u][q] = Hc[(long long)(j * NB + q) * Pp];
                    xv[u][q] = xget(x + j * NB + q4 + q);
                }
            }
#pragma unroll
            for (int u = 0; u < SOLVE_GROUP; ++u)
                if (b1 - 1 - u >= e0)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc = fma(lv[u][q], xv[u][q], acc);
        }
    }
    solve_wait_set([&](int t) { return flag + act[a0 + t]; }, nnear, tid, ctl);
    {
        double xv[SOLVE_NEAR][4];
#pragma unroll
        for (int u = 0; u < SOLVE_NEAR; ++u) {
            const int j = a1 > a0 ? act[min(a0 + u, a1 - 1)] : i;
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[u][q] = xget(x + j * NB + q4 + q);
        }
#pragma unroll
        for (int u = 0; u < SOLVE_NEAR; ++u)
            if (u < nnear)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc = fma(ln[u][q], xv[u][q], acc);
    }
    part[tid >> 5][c] = acc;
    __syncthreads();
    if (tid < NB) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) s += part[k][tid];
        rr[tid] = y[i * NB + tid] - s;
    }
    __syncthreads();
    if (tid < NB) {
        const double* Li = Linv + (long long)i * NB * NB + tid;
        double v = 0.0;
        for (int q = tid; q < NB; ++q) v = fma(Li[q * NB], rr[q], v);
        xput(x + i * NB + tid, v);
    }
    solve_publish(flag + i, tid);
}

void launch_chol_solve(const double* H, const double* Linv, const double* g, double* delta, int Pp,
                       const int* first, const int* act_off, const int* act, double* ywork, int* sync,
                       Control* ctl, hipStream_t s) {
    const int nb = Pp / NB;
    hipLaunchKernelGGL(k_solve_reset, dim3(1), dim3(64), 0, s, sync, 2 + 2 * nb, ctl);
    hipLaunchKernelGGL(k_solve_fwd, dim3(nb), dim3(256), 0, s, H, Linv, g, ywork, Pp, first, sync, ctl);
    hipLaunchKernelGGL(k_solve_bwd, dim3(nb), dim3(256), 0, s, H, Linv, ywork, delta, Pp, act_off, act, sync, ctl);
}

}  // namespace spgo
