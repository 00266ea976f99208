// Residual / Jacobian evaluation, block assembly and the landmark Schur
// contraction of one LM linearisation (the work of Solver_cuda.cu:49-345 and
// the SGEMMs of Solver_cuda.cu:636-712, restructured).
//
// One workgroup owns one landmark group: landmarks that share an anchor frame
// a.  Everything is expressed in the anchor frame, where the reference's pose
// Jacobian factors as
//     dz/dxi_i = J_a Ad(G_a^-1 G_i),   J_a = M R_{a->t} [-[P_a]x, I]
// for every state i in [a, t) (Solver_cuda.cu:248-290).  So per measurement
// only J_a is formed, and the group reduces
//   * pair sums S(t) = sum J_a^T J_a, s(t) = sum J_a^T r per target frame t,
//     whose suffix sums over t give the pose-pose block Q(i,j) = sum_{t>max(i,j)} S(t),
//   * per landmark C, g_alpha and v_l(i) = sum_{t>i} J_a^T j_alpha (its B column),
//   * Y = sum_l v_l v_l^T / (C_l + eps) and y = sum_l v_l g_l / (C_l + eps)
//     on MFMA (the landmark Schur contraction, Solver_cuda.cu:690-712),
// and writes D = Q - Y, d = q - y mapped to the incremental-pose parameters
// with the block-diagonal adjoints Ad(G_a^-1 G_i).  The pose system is the sum
// of these blocks over the groups that cover each pose pair (k_form_system).  Reductions are deterministic: wavefront-private
// LDS partials combined in a fixed order, no global atomics.
#include "spgo_device.h"
#include "spgo_kernels.h"

namespace spgo {

namespace {

constexpr int VS = 66;                 // LDS row stride of the V tile (doubles)
constexpr int VR = 6 * KMAX;           // rows of the V tile (96)
constexpr int SCR = 27;                // per-lane stride of the pair scratch
constexpr int NWAVE = WG / 64;
constexpr int VSH = (VR * VR > NWAVE * 64 * SCR) ? VR * VR : NWAVE * 64 * SCR;   // >= VR * VS
constexpr int MAXTILE = ((VR / 16) * (VR / 16 + 1) / 2 + NWAVE - 1) / NWAVE;   // tiles per wave

template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef double __attribute__((ext_vector_type(4))) type; };
template <> struct Vec4<float> { typedef float __attribute__((ext_vector_type(4))) type; };

__device__ __forceinline__ Vec4<double>::type mfma16x16x4(double a, double b, Vec4<double>::type c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Vec4<float>::type mfma16x16x4(float a, float b, Vec4<float>::type c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ void robust_eval(const Robust& rb, T s, T kappa, T& w, double& cost) {
    T rho, drho;
    if (rb.kind == 0) {
        const T c2 = (T)rb.param;
        rho = c2 * log1p(s / c2);
        drho = (T)1 / ((T)1 + s / c2);
    } else if (rb.kind == 1) {
        const T d = (T)rb.param;
        const T rs = sqrt(s);
        if (rs <= d) {
            rho = s;
            drho = (T)1;
        } else {
            rho = (T)2 * d * rs - d * d;
            drho = d / rs;
        }
    } else {
        rho = s;
        drho = (T)1;
    }
    w = sqrt(kappa * drho);
    cost += 0.5 * (double)kappa * (double)rho;
}

}  // namespace

template <typename T>
__global__ __launch_bounds__(WG) void k_assemble(AsmArgs A, StateBuf<T> B0, StateBuf<T> B1) {
    const Layout& lay = A.lay;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;

    const int cur = A.ctl->cur;
    const StateBuf<T>& bin = cur ? B1 : B0;                    // linearisation point
    const StateBuf<T>& bout = A.init ? bin : (cur ? B0 : B1);  // state being linearised
    const bool update = !A.init;
    const bool chol_ok = update ? (A.ctl->chol_ok != 0) : true;

    const int a = lay.grp_anchor[g];
    const int K = lay.grp_K[g];
    const int K6 = 6 * K;
    const int nT = (K6 + 15) >> 4;
    const int nTiles = nT * (nT + 1) / 2;

    __shared__ T S_w[NWAVE][KMAX + 1][27];      // per-wave pair partials, slot = t - a
    __shared__ T Trel[2][KMAX + 1][12];         // (G^c_{a+s})^-1 G_a of the state being linearised
    __shared__ T dl[KMAX][6];                   // Ad(G_a^-1 G_{a+k}) delta_{a+k} (back-substitution)
    __shared__ T stage[CHUNK_OBS][8];           // per measurement: J_a^T j_alpha (6), j_alpha^2, j_alpha r
    __shared__ int stslot[CHUNK_OBS];
    __shared__ T Vsh[VSH];
    __shared__ int sslot[NWAVE][64];
    __shared__ T gt[CHUNK_LM];
    __shared__ T lmP[CHUNK_LM][6];              // P_a and -P_a/alpha in the anchor frame
    __shared__ double red[NWAVE];
    __shared__ double Adk[KMAX][36];            // Ad(G_a^-1 G_{a+k}) of the state being linearised

    // ---- group set-up ------------------------------------------------------
    for (int i = tid; i < NWAVE * (KMAX + 1) * 27; i += WG) (&S_w[0][0][0])[i] = (T)0;
    for (int i = tid; i < VR * VS; i += WG) Vsh[i] = (T)0;
    if (tid < 2 * (K + 1)) {
        const int c = tid / (K + 1), s = tid % (K + 1);
        double Ga[12], Gti[12], Tr[12];
        pose_load(bout.G + 12 * a, Ga);
        pose_load((c ? bout.GinvR : bout.GinvL) + 12 * (a + s), Gti);
        pose_mul(Gti, Ga, Tr);
#pragma unroll
        for (int i = 0; i < 12; ++i) Trel[c][s][i] = (T)Tr[i];
    }
    if (tid >= 64 && tid < 64 + K) {
        const int k = tid - 64;
        double Gai[12], Gi[12], E[12];
        pose_load(bout.GinvL + 12 * a, Gai);
        pose_load(bout.G + 12 * (a + k), Gi);
        pose_mul(Gai, Gi, E);
        for (int e = 0; e < 36; ++e) Adk[k][e] = adjoint_entry(E, e / 6, e % 6);
    }
    if (update && tid < K) {
        double Gai[12], Gi[12], E[12], x[6], out[6];
        pose_load(bin.GinvL + 12 * a, Gai);
        pose_load(bin.G + 12 * (a + tid), Gi);
        pose_mul(Gai, Gi, E);
#pragma unroll
        for (int q = 0; q < 6; ++q) x[q] = A.delta[6 * (a + tid) + q];
        adjoint_apply(E, x, out);
#pragma unroll
        for (int q = 0; q < 6; ++q) dl[tid][q] = (T)out[q];
    }

    typename Vec4<T>::type acc[MAXTILE];
#pragma unroll
    for (int i = 0; i < MAXTILE; ++i) acc[i] = (T)0;
    T yr = (T)0;
    double cost = 0.0;

    const T fx0 = (T)A.cam.fx[0], fy0 = (T)A.cam.fy[0], cx0 = (T)A.cam.cx[0], cy0 = (T)A.cam.cy[0];
    const int2 cr = lay.grp_chunks[g];
    const int glm0 = lay.grp_lm_begin[g];
    const long long voff = lay.grp_v_off[g];
    __syncthreads();

    for (int ci = cr.x; ci < cr.y; ++ci) {
        const int4 ch = lay.chunk[ci];
        const int lm0 = ch.x, nl = ch.y, ob0 = ch.z, nob = ch.w;

        // ---- phase 0: per landmark: back-substitution and anchor-frame point ----
        if (tid < nl) {
            const int act = lm0 + tid;
            const int id = lay.lm_id[act];
            double al = bin.alpha[id];
            if (update) {
                double dal = 0.0;
                if (chol_ok) {
                    const T* v = bin.v + voff + (long long)(act - glm0) * K6;
                    double s = (double)bin.g[act];
                    for (int k = 0; k < K; ++k)
#pragma unroll
                        for (int q = 0; q < 6; ++q) s -= (double)v[6 * k + q] * (double)dl[k][q];
                    dal = s / (double)bin.Ct[act];
                }
                al = fmin(fmax(al + dal, A.alpha_min), A.alpha_max);
                bout.alpha[id] = al;
            }
            const float2 ray = lay.lm_ray[act];
            const double x = ((double)ray.x - A.cam.cx[0]) / A.cam.fx[0];
            const double y = ((double)ray.y - A.cam.cy[0]) / A.cam.fy[0];
            const double P0 = x / al, P1 = y / al, P2 = 1.0 / al;
            lmP[tid][0] = (T)P0;
            lmP[tid][1] = (T)P1;
            lmP[tid][2] = (T)P2;
            lmP[tid][3] = (T)(-P0 / al);
            lmP[tid][4] = (T)(-P1 / al);
            lmP[tid][5] = (T)(-P2 / al);
        }
        __syncthreads();

        // ---- phase 1: per measurement ----------------------------------------
        for (int base = 0; base < nob; base += WG) {
            const int o = base + tid;
            const bool valid = o < nob;
            T pr[27];
#pragma unroll
            for (int q = 0; q < 27; ++q) pr[q] = (T)0;
            int slot = 0;
            if (valid) {
                const int t = lay.ob_t[ob0 + o];
                const uint32_t meta = lay.ob_meta[ob0 + o];
                const int pos = (int)(meta & 0xffffu);
                const int j = (int)(meta >> 16);
                const float4 uv = lay.ob_uv[ob0 + o];
                slot = t > a ? t - a : 0;
                const T Pa[3] = {lmP[j][0], lmP[j][1], lmP[j][2]};
                const T dv[3] = {lmP[j][3], lmP[j][4], lmP[j][5]};
                const T kappa = (A.rb.near_boost && abs(t - a) < 2) ? (T)16 : (T)1;
                T u[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
                T cc = (T)0, gg = (T)0;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    T R[12];
                    if (t >= a) {
#pragma unroll
                        for (int i = 0; i < 12; ++i) R[i] = Trel[c][t - a][i];
                    } else {
                        double Ga[12], Gti[12], Tr[12];
                        pose_load(bout.G + 12 * a, Ga);
                        pose_load((c ? bout.GinvR : bout.GinvL) + 12 * t, Gti);
                        pose_mul(Gti, Ga, Tr);
#pragma unroll
                        for (int i = 0; i < 12; ++i) R[i] = (T)Tr[i];
                    }
                    const T fx = c ? (T)A.cam.fx[1] : fx0, fy = c ? (T)A.cam.fy[1] : fy0;
                    const T cx = c ? (T)A.cam.cx[1] : cx0, cy = c ? (T)A.cam.cy[1] : cy0;
                    const T Px = R[0] * Pa[0] + R[1] * Pa[1] + R[2] * Pa[2] + R[3];
                    const T Py = R[4] * Pa[0] + R[5] * Pa[1] + R[6] * Pa[2] + R[7];
                    const T Pz = R[8] * Pa[0] + R[9] * Pa[1] + R[10] * Pa[2] + R[11];
                    const T iz = (T)1 / Pz;
                    const T ex = (c ? (T)uv.z : (T)uv.x) - (fx * Px * iz + cx);
                    const T ey = (c ? (T)uv.w : (T)uv.y) - (fy * Py * iz + cy);
                    T w;
                    robust_eval(A.rb, ex * ex + ey * ey, kappa, w, cost);
                    // rows of M R_{a->t}: M = K_c Pi'(P_t) (LMutils.cu:615-628)
                    const T m0a = fx * iz, m0c = -fx * Px * iz * iz;
                    const T m1b = fy * iz, m1c = -fy * Py * iz * iz;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        T n[3];
                        if (k == 0) {
                            n[0] = m0a * R[0] + m0c * R[8];
                            n[1] = m0a * R[1] + m0c * R[9];
                            n[2] = m0a * R[2] + m0c * R[10];
                        } else {
                            n[0] = m1b * R[4] + m1c * R[8];
                            n[1] = m1b * R[5] + m1c * R[9];
                            n[2] = m1b * R[6] + m1c * R[10];
                        }
                        // weighted row [P_a x n, n] (the -[P]x of compute_del_tn_del_xi)
                        T J[6];
                        J[0] = w * (Pa[1] * n[2] - Pa[2] * n[1]);
                        J[1] = w * (Pa[2] * n[0] - Pa[0] * n[2]);
                        J[2] = w * (Pa[0] * n[1] - Pa[1] * n[0]);
                        J[3] = w * n[0];
                        J[4] = w * n[1];
                        J[5] = w * n[2];
                        const T ja = w * (n[0] * dv[0] + n[1] * dv[1] + n[2] * dv[2]);
                        const T rk = w * (k == 0 ? ex : ey);
                        int idx = 0;
#pragma unroll
                        for (int p = 0; p < 6; ++p) {
#pragma unroll
                            for (int q = p; q < 6; ++q) pr[idx++] += J[p] * J[q];
                            pr[21 + p] += J[p] * rk;
                            u[p] += J[p] * ja;
                        }
                        cc += ja * ja;
                        gg += ja * rk;
                    }
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) stage[pos][q] = u[q];
                stage[pos][6] = cc;
                stage[pos][7] = gg;
                stslot[pos] = slot;
            }
            // pair sums: per-wave segmented reduction over lanes (deterministic)
            T* scr = Vsh + (wv * 64 + lane) * SCR;
#pragma unroll
            for (int q = 0; q < 27; ++q) scr[q] = pr[q];
            sslot[wv][lane] = slot;
            __syncthreads();
            if (lane < 27) {
                const T* col = Vsh + wv * 64 * SCR + lane;
                T s = (T)0;
                int cs = sslot[wv][0];
                for (int jj = 0; jj < 64; ++jj) {
                    const int sj = sslot[wv][jj];
                    if (sj != cs) {
                        if (cs > 0) S_w[wv][cs][lane] += s;
                        s = (T)0;
                        cs = sj;
                    }
                    s += col[jj * SCR];
                }
                if (cs > 0) S_w[wv][cs][lane] += s;
            }
            __syncthreads();
        }

        // ---- phase 2: per landmark: C, g_alpha, B column v_l -------------------
        for (int i = tid; i < VR * VS; i += WG) Vsh[i] = (T)0;   // scratch -> V tile
        __syncthreads();
        if (tid < nl) {
            const int act = lm0 + tid;
            const int2 lp = lay.lm_pos[act];
            T C = (T)0, G = (T)0;
            for (int p = lp.x; p < lp.x + lp.y; ++p) {
                C += stage[p][6];
                G += stage[p][7];
            }
            const T Ct = C + (T)A.eps;
            const T isc = (T)1 / sqrt(Ct);
            T* vo = bout.v + voff + (long long)(act - glm0) * K6;
            T s6[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
            int p = lp.x + lp.y - 1;
            for (int k = K - 1; k >= 0; --k) {
                while (p >= lp.x && stslot[p] > k) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) s6[q] += stage[p][q];
                    --p;
                }
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    vo[6 * k + q] = s6[q];
                    Vsh[(6 * k + q) * VS + tid] = s6[q] * isc;
                }
            }
            bout.Ct[act] = Ct;
            bout.g[act] = G;
            gt[tid] = G * isc;
        } else if (tid < CHUNK_LM) {
            gt[tid] = (T)0;
        }
        __syncthreads();

        // ---- phase 3: Y += V V^T (MFMA), y += V gt -----------------------------
        const int kend = (nl + 3) & ~3;
#pragma unroll
        for (int ti = 0; ti < MAXTILE; ++ti) {
            const int tile = wv + ti * NWAVE;
            if (tile < nTiles) {
                int r = 0;
                while ((r + 1) * (r + 2) / 2 <= tile) ++r;
                const int c = tile - r * (r + 1) / 2;
                const T* Ar = Vsh + (r * 16 + (lane & 15)) * VS + (lane >> 4);
                const T* Br = Vsh + (c * 16 + (lane & 15)) * VS + (lane >> 4);
                for (int kk = 0; kk < kend; kk += 4) acc[ti] = mfma16x16x4(Ar[kk], Br[kk], acc[ti]);
            }
        }
        if (tid < K6) {
            const T* row = Vsh + tid * VS;
            for (int jj = 0; jj < nl; ++jj) yr += row[jj] * gt[jj];
        }
        __syncthreads();
    }

    // ---- finalize: D = Q - Y, d = q - y ----------------------------------------
    T* Rsuf = &stage[0][0];            // [KMAX][27] suffix sums of the pair partials
    if (tid < 27) {
        T run = (T)0;
        for (int s = K; s >= 1; --s) {
#pragma unroll
            for (int w = 0; w < NWAVE; ++w) run += S_w[w][s][tid];
            Rsuf[(s - 1) * 27 + tid] = run;
        }
    }
    __syncthreads();
    // D = Q - Y into LDS (row stride 6K), d = q - y beside it
    T* Dl = Vsh;
    T* dlv = &stage[0][0] + KMAX * 27;
#pragma unroll
    for (int ti = 0; ti < MAXTILE; ++ti) {
        const int tile = wv + ti * NWAVE;
        if (tile < nTiles) {
            int r = 0;
            while ((r + 1) * (r + 2) / 2 <= tile) ++r;
            const int c = tile - r * (r + 1) / 2;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = r * 16 + mfma16_row((T)0, lane, v);
                const int col = c * 16 + (lane & 15);
                if (row < K6 && col < K6) {
                    const int kb = max(row / 6, col / 6);
                    const T q = Rsuf[kb * 27 + sym6(row % 6, col % 6)];
                    const T val = q - acc[ti][v];
                    Dl[row * K6 + col] = val;
                    Dl[col * K6 + row] = val;
                }
            }
        }
    }
    if (tid < K6) dlv[tid] = Rsuf[(tid / 6) * 27 + 21 + (tid % 6)] - yr;
    __syncthreads();
    // to incremental-pose coordinates: block (i, j) <- Ad_i^T D(i, j) Ad_j, d_i <- Ad_i^T d_i,
    // Ad_k = Ad(G_a^-1 G_{a+k}); lower blocks only (the reduced system reads i >= j)
    T* Dg = bout.D + lay.grp_D_off[g];
    for (int pr = tid; pr < K * (K + 1) / 2; pr += WG) {
        int bi = 0;
        while ((bi + 1) * (bi + 2) / 2 <= pr) ++bi;
        const int bj = pr - bi * (bi + 1) / 2;
        double M1[36];
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) s += (double)Dl[(6 * bi + r) * K6 + 6 * bj + q] * Adk[bj][6 * q + c];
                M1[6 * r + c] = s;
            }
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 6; ++q) s += Adk[bi][6 * q + r] * M1[6 * q + c];
                Dg[(long long)(6 * bi + r) * K6 + 6 * bj + c] = (T)s;
            }
    }
    if (tid < K6) {
        const int k = tid / 6, r = tid % 6;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) s += Adk[k][6 * q + r] * (double)dlv[6 * k + q];
        bout.d[lay.grp_d_off[g] + tid] = (T)s;
    }
    cost = wave_sum(cost);
    if (lane == 0) red[wv] = cost;
    __syncthreads();
    if (tid == 0) {
        double c = 0.0;
        for (int w = 0; w < NWAVE; ++w) c += red[w];
        bout.cost[g] = c;
    }
}

template <typename T>
void launch_assemble(const AsmArgs& a, const StateBuf<T>* bufs, hipStream_t s) {
    if (a.lay.n_groups == 0) return;
    hipLaunchKernelGGL(k_assemble<T>, dim3(a.lay.n_groups), dim3(WG), 0, s, a, bufs[0], bufs[1]);
}

template void launch_assemble<double>(const AsmArgs&, const StateBuf<double>*, hipStream_t);
template void launch_assemble<float>(const AsmArgs&, const StateBuf<float>*, hipStream_t);

}  // namespace spgo
