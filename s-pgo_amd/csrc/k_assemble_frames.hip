// Linearisation in frame coordinates: the work of Solver_cuda.cu:49-345 and
// the SGEMMs of Solver_cuda.cu:636-712 for layouts whose landmarks span many
// poses (loop revisits, long tracks).
//
// The reduced system is written in the variables z_k = sum_{i>=k} Ad(G_i)
// delta_i (z_N = 0), an exact change of variables of the incremental pose
// increments delta.  The reference's pose Jacobian of a measurement with anchor
// a and target t > a, dz/dxi_i = J_a Ad(G_a^-1 G_i) for every i in [a, t)
// (Solver_cuda.cu:248-290), becomes J_w (z_a - z_t) with J_w = J_a Ad(G_a^-1):
// one block at the anchor frame and one at the target frame, however far apart
// they are.  A landmark group (landmarks of one anchor, one LDS chunk) keeps
// its slots for the anchor and its distinct target frames f_1 < ... < f_m, and
// reduces in the anchor frame as the incremental kernel does:
//   * pair sums S(s) = sum J_a^T J_a, s(s) = sum J_a^T r per target slot s,
//   * per landmark C, g_alpha and u_l(s) = sum J_a^T j_alpha per slot,
// then forms, still in the anchor's axes, the target-frame block and gradient
//   B_a = diag(S) - sum_l u_l u_l^T / (C_l + eps),  b_a = s - sum_l u_l g_l / (C_l + eps)
// (6m x 6m and 6m; the second terms one MFMA product of the rows
// [u_1; ..; u_m; g] scaled by 1 / sqrt(C + eps)), and maps them to world axes
// in double: B = Ad(G_a^-1)^T B_a Ad(G_a^-1), b = Ad(G_a^-1)^T b_a.  The world
// axes carry lever arms of the trajectory's size, so the mapped blocks are kept
// in double whatever the precision of the reduction.  With P = [1 | -I] the
// group's part of the system is P^T B P and its gradient P^T b: the anchor's
// blocks are the row sums R(s) = sum_s' B(s, s'), their total and the gradient
// total, formed here in double from the stored blocks, so the anchor rows
// cancel the target rows exactly in the system.  k_form_frames sums them.
// The back-substitution of the next linearisation needs only z_a - z_{f_s}:
// delta_alpha = (g_alpha - sum_s u_l(s) . Ad(G_a^-1) (z_a - z_{f_s})) / (C + eps).
#include "spgo_asm_common.h"
#include "spgo_device.h"
#include "spgo_kernels.h"

namespace spgo {

namespace {

// LDS shapes for groups of at most KT target frames
template <typename T, int KT>
struct FrShape {
    // landmarks per group: 32 in FP64 at 16 frames, so that two workgroups fit a CU
    static constexpr int CL = KT <= 16 ? (sizeof(T) == 8 ? 32 : 64) : (KT <= 32 ? 32 : 8);
    static constexpr int VS = CL + 2;                                 // V tile row stride
    static constexpr int K6 = 6 * KT;                                  // block rows: the target frames
    static constexpr int VROWS = 16 * ((K6 + 1 + 15) / 16);            // and the gradient row, whole tiles
    static constexpr int US = 6 * KT + 1;                              // U row stride (odd: banks)
};

}  // namespace

template <typename T, int KT>
__global__ __launch_bounds__(WG, KT <= 16 ? 2 : 1) void k_assemble_frames(AsmArgs A, StateBuf<T> B0, StateBuf<T> B1) {
    using SH = FrShape<T, KT>;
    constexpr int CL = SH::CL, VS = SH::VS, US = SH::US;
    const Layout& lay = A.lay;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;

    // every global load is issued before the kernel's first store (a load after
    // a store waits for it)
    const GroupDesc* gdp = lay.grp_desc + g;
    const int a = gdp->anchor;
    const int m = gdp->K;                      // target slots 1..m
    const int prob = gdp->prob;
    const int glm0 = gdp->lm_begin, d_off = gdp->d_off, gid = gdp->gid;
    const long long voff = gdp->v_off, D_off = gdp->D_off, delta_off = gdp->delta_off, r_off = gdp->r_off;
    const int4 ch = gdp->chunk0;
    const int lm0 = ch.x, nl = ch.y, ob0 = ch.z;
    const int* frames = lay.grp_frames + (long long)gid * (KTF + 1);
    const int* soff = lay.chunk_soff + (long long)gdp->chunk_begin * lay.soff_stride;
    const int NF = lay.N;                      // z of frame N is zero
    const Control* ctl = A.ctl + prob;
    const int cur = ctl->cur;
    const StateBuf<T> bin = cur ? B1 : B0;                    // linearisation point
    const StateBuf<T> bout = A.init ? bin : (cur ? B0 : B1);  // state being linearised
    const bool update = !A.init;
    const bool chol_ok = update ? (ctl->chol_ok != 0) : true;
    const int K6 = 6 * m;                      // rows of the target-frame block B
    const int M6 = K6;
    const int nT = (K6 + 16) >> 4;             // tile rows: K6 block rows and the gradient row
    const int nTiles = nT * (nT + 1) / 2;

    __shared__ T S[KT + 1][27];                 // pair sums per target slot (anchor axes)
    __shared__ T Trel[2][KT + 1][12];           // (G^c_f)^-1 G_a for the group's frames
    __shared__ double Ainv[36];                 // Ad(G_a^-1) of the state being linearised, row-major
    __shared__ double AinvB[36];                // the same at the linearisation point (back-substitution)
    __shared__ double dd[KT + 1][6];            // z_a - z_{f_s} (back-substitution)
    __shared__ T U[CL * US];                    // per (landmark, slot >= 1): J_a^T j_alpha
    __shared__ T iscl[CL];
    __shared__ T Vsh[SH::VROWS * VS];           // B columns in / per-measurement sums / V tile
    // per measurement j_alpha^2 and j_alpha r, written in phase 1 and summed per
    // landmark in phase 2 before the V tile takes the buffer over
    static_assert(SH::VROWS * VS >= 2 * CHUNK_OBS, "the per-measurement sums fit the V buffer");
    T* const ccp = Vsh;
    T* const ggp = Vsh + CHUNK_OBS;
    __shared__ T lmP[CL][6];
    __shared__ int2 lpos[CL];
    __shared__ double red[NWAVE];
    __shared__ int soffL[KT + 2];
    __shared__ int frl[KT + 1];
    __shared__ double alnew[CL];
    __shared__ int lidl[CL];
    __shared__ double Rsh[KT][36];              // R(s) = sum_s' B(s, s') as stored

    double cost = 0.0;
    const T fx0 = (T)A.cam.fx[0], fy0 = (T)A.cam.fy[0], cx0 = (T)A.cam.cx[0], cy0 = (T)A.cam.cy[0];
    const T inv_param = A.rb.param != 0.0 ? (T)(1.0 / A.rb.param) : (T)0;
    const int pair_q = wave_sum27_entry<T>(lane);

    // the measurements of this wavefront's first slot segments and the landmark
    // records, requested up front
    MeasIn pre[PRE_SEG];
#pragma unroll
    for (int r = 0; r < PRE_SEG; ++r) {
        const int seg = wv + r * NWAVE;
        const int sb = seg <= m ? soff[seg] : 0, se = seg <= m ? soff[seg + 1] : 0;
        pre[r] = meas_load(lay, ob0, sb + lane, sb + lane < se);
    }
    int lid = 0;
    float2 lray = make_float2(0.f, 0.f);
    int2 lps = make_int2(0, 0);
    if (tid < nl) {
        lid = lay.lm_id[lm0 + tid];
        lray = lay.lm_ray[lm0 + tid];
        lps = lay.lm_pos[lm0 + tid];
    }
    if (tid < KT + 2) soffL[tid] = soff[tid];
    if (tid < KT + 1) frl[tid] = tid <= m ? frames[tid] : -1;
    for (int i = tid; i < (KT + 1) * 27; i += WG) (&S[0][0])[i] = (T)0;
    // the chunk's stored u columns (nl x 6m, contiguous) for the back-substitution
    const T* vin = bin.v + voff;
    const int nv = nl * M6;
    if (update && chol_ok && nv > 0) {
        constexpr int NV = (CL * 6 * KT + WG - 1) / WG;
        constexpr int NBV = sizeof(T) == 8 ? 4 : 6;
#pragma unroll
        for (int r0 = 0; r0 < NV; r0 += NBV) {
            T vv[NBV];
#pragma unroll
            for (int r = 0; r < NBV; ++r) {
                const int i = tid + (r0 + r) * WG;
                if (r0 + r < NV) vv[r] = vin[min(i, nv - 1)];
            }
#pragma unroll
            for (int r = 0; r < NBV; ++r) {
                const int i = tid + (r0 + r) * WG;
                if (r0 + r < NV && i < nv) Vsh[i] = vv[r];
            }
        }
    }
    for (int i = tid; i < CL * US; i += WG) U[i] = (T)0;
    lds_barrier();

    // ---- set-up: relative poses, world axes, the step at the group's frames ----
    if (tid < 2 * (m + 1)) {
        const int c = tid / (m + 1), s = tid % (m + 1);
        double Ga[12], Gti[12], Tr[12];
        pose_load(bout.G + 12 * a, Ga);
        pose_load((c ? bout.GinvR : bout.GinvL) + 12 * frl[s], Gti);
        pose_mul(Gti, Ga, Tr);
#pragma unroll
        for (int i = 0; i < 12; ++i) Trel[c][s][i] = (T)Tr[i];
    }
    if (tid >= 64 && tid < 136) {
        const int e = (tid - 64) % 36;
        double Gai[12];
        pose_load((tid < 100 ? bout.GinvL : bin.GinvL) + 12 * a, Gai);
        (tid < 100 ? Ainv : AinvB)[e] = adjoint_entry(Gai, e / 6, e % 6);
    }
    lds_barrier();
    if (update && tid >= 128 && tid < 128 + (m + 1)) {
        // z_a - z_{f_s} in the anchor's axes at the linearisation point
        const int s = tid - 128;
        const int f = frl[s];
        double w[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const double za = a < NF ? A.delta[delta_off + 6 * a + q] : 0.0;
            const double zf = f < NF ? A.delta[delta_off + 6 * f + q] : 0.0;
            w[q] = za - zf;
        }
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) t += AinvB[6 * p + q] * w[q];
            dd[s][p] = t;
        }
    }
    lds_barrier();

    // ---- phase 0: back-substitution and the anchor-frame point per landmark -----
    if (tid < nl) {
        const int act = lm0 + tid;
        lpos[tid] = lps;
        double al = bin.alpha[lid];
        if (update) {
            double dal = 0.0;
            if (chol_ok) {
                const T* v = Vsh + tid * M6;
                double s = (double)bin.g[act];
                for (int k = 1; k <= m; ++k)
#pragma unroll
                    for (int q = 0; q < 6; ++q) s -= (double)v[6 * (k - 1) + q] * dd[k][q];
                dal = s / (double)bin.Ct[act];
            }
            al = fmin(fmax(al + dal, A.alpha_min), A.alpha_max);
        }
        alnew[tid] = al;
        lidl[tid] = lid;
        const double x = ((double)lray.x - A.cam.cx[0]) / A.cam.fx[0];
        const double y = ((double)lray.y - A.cam.cy[0]) / A.cam.fy[0];
        const double P0 = x / al, P1 = y / al, P2 = 1.0 / al;
        lmP[tid][0] = (T)P0;
        lmP[tid][1] = (T)P1;
        lmP[tid][2] = (T)P2;
        lmP[tid][3] = (T)(-P0 / al);
        lmP[tid][4] = (T)(-P1 / al);
        lmP[tid][5] = (T)(-P2 / al);
    }
    lds_barrier();

    // ---- phase 1: per measurement, one target-slot segment per wavefront --------
    for (int seg = wv; seg <= m; seg += NWAVE) {
        const int sb = soffL[seg], se = soffL[seg + 1];
        T pr[27];
#pragma unroll
        for (int q = 0; q < 27; ++q) pr[q] = (T)0;
        const int round = (seg - wv) / NWAVE;
        for (int o = sb + lane; o - lane < se; o += 64) {
            const bool valid = o < se;
            MeasIn mi;
            if (o == sb + lane && round < PRE_SEG) {
                mi = pre[0];
#pragma unroll
                for (int r = 1; r < PRE_SEG; ++r)
                    if (round == r) mi = pre[r];
            } else {
                mi = meas_load(lay, ob0, o, valid);
            }
            if (valid) {
                const int t = mi.t;
                const int pos = (int)(mi.meta & 0xffffu);
                const int j = (int)(mi.meta >> 16);
                const float4 uv = mi.uv;
                const T Pa[3] = {lmP[j][0], lmP[j][1], lmP[j][2]};
                const T dv[3] = {lmP[j][3], lmP[j][4], lmP[j][5]};
                const T kappa = (A.rb.near_boost && abs(t - a) < 2) ? (T)16 : (T)1;
                T u[6] = {(T)0, (T)0, (T)0, (T)0, (T)0, (T)0};
                T cc = (T)0, gg = (T)0;
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    T R[12];
                    if (seg > 0 || t == a) {
#pragma unroll
                        for (int i = 0; i < 12; ++i) R[i] = Trel[c][seg][i];
                    } else {
                        // target before the anchor: no pose Jacobian (Solver_cuda.cu:248)
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
                    const T iz = rcp_refined(Pz);
                    const T ex = (c ? (T)uv.z : (T)uv.x) - (fx * Px * iz + cx);
                    const T ey = (c ? (T)uv.w : (T)uv.y) - (fy * Py * iz + cy);
                    T w;
                    robust_eval(A.rb, inv_param, ex * ex + ey * ey, kappa, w, cost);
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
                if (seg > 0) {
                    // two measurements of one landmark in one slot both count: add
#pragma unroll
                    for (int q = 0; q < 6; ++q) atomicAdd(&U[j * US + 6 * (seg - 1) + q], u[q]);
                }
                ccp[pos] = cc;
                ggp[pos] = gg;
            }
        }
        if (seg > 0) {
            const T tot = wave_sum27(pr, lane);
            if (!(lane & 1) && pair_q >= 0) S[seg][pair_q] += tot;
        }
    }
    lds_barrier();

    // ---- phase 2: C and g_alpha per landmark, the rows of V --------------------
    T gl_grad = (T)0;
    if (tid < nl) {
        const int2 lp = lpos[tid];
        T C = (T)0, G = (T)0;
        int p = lp.x + (lp.y > 0 ? tid % lp.y : 0);
        for (int i = 0; i < lp.y; ++i) {
            C += ccp[p];
            G += ggp[p];
            if (++p == lp.x + lp.y) p = lp.x;
        }
        const int act = lm0 + tid;
        const T Ct = C + (T)A.eps;
        const T isc = (T)1 / sqrt(Ct);
        if (update) bout.alpha[lidl[tid]] = alnew[tid];
        bout.Ct[act] = Ct;
        bout.g[act] = G;
        iscl[tid] = isc;
        gl_grad = G * isc;
    }
    lds_barrier();   // the sums are read before the V tile overwrites them
    for (int i = tid; i < SH::VROWS * VS; i += WG) Vsh[i] = (T)0;
    lds_barrier();
    if (tid < nl) Vsh[K6 * VS + tid] = gl_grad;   // the gradient row
    // u in the anchor's axes, stored for the back-substitution, and the rows of
    // V: [u_1; ...; u_m] / sqrt(C + eps)
    for (int it = tid; it < nl * M6; it += WG) {
        const int j = it / M6, k = it - j * M6;
        const T u = U[j * US + k];
        bout.v[voff + it] = u;
        Vsh[k * VS + j] = u * iscl[j];
    }
    lds_barrier();

    // ---- phase 3: V V^T and V g on MFMA, tiles stored as B_a = diag(S) - Y ----
    const int kend = (nl + 3) & ~3;
    T* Da = bout.D + D_off;
    for (int tile = wv; tile < nTiles; tile += NWAVE) {
        int r = 0;
        while ((r + 1) * (r + 2) / 2 <= tile) ++r;
        const int c = tile - r * (r + 1) / 2;
        typename Vec4<T>::type acc = (T)0;
        const T* Ar = Vsh + (r * 16 + (lane & 15)) * VS + (lane >> 4);
        const T* Br = Vsh + (c * 16 + (lane & 15)) * VS + (lane >> 4);
        for (int kk = 0; kk < kend; kk += 4) acc = mfma16x16x4(Ar[kk], Br[kk], acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = r * 16 + mfma16_row((T)0, lane, v);
            const int col = c * 16 + (lane & 15);
            if (row < K6 && col < K6) {
                const int sr = row / 6, sc = col / 6;
                const T q = sr == sc ? S[sr + 1][sym6(row % 6, col % 6)] : (T)0;
                Da[(long long)row * K6 + col] = q - acc[v];
            } else if (row == K6 && col < K6) {
                bout.d[d_off + col] = S[col / 6 + 1][21 + col % 6] - acc[v];
            }
        }
    }
    // to world axes in double (the barrier makes this workgroup's global stores
    // visible to its own loads): the lower blocks of B and the gradient b
    __syncthreads();
    double* Dw = bout.Dw + D_off;
    double* Rg = bout.Rz + r_off;
    const int nlow = m * (m + 1) / 2;
    for (int it = tid; it < 36 * nlow; it += WG) {
        const int blk = it / 36, e = (it / 6) % 6, j = it % 6;
        int sr = 0;
        while ((sr + 1) * (sr + 2) / 2 <= blk) ++sr;
        const int sc = blk - sr * (sr + 1) / 2;
        double acc = 0.0;
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int x = 6 * sr + p, y = 6 * sc + q;
                t += (double)(x >= y ? Da[(long long)x * K6 + y] : Da[(long long)y * K6 + x]) * Ainv[6 * q + j];
            }
            acc += Ainv[6 * p + e] * t;
        }
        Dw[(long long)(6 * sr + e) * K6 + 6 * sc + j] = acc;
    }
    for (int it = tid; it < M6; it += WG) {
        const int s0 = it / 6, e = it % 6;
        double acc = 0.0;
#pragma unroll
        for (int p = 0; p < 6; ++p) acc += Ainv[6 * p + e] * (double)bout.d[d_off + 6 * s0 + p];
        Rg[36 * (m + 1) + 6 + it] = acc;
    }
    __syncthreads();
    // the anchor's blocks: R(s) = sum_s' B(s, s') (upper blocks as transposes),
    // their total and the gradient total
    for (int it = tid; it < 36 * m; it += WG) {
        const int s0 = it / 36, e = (it / 6) % 6, j = it % 6;
        double acc = 0.0;
        for (int s1 = 0; s1 < m; ++s1)
            acc += s1 <= s0 ? Dw[(long long)(6 * s0 + e) * K6 + 6 * s1 + j] : Dw[(long long)(6 * s1 + j) * K6 + 6 * s0 + e];
        Rsh[s0][6 * e + j] = acc;
        Rg[36 * (s0 + 1) + 6 * e + j] = acc;
    }
    if (tid >= 192 && tid < 198) {
        double bt = 0.0;
        for (int s1 = 0; s1 < m; ++s1) bt += Rg[36 * (m + 1) + 6 + 6 * s1 + (tid - 192)];
        Rg[36 * (m + 1) + (tid - 192)] = bt;
    }
    lds_barrier();
    if (tid < 36) {
        double tot = 0.0;
        for (int s1 = 0; s1 < m; ++s1) tot += Rsh[s1][tid];
        Rg[tid] = tot;
    }
    cost = wave_sum(cost);
    if (lane == 0) red[wv] = cost;
    lds_barrier();
    if (tid == 0) {
        double c = 0.0;
        for (int w = 0; w < NWAVE; ++w) c += red[w];
        bout.cost[gid] = c;
    }
    (void)glm0;
}

template <typename T>
void launch_assemble_frames(const AsmArgs& a, const StateBuf<T>* bufs, hipStream_t s) {
    if (a.lay.n_groups == 0) return;
    const dim3 grid(a.lay.n_groups), block(WG);
    if (a.lay.kmax <= 16) hipLaunchKernelGGL((k_assemble_frames<T, 16>), grid, block, 0, s, a, bufs[0], bufs[1]);
    else if (a.lay.kmax <= 32) hipLaunchKernelGGL((k_assemble_frames<T, 32>), grid, block, 0, s, a, bufs[0], bufs[1]);
    else hipLaunchKernelGGL((k_assemble_frames<T, KTF>), grid, block, 0, s, a, bufs[0], bufs[1]);
}

template void launch_assemble_frames<double>(const AsmArgs&, const StateBuf<double>*, hipStream_t);
template void launch_assemble_frames<float>(const AsmArgs&, const StateBuf<float>*, hipStream_t);

int frames_landmarks_per_group(int kt, int fp64) {
    return kt <= 16 ? (fp64 ? FrShape<double, 16>::CL : FrShape<float, 16>::CL) : (kt <= 32 ? 32 : 8);
}

}  // namespace spgo
