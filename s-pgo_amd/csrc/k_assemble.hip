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
// of these blocks over the groups that cover each pose pair (k_form_system).
// Reductions are deterministic: fixed cross-lane exchange patterns and
// wavefront-private LDS partials combined in a fixed order, no global atomics.
#include <type_traits>

#include "spgo_asm_common.h"
#include "spgo_device.h"
#include "spgo_kernels.h"

namespace spgo {

namespace {

constexpr int VS = 66;                 // LDS row stride of the V tile (landmarks + padding)

// LDS shapes for groups whose pose span is at most KT
template <int KT>
struct AsmShape {
    static constexpr int VROWS = 16 * ((6 * KT + 16) / 16);   // B columns + the gradient row, whole tiles
    static constexpr int NTILE = (VROWS / 16) * (VROWS / 16 + 1) / 2;
    static constexpr int MAXTILE = (NTILE + NWAVE - 1) / NWAVE;   // MFMA tiles per wavefront
    static constexpr int A0 = VROWS * VS, A2 = (6 * KT) * (6 * KT + 1);
    static constexpr int VSH = A0 > A2 ? A0 : A2;
};

}  // namespace

// MC: some group has several chunks, so the Schur tiles accumulate across the
// chunk loop in registers; with one chunk per group they live only from phase 3
// on, which keeps them out of the register peak of phase 1
template <typename T, int KT, bool MC>
__global__ __launch_bounds__(WG, (KT <= 12 && sizeof(T) == 4) ? 3 : (KT <= 12 ? 2 : 1)) void k_assemble(AsmArgs A, StateBuf<T> B0, StateBuf<T> B1) {
    using SH = AsmShape<KT>;
    const Layout& lay = A.lay;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;

    // The group record is copied into registers before the kernel's first global
    // store.  A load issued after a store returns only after that store has
    // completed (the vector memory counter retires in order), and the compiler
    // must reload any field it cannot prove unchanged by the stores.  So every
    // global load of this kernel sits before its stores.
    const GroupDesc* gdp = lay.grp_desc + g;
    const int a = gdp->anchor;
    const int K = gdp->K;
    const int prob = gdp->prob, a_loc = gdp->a_loc;
    const int chunk_begin = gdp->chunk_begin, chunk_end = gdp->chunk_end;
    const int glm0 = gdp->lm_begin, d_off = gdp->d_off, gid = gdp->gid;
    const long long voff = gdp->v_off, D_off = gdp->D_off, delta_off = gdp->delta_off;
    const int4 chunk0 = gdp->chunk0;
    const Control* ctl = A.ctl + prob;          // the group's problem (a batch holds several)
    const int cur = ctl->cur;
    // by value: a selected reference to the kernel arguments would put both
    // structures in scratch memory and read every pointer back from there
    const StateBuf<T> bin = cur ? B1 : B0;                    // linearisation point
    const StateBuf<T> bout = A.init ? bin : (cur ? B0 : B1);  // state being linearised
    const bool update = !A.init;
    const bool chol_ok = update ? (ctl->chol_ok != 0) : true;
    const int K6 = 6 * K;
    const int nT = (K6 + 16) >> 4;             // tile rows: B columns and the gradient row K6
    const int nTiles = nT * (nT + 1) / 2;

    __shared__ T S[KT + 1][27];                 // pair sums per slot t - a
    __shared__ T Trel[2][KT + 1][12];           // (G^c_{a+s})^-1 G_a of the state being linearised
    __shared__ T dl[KT][6];                     // Ad(G_a^-1 G_{a+k}) delta_{a+k} (back-substitution)
    // per (landmark, slot s >= 1): u = J_a^T j_alpha of its measurements (zero if none);
    // row stride 6 KT + 1 keeps the landmarks of one wavefront in different banks
    __shared__ T U[CHUNK_LM * (6 * KT + 1)];
    __shared__ T iscl[CHUNK_LM];                // 1 / sqrt(C_l + eps)
    __shared__ T Vsh[SH::VSH];                  // B columns in / per-measurement sums / V tile / D
    // per measurement j_alpha^2 and j_alpha r, written in phase 1 and summed per
    // landmark in phase 2 before the V tile takes the buffer over
    static_assert(SH::VSH >= 2 * CHUNK_OBS, "the per-measurement sums fit the V buffer");
    T* const ccp = Vsh;
    T* const ggp = Vsh + CHUNK_OBS;
    __shared__ T lmP[CHUNK_LM][6];              // P_a and -P_a/alpha in the anchor frame
    __shared__ int2 lpos[CHUNK_LM];
    __shared__ T dly[6 * KT];
    __shared__ double red[NWAVE];
    __shared__ double Adk[KT][36];              // Ad(G_a^-1 G_{a+k}) of the state being linearised
    __shared__ int soffL[KMAX + 2];             // slot offsets of the current chunk
    __shared__ double alnew[CHUNK_LM];          // new inverse depths, stored in phase 2
    __shared__ int lidl[CHUNK_LM];              // their slots

    // ---- group set-up: relative poses, adjoints, the update mapped to the anchor
    // frame
    auto group_setup = [&]() __attribute__((always_inline)) {
        for (int i = tid; i < (KT + 1) * 27; i += WG) (&S[0][0])[i] = (T)0;
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
#pragma unroll
            for (int e = 0; e < 36; ++e) Adk[k][e] = adjoint_entry(E, e / 6, e % 6);
        }
        if (update && tid >= 128 && tid < 128 + K) {
            const int k = tid - 128;
            double Gai[12], Gi[12], E[12], x[6], out[6];
            pose_load(bin.GinvL + 12 * a, Gai);
            pose_load(bin.G + 12 * (a + k), Gi);
            pose_mul(Gai, Gi, E);
#pragma unroll
            for (int q = 0; q < 6; ++q) x[q] = A.delta[delta_off + 6 * (a_loc + k) + q];
            adjoint_apply(E, x, out);
#pragma unroll
            for (int q = 0; q < 6; ++q) dl[k][q] = (T)out[q];
        }
    };

    typename Vec4<T>::type acc[SH::MAXTILE];
    if (MC) {
#pragma unroll
        for (int i = 0; i < SH::MAXTILE; ++i) acc[i] = (T)0;
    }
    double cost = 0.0;
    T gl_grad = (T)0;

    const T fx0 = (T)A.cam.fx[0], fy0 = (T)A.cam.fy[0], cx0 = (T)A.cam.cx[0], cy0 = (T)A.cam.cy[0];
    const T inv_param = A.rb.param != 0.0 ? (T)(1.0 / A.rb.param) : (T)0;

    const int pair_q = wave_sum27_entry<T>(lane);   // the pair-sum entry this lane receives
    group_setup();
    for (int ci = chunk_begin; ci < chunk_end; ++ci) {
        const bool first = ci == chunk_begin;
        const int4 ch = first ? chunk0 : lay.chunk[ci];
        const int lm0 = ch.x, nl = ch.y, ob0 = ch.z;
        const int* soff = first ? gdp->soff0 : lay.chunk_soff + (long long)ci * (KMAX + 2);
        if (tid < KMAX + 2) soffL[tid] = soff[tid];
        // everything phases 0 and 1 read from global memory is requested now:
        // the measurements of this wavefront's first slot segments and the
        // landmark records
        MeasIn pre[PRE_SEG];
#pragma unroll
        for (int r = 0; r < PRE_SEG; ++r) {
            const int seg = wv + r * NWAVE;
            const int sb = seg <= K ? soff[seg] : 0, se = seg <= K ? soff[seg + 1] : 0;
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
        // ---- phase 0: back-substitution and anchor-frame point per landmark ----
        // the chunk's B columns (nl x 6K, contiguous) into LDS with coalesced loads
        const T* vin = bin.v + voff + (long long)(lm0 - glm0) * K6;
        const int nv = nl * K6;
        if (update && chol_ok && nv > 0) {
            // loads in flight in batches of NB before their stores (two round trips
            // for spans up to 8); indices clamped so every load is unconditional
            constexpr int NV = (CHUNK_LM * 6 * KT + WG - 1) / WG;
            constexpr int NBV = sizeof(T) == 8 ? 4 : 6;   // FP64: fewer values in flight, no spills
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
        for (int i = tid; i < CHUNK_LM * (6 * KT + 1); i += WG) U[i] = (T)0;
        lds_barrier();   // also orders the set-up above
        if (tid < nl) {
            const int act = lm0 + tid;
            const int id = lid;
            lpos[tid] = lps;
            double al = bin.alpha[id];
            if (update) {
                double dal = 0.0;
                if (chol_ok) {
                    const T* v = Vsh + tid * K6;
                    double s = (double)bin.g[act];
                    for (int k = 0; k < K; ++k)
#pragma unroll
                        for (int q = 0; q < 6; ++q) s -= (double)v[6 * k + q] * (double)dl[k][q];
                    dal = s / (double)bin.Ct[act];
                }
                al = fmin(fmax(al + dal, A.alpha_min), A.alpha_max);
            }
            // kept in LDS until phase 2 stores it with C and g_alpha
            alnew[tid] = al;
            lidl[tid] = id;
            const float2 ray = lray;
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
        lds_barrier();

        // ---- phase 1: per measurement, one slot segment per wavefront ----------
        // Measurements are in slot-major order, so slot s = t - a (0 for t <= a)
        // is the contiguous segment soff[s] .. soff[s+1]; its pair sums are one
        // reduction over the wavefront (no atomics, fixed order).
        // One measurement: its residual, robust weight and cost, its share of
        // the segment's pair sums and of u, and its j_alpha^2, j_alpha r.  A
        // target before the anchor frame (slot 0 only) takes its relative pose
        // from the global poses, in double; that path runs in a pass of its own
        // so that its registers stay out of the main loop's budget.
        auto measure = [&](const MeasIn& mi, T (&pr)[27], auto before, int seg) __attribute__((always_inline)) {
            constexpr bool BEFORE = decltype(before)::value;
            const int t = mi.t;
            const int pos = (int)(mi.meta & 0xffffu);
            const int j = (int)(mi.meta >> 16);
            const float4 uv = mi.uv;
            const T Pa[3] = {lmP[j][0], lmP[j][1], lmP[j][2]};
            const T dv[3] = {lmP[j][3], lmP[j][4], lmP[j][5]};
            const T kappa = (A.rb.near_boost && abs(t - a) < 2) ? (T)16 : (T)1;
            // The measurement's four weighted rows are J = [P_a x v, v] with v = w n
            // (n: a row of M R_{a->t}).  Its sums are formed from v alone --
            // Mv = sum v v^T, r = sum rk v, q = sum ja v -- and expanded once with
            // A = [P_a]x: J^T J = [[A Mv A^T, A Mv], [., Mv]], J^T rk = [A r, r],
            // J^T ja = [A q, q].
            T M00 = (T)0, M01 = (T)0, M02 = (T)0, M11 = (T)0, M12 = (T)0, M22 = (T)0;
            T r0 = (T)0, r1 = (T)0, r2 = (T)0, q0 = (T)0, q1 = (T)0, q2 = (T)0;
            T cc = (T)0, gg = (T)0;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                T R[12];
                if constexpr (!BEFORE) {
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
                const T iz = rcp_refined(Pz);
                const T ex = (c ? (T)uv.z : (T)uv.x) - (fx * Px * iz + cx);
                const T ey = (c ? (T)uv.w : (T)uv.y) - (fy * Py * iz + cy);
                T w;
                robust_eval(A.rb, inv_param, ex * ex + ey * ey, kappa, w, cost);
                // rows of M R_{a->t}: M = K_c Pi'(P_t) (LMutils.cu:615-628), weighted
                const T wfx = w * fx, wfy = w * fy;
                const T m0a = wfx * iz, m0c = -wfx * Px * iz * iz;
                const T m1b = wfy * iz, m1c = -wfy * Py * iz * iz;
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    T v0, v1, v2;
                    if (k == 0) {
                        v0 = m0a * R[0] + m0c * R[8];
                        v1 = m0a * R[1] + m0c * R[9];
                        v2 = m0a * R[2] + m0c * R[10];
                    } else {
                        v0 = m1b * R[4] + m1c * R[8];
                        v1 = m1b * R[5] + m1c * R[9];
                        v2 = m1b * R[6] + m1c * R[10];
                    }
                    const T ja = v0 * dv[0] + v1 * dv[1] + v2 * dv[2];
                    const T rk = w * (k == 0 ? ex : ey);
                    M00 += v0 * v0;
                    M01 += v0 * v1;
                    M02 += v0 * v2;
                    M11 += v1 * v1;
                    M12 += v1 * v2;
                    M22 += v2 * v2;
                    r0 += rk * v0;
                    r1 += rk * v1;
                    r2 += rk * v2;
                    q0 += ja * v0;
                    q1 += ja * v1;
                    q2 += ja * v2;
                    cc += ja * ja;
                    gg += ja * rk;
                }
            }
            // the expansion with A = [P_a]x (A x = P_a x x, the -[P]x of
            // compute_del_tn_del_xi): B = A Mv, then A Mv A^T = B A^T
            const T a0 = Pa[0], a1 = Pa[1], a2 = Pa[2];
            const T B00 = a1 * M02 - a2 * M01, B01 = a1 * M12 - a2 * M11, B02 = a1 * M22 - a2 * M12;
            const T B10 = a2 * M00 - a0 * M02, B11 = a2 * M01 - a0 * M12, B12 = a2 * M02 - a0 * M22;
            const T B20 = a0 * M01 - a1 * M00, B21 = a0 * M11 - a1 * M01, B22 = a0 * M12 - a1 * M02;
            const T JJ[21] = {a1 * B02 - a2 * B01, a2 * B00 - a0 * B02, a0 * B01 - a1 * B00, B00, B01, B02,
                              a2 * B10 - a0 * B12, a0 * B11 - a1 * B10, B10, B11, B12,
                              a0 * B21 - a1 * B20, B20, B21, B22,
                              M00, M01, M02, M11, M12, M22};
#pragma unroll
            for (int i = 0; i < 21; ++i) pr[i] += JJ[i];
            pr[21] += a1 * r2 - a2 * r1;
            pr[22] += a2 * r0 - a0 * r2;
            pr[23] += a0 * r1 - a1 * r0;
            pr[24] += r0;
            pr[25] += r1;
            pr[26] += r2;
            const T u[6] = {a1 * q2 - a2 * q1, a2 * q0 - a0 * q2, a0 * q1 - a1 * q0, q0, q1, q2};
            if (seg > 0) {
                // a cell written twice is two measurements of one landmark in one
                // slot (the reference counts both): add, do not overwrite.  The
                // slot starts at zero, so one or two contributions give the
                // same sum in either order.
#pragma unroll
                for (int q = 0; q < 6; ++q) atomicAdd(&U[j * (6 * KT + 1) + 6 * (seg - 1) + q], u[q]);
            }
            ccp[pos] = cc;
            ggp[pos] = gg;
        };
        for (int seg = wv; seg <= K; seg += NWAVE) {
            const int sb = soffL[seg], se = soffL[seg + 1];
            T pr[27];
#pragma unroll
            for (int q = 0; q < 27; ++q) pr[q] = (T)0;
            const int round = (seg - wv) / NWAVE;
            bool before = false;
            for (int o = sb + lane; o - lane < se; o += 64) {
                const bool valid = o < se;
                MeasIn mi;
                if (o == sb + lane && round < PRE_SEG) {
                    // a register array indexed by a wavefront-uniform value: select, no scratch
                    mi = pre[0];
#pragma unroll
                    for (int r = 1; r < PRE_SEG; ++r)
                        if (round == r) mi = pre[r];
                } else {
                    mi = meas_load(lay, ob0, o, valid);
                }
                if (valid) {
                    if (mi.t >= a) measure(mi, pr, std::false_type{}, seg);
                    else before = true;
                }
            }
            if (seg == 0 && __builtin_amdgcn_ballot_w64(before) != 0) {
                for (int o = sb + lane; o - lane < se; o += 64) {
                    const bool valid = o < se;
                    const MeasIn mi = meas_load(lay, ob0, o, valid);
                    if (valid && mi.t < a) measure(mi, pr, std::true_type{}, seg);
                }
            }
            if (seg > 0) {
                // pair sums of the segment, reduced over the wavefront in registers
                const T tot = wave_sum27(pr, lane);
                if (!(lane & 1) && pair_q >= 0) S[seg][pair_q] += tot;
            }
        }
        lds_barrier();

        // ---- phase 2: per landmark C, g_alpha and its B column v_l ---------------
        if (tid < nl) {
            // sums over the landmark's measurements, starting at a landmark-dependent
            // position so the lanes of a wavefront read different banks
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
        if (tid < nl) Vsh[K6 * VS + tid] = gl_grad;   // the gradient row: y comes out of the same products
        // v_l(k) = sum of u over the slots s > k: a suffix scan per (landmark, component)
        for (int it = tid; it < nl * 6; it += WG) {
            const int j = it / 6, q = it - j * 6;
            const T isc = iscl[j];
            T* vo = bout.v + voff + (long long)(lm0 + j - glm0) * K6 + q;
            const T* uj = U + j * (6 * KT + 1) + q;
            T run = (T)0;
            for (int k = K - 1; k >= 0; --k) {
                run += uj[6 * k];
                vo[6 * k] = run;
                Vsh[(6 * k + q) * VS + j] = run * isc;
            }
        }
        lds_barrier();

        // ---- phase 3: [V; g]^T-products on MFMA: Y = V V^T and y = V g -----------
        if (!MC) {
#pragma unroll
            for (int i = 0; i < SH::MAXTILE; ++i) acc[i] = (T)0;
        }
        const int kend = (nl + 3) & ~3;
#pragma unroll
        for (int ti = 0; ti < SH::MAXTILE; ++ti) {
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
        lds_barrier();
    }

    // ---- finalize: D = Q - Y, d = q - y ----------------------------------------
    // suffix sums over slots: Q(i, j) = sum_{s > max(i, j)} S(s)
    T* Rsuf = U;                       // [K][27]
    for (int it = tid; it < K * 27; it += WG) {
        const int s0 = it / 27 + 1, q = it - (s0 - 1) * 27;
        T run = (T)0;
        for (int s = K; s >= s0; --s) run += S[s][q];
        Rsuf[(s0 - 1) * 27 + q] = run;
    }
    lds_barrier();
    // K6 x K6, full, rows K6 + 1 apart: an odd stride puts the 16 rows of a
    // transposed tile store (one per lane) in different banks
    T* Dl = Vsh;
    const int DS = K6 + 1;
#pragma unroll
    for (int ti = 0; ti < SH::MAXTILE; ++ti) {
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
                    const T val = Rsuf[kb * 27 + sym6(row % 6, col % 6)] - acc[ti][v];
                    Dl[row * DS + col] = val;
                    Dl[col * DS + row] = val;
                } else if (row == K6 && col < K6) {
                    dly[col] = Rsuf[(col / 6) * 27 + 21 + (col % 6)] - acc[ti][v];
                }
            }
        }
    }
    lds_barrier();
    // to incremental-pose coordinates, straight to the reduced blocks:
    // D(i, j) <- Ad_i^T D(i, j) Ad_j with Ad_k = Ad(G_a^-1 G_{a+k}), for the
    // lower blocks i >= j (the reduced system reads only those).  One thread
    // forms row r of one block: t = (column r of Ad_i)^T D(i, j), then t Ad_j
    // (6 K (K + 1) / 2 rows: more than one per thread for spans past 10).
    T* Dg = bout.D + D_off;
    const int nlow = K * (K + 1) / 2;
    for (int it = tid; it < 6 * nlow; it += WG) {
        const int blk = it / 6, r = it - 6 * blk;
        int bi = 0;
        while ((bi + 1) * (bi + 2) / 2 <= blk) ++bi;
        const int bj = blk - bi * (bi + 1) / 2;
        const T* Db = Dl + 6 * bi * DS + 6 * bj;
        double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int p = 0; p < 6; ++p) {
            const double a = Adk[bi][6 * p + r];
#pragma unroll
            for (int q = 0; q < 6; ++q) t[q] += a * (double)Db[p * DS + q];
        }
        T* out = Dg + (6 * bi + r) * K6 + 6 * bj;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) s += t[q] * Adk[bj][6 * q + c];
            out[c] = (T)s;
        }
    }
    if (tid < K6) {
        const int k = tid / 6, r = tid % 6;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) s += Adk[k][6 * q + r] * (double)dly[6 * k + q];
        bout.d[d_off + tid] = (T)s;
    }
    cost = wave_sum(cost);
    if (lane == 0) red[wv] = cost;
    lds_barrier();
    if (tid == 0) {
        double c = 0.0;
        for (int w = 0; w < NWAVE; ++w) c += red[w];
        bout.cost[gid] = c;
    }
}

template <typename T>
void launch_assemble(const AsmArgs& a, const StateBuf<T>* bufs, hipStream_t s) {
    if (a.lay.frame_mode) {
        launch_assemble_frames<T>(a, bufs, s);
        return;
    }
    if (a.lay.n_groups == 0) return;
    const dim3 grid(a.lay.n_groups), block(WG);
    if (a.lay.kmax <= 8) {
        if (a.lay.multi_chunk) hipLaunchKernelGGL((k_assemble<T, 8, true>), grid, block, 0, s, a, bufs[0], bufs[1]);
        else hipLaunchKernelGGL((k_assemble<T, 8, false>), grid, block, 0, s, a, bufs[0], bufs[1]);
    } else if (a.lay.kmax <= 12) {
        if (a.lay.multi_chunk) hipLaunchKernelGGL((k_assemble<T, 12, true>), grid, block, 0, s, a, bufs[0], bufs[1]);
        else hipLaunchKernelGGL((k_assemble<T, 12, false>), grid, block, 0, s, a, bufs[0], bufs[1]);
    } else {
        if (a.lay.multi_chunk) hipLaunchKernelGGL((k_assemble<T, KMAX, true>), grid, block, 0, s, a, bufs[0], bufs[1]);
        else hipLaunchKernelGGL((k_assemble<T, KMAX, false>), grid, block, 0, s, a, bufs[0], bufs[1]);
    }
}

// Workgroups of the assembly kernel for groups of span kmax resident on the
// current device at once (per-CU occupancy x CUs); 0 if the device cannot say.
int assemble_resident_blocks(int kmax, int fp32, int multi_chunk) {
    const void* f;
    if (fp32) {
        if (kmax <= 8) f = multi_chunk ? (const void*)k_assemble<float, 8, true> : (const void*)k_assemble<float, 8, false>;
        else if (kmax <= 12)
            f = multi_chunk ? (const void*)k_assemble<float, 12, true> : (const void*)k_assemble<float, 12, false>;
        else f = multi_chunk ? (const void*)k_assemble<float, KMAX, true> : (const void*)k_assemble<float, KMAX, false>;
    } else {
        if (kmax <= 8) f = multi_chunk ? (const void*)k_assemble<double, 8, true> : (const void*)k_assemble<double, 8, false>;
        else if (kmax <= 12)
            f = multi_chunk ? (const void*)k_assemble<double, 12, true> : (const void*)k_assemble<double, 12, false>;
        else f = multi_chunk ? (const void*)k_assemble<double, KMAX, true> : (const void*)k_assemble<double, KMAX, false>;
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, WG, 0) != hipSuccess) return 0;
    return per_cu * cus;
}

template void launch_assemble<double>(const AsmArgs&, const StateBuf<double>*, hipStream_t);
template void launch_assemble<float>(const AsmArgs&, const StateBuf<float>*, hipStream_t);

}  // namespace spgo
