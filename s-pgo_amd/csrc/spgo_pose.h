// The pose update of one problem by one workgroup of WG threads, shared by
// k_pose_update and the cyclic reduction's last back-substitution workgroup:
// X_k <- X_k Exp(s delta_k) (Solver_cuda.cu:511-536) and the global chain
// G_0 = I, G_{k+1} = G_k X_k^-1, G^R = G T_r_to_l (Solver_cuda.cu:440-493),
// computed as a parallel prefix product instead of the reference's
// single-thread loop: every thread multiplies its chunk of poses, the chunk
// products are scanned across each wavefront by DPP lane moves and across the
// wavefronts in LDS, and each thread re-walks its chunk from its exclusive
// prefix (with one pose per thread the prefix is the result).
// COHERENT: the step delta is read with device-coherent loads (it was written
// by other workgroups of the calling kernel).
#pragma once
#include "spgo_device.h"

namespace spgo {

// A pose (12 doubles) from another lane by DPP: CTRL selects the source lane,
// ROW_MASK the 16-lane rows written; the other lanes keep their own value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void pose_dpp(const double* P, double* Q) {
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const long long b = __double_as_longlong(P[q]);
        const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, CTRL, ROW_MASK, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
        Q[q] = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    }
}
// P <- Q P where take holds
__device__ __forceinline__ void pose_lmul_if(bool take, const double* Q, double* P) {
    if (take) {
        double t[12];
        pose_mul(Q, P, t);
#pragma unroll
        for (int q = 0; q < 12; ++q) P[q] = t[q];
    }
}

template <bool COHERENT>
__device__ __forceinline__ void pose_update_block(const ProbDesc pd, int frame_mode, const Camera& cam,
                                                  const PoseBufs& B, const Control* ctl, const double* delta,
                                                  double step, int init) {
    __shared__ double wtot[WG / 64][12];   // per wavefront: the product of its chunks
    const int tid = threadIdx.x;
    const int N = pd.N;
    const double* dlt = delta + pd.vec_off;
    auto dl = [&](int i) { return COHERENT ? __hip_atomic_load(dlt + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : dlt[i]; };
    // one pose per thread (N <= WG): its pose in both buffers and its step are
    // requested before the control block that selects between them arrives
    // In frame coordinates the solution is z and delta_k = Ad(G_k^-1) (z_k - z_{k+1})
    // (z_N = 0) at the linearisation point.
    const bool frame = frame_mode != 0;
    double Xa[12], Xb[12], dpre[6];
    const bool one = N <= WG && tid < N;
    if (one) {
        pose_load(B.X[0] + 12 * (pd.pose_base + tid), Xa);
        pose_load(B.X[1] + 12 * (pd.pose_base + tid), Xb);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            dpre[q] = init ? 0.0 : dl(6 * tid + q) - ((frame && tid + 1 < N) ? dl(6 * (tid + 1) + q) : 0.0);
    }
    const int cur = ctl->cur;
    const int bi = cur, bo = init ? cur : cur ^ 1;   // linearisation point, state written
    // the selected buffers' arrays, chosen by value (an array indexed at run
    // time would leave the pointers without their global address space and
    // turn every access into a flat one)
    const double* const Xin = bi ? B.X[1] : B.X[0];
    const double* const GiLin = bi ? B.GinvL[1] : B.GinvL[0];
    double* const Xout = bo ? B.X[1] : B.X[0];
    double* const Gout = bo ? B.G[1] : B.G[0];
    double* const GiLout = bo ? B.GinvL[1] : B.GinvL[0];
    double* const GiRout = bo ? B.GinvR[1] : B.GinvR[0];
    const int okv = COHERENT ? __hip_atomic_load(&ctl->chol_ok, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ctl->chol_ok;
    const bool ok = init ? false : (okv != 0);
    if (one && frame && !init) {
        double Gi[12], x[6];
        pose_load(GiLin + 12 * (pd.frame_base + tid), Gi);
#pragma unroll
        for (int q = 0; q < 6; ++q) x[q] = dpre[q];
        adjoint_apply(Gi, x, dpre);
    }

    const int chunk = (N + WG - 1) / WG;
    const int nact = (N + chunk - 1) / chunk;   // threads holding poses
    const int k0 = min(N, tid * chunk), k1 = min(N, k0 + chunk);
    double P[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (int k = k0; k < k1; ++k) {
        double X[12], Xi[12], tmp[12];
        if (one) {
#pragma unroll
            for (int q = 0; q < 12; ++q) X[q] = cur ? Xb[q] : Xa[q];
        } else {
            pose_load(Xin + 12 * (pd.pose_base + k), X);
        }
        if (!init) {
            double xi[6], E[12], dk[6];
            if (one) {
#pragma unroll
                for (int q = 0; q < 6; ++q) dk[q] = dpre[q];
            } else if (frame) {
                double Gi[12], x[6];
                pose_load(GiLin + 12 * (pd.frame_base + k), Gi);
#pragma unroll
                for (int q = 0; q < 6; ++q) x[q] = dl(6 * k + q) - (k + 1 < N ? dl(6 * (k + 1) + q) : 0.0);
                adjoint_apply(Gi, x, dk);
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) dk[q] = dl(6 * k + q);
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) xi[q] = ok ? step * dk[q] : 0.0;
            exp_se3(xi, E);
            pose_mul(X, E, tmp);
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                X[q] = tmp[q];
                Xout[12 * (pd.pose_base + k) + q] = tmp[q];
            }
        }
        pose_inv(X, Xi);
        pose_mul(P, Xi, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) P[q] = tmp[q];
    }
    // inclusive scan of the chunk products over the workgroup: a scan across
    // each wavefront's lanes by DPP lane moves (no LDS, no barriers), then the
    // wavefront totals, combined in LDS.  Threads past the last chunk hold the
    // identity.  Only the wavefronts holding chunks scan, and only the steps
    // that reach their last chunk.
    const int lane = tid & 63, wv = tid >> 6;
    const int nw = (nact + 63) >> 6;   // wavefronts holding chunks
    if (wv < nw) {
        // within each 16-lane row (row_shr), then lane 15 of a row into the
        // next row (row_bcast:15, rows 1 and 3) and lane 31 into rows 2 and 3
        // (row_bcast:31)
        double Q[12];
        const int r16 = lane & 15;
        if (nact > 1) {
            pose_dpp<0x111, 0xF>(P, Q);
            pose_lmul_if(r16 >= 1, Q, P);
        }
        if (nact > 2) {
            pose_dpp<0x112, 0xF>(P, Q);
            pose_lmul_if(r16 >= 2, Q, P);
        }
        if (nact > 4) {
            pose_dpp<0x114, 0xF>(P, Q);
            pose_lmul_if(r16 >= 4, Q, P);
        }
        if (nact > 8) {
            pose_dpp<0x118, 0xF>(P, Q);
            pose_lmul_if(r16 >= 8, Q, P);
        }
        if (nact > 16) {
            pose_dpp<0x142, 0xA>(P, Q);
            pose_lmul_if((lane >> 4) & 1, Q, P);
        }
        if (nact > 32) {
            pose_dpp<0x143, 0xC>(P, Q);
            pose_lmul_if(lane >= 32, Q, P);
        }
        if (lane == 63 && nw > 1) {
#pragma unroll
            for (int q = 0; q < 12; ++q) wtot[wv][q] = P[q];
        }
    }
    if (nw > 1) __syncthreads();
    // product of the totals of the wavefronts before this one
    double W[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (int w = 0; w < wv && wv < nw; ++w) {
        double tmp[12];
        pose_mul(W, wtot[w], tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) W[q] = tmp[q];
    }
    if (wv > 0 && wv < nw) {
        double tmp[12];
        pose_mul(W, P, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) P[q] = tmp[q];
    }
    // exclusive prefix = inclusive scan of the previous thread (wave_shr:1)
    double G[12];
    pose_dpp<0x138, 0xF>(P, G);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 12; ++q) G[q] = W[q];
    }
    double Trl[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) Trl[q] = cam.Trl[q];
    auto emit = [&](int f, const double* Gf) {
        double Gi[12], GR[12], GRi[12];
        pose_inv(Gf, Gi);
        pose_mul(Gf, Trl, GR);
        pose_inv(GR, GRi);
        const long long o = 12LL * (pd.frame_base + f);
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            Gout[o + q] = Gf[q];
            GiLout[o + q] = Gi[q];
            GiRout[o + q] = GRi[q];
        }
    };
    if (tid == 0) emit(0, G);
    if (chunk == 1) {
        // one pose per thread: G_{k+1} is this thread's inclusive prefix
        if (k0 < k1) emit(k0 + 1, P);
        return;
    }
    for (int k = k0; k < k1; ++k) {
        double X[12], Xi[12], tmp[12];
        pose_load(Xout + 12 * (pd.pose_base + k), X);   // this thread's own stores: program order
        pose_inv(X, Xi);
        pose_mul(G, Xi, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) G[q] = tmp[q];
        emit(k + 1, G);
    }
}

}  // namespace spgo
