// State update, global pose chain and the LM accept/reject decision.
//
// k_pose_update: X_k <- X_k Exp(s delta_k) (Solver_cuda.cu:511-536) and the
// global chain G_0 = I, G_{k+1} = G_k X_k^-1, G^R = G T_r_to_l
// (Solver_cuda.cu:440-493), computed as a parallel prefix product instead of
// the reference's single-thread loop: every thread multiplies its chunk of
// poses, the chunk products are scanned across each wavefront by shuffles and
// across the wavefronts in LDS, and each thread re-walks its chunk from its
// exclusive prefix (with one pose per thread the prefix is the result).
// k_accept: robust cost of the candidate (sum of the group partials in a fixed
// order), accept if it decreased, lambda update, per-iteration record.
#include "spgo_device.h"
#include "spgo_kernels.h"

namespace spgo {

template <typename T>
__global__ __launch_bounds__(WG) void k_pose_update(Layout lay, Camera cam, StateBuf<T> B0, StateBuf<T> B1,
                                                    const Control* ctls, const double* delta, double step,
                                                    int init) {
    // one workgroup per problem: its N poses start at pose_base, its frames at frame_base
    __shared__ double wtot[WG / 64][12];   // per wavefront: the product of its chunks
    const int tid = threadIdx.x;
    const ProbDesc pd = lay.probs[blockIdx.x];
    const Control* ctl = ctls + blockIdx.x;
    const int N = pd.N;
    const double* dlt = delta + pd.vec_off;
    // one pose per thread (N <= WG): its pose in both buffers and its step are
    // requested before the control block that selects between them arrives
    // In frame coordinates the solution is z and delta_k = Ad(G_k^-1) (z_k - z_{k+1})
    // (z_N = 0) at the linearisation point.
    const bool frame = lay.frame_mode != 0;
    double Xa[12], Xb[12], dpre[6];
    const bool one = N <= WG && tid < N;
    if (one) {
        pose_load(B0.X + 12 * (pd.pose_base + tid), Xa);
        pose_load(B1.X + 12 * (pd.pose_base + tid), Xb);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            dpre[q] = init ? 0.0 : dlt[6 * tid + q] - ((frame && tid + 1 < N) ? dlt[6 * (tid + 1) + q] : 0.0);
    }
    const int cur = ctl->cur;
    const StateBuf<T> bin = cur ? B1 : B0;
    const StateBuf<T> bout = init ? bin : (cur ? B0 : B1);
    const bool ok = init ? false : (ctl->chol_ok != 0);
    if (one && frame && !init) {
        double Gi[12], x[6];
        pose_load(bin.GinvL + 12 * (pd.frame_base + tid), Gi);
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
            pose_load(bin.X + 12 * (pd.pose_base + k), X);
        }
        if (!init) {
            double xi[6], E[12], dk[6];
            if (one) {
#pragma unroll
                for (int q = 0; q < 6; ++q) dk[q] = dpre[q];
            } else if (frame) {
                double Gi[12], x[6];
                pose_load(bin.GinvL + 12 * (pd.frame_base + k), Gi);
#pragma unroll
                for (int q = 0; q < 6; ++q) x[q] = dlt[6 * k + q] - (k + 1 < N ? dlt[6 * (k + 1) + q] : 0.0);
                adjoint_apply(Gi, x, dk);
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) dk[q] = dlt[6 * k + q];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) xi[q] = ok ? step * dk[q] : 0.0;
            exp_se3(xi, E);
            pose_mul(X, E, tmp);
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                X[q] = tmp[q];
                bout.X[12 * (pd.pose_base + k) + q] = tmp[q];
            }
        }
        pose_inv(X, Xi);
        pose_mul(P, Xi, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) P[q] = tmp[q];
    }
    // inclusive scan of the chunk products over the workgroup: a Kogge-Stone
    // scan across each wavefront's lanes by shuffles (no LDS, no barriers), then
    // the wavefront totals, combined in LDS.  Threads past the last chunk hold
    // the identity.
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        double Q[12], tmp[12];
#pragma unroll
        for (int q = 0; q < 12; ++q) Q[q] = __shfl_up(P[q], d, 64);
        if (lane >= d) {
            pose_mul(Q, P, tmp);
#pragma unroll
            for (int q = 0; q < 12; ++q) P[q] = tmp[q];
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int q = 0; q < 12; ++q) wtot[wv][q] = P[q];
    }
    __syncthreads();
    // product of the totals of the wavefronts before this one
    double W[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (int w = 0; w < wv; ++w) {
        double tmp[12];
        pose_mul(W, wtot[w], tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) W[q] = tmp[q];
    }
    if (wv > 0) {
        double tmp[12];
        pose_mul(W, P, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) P[q] = tmp[q];
    }
    // exclusive prefix = inclusive scan of the previous thread
    double G[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const double up = __shfl_up(P[q], 1, 64);
        G[q] = lane > 0 ? up : W[q];
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
            bout.G[o + q] = Gf[q];
            bout.GinvL[o + q] = Gi[q];
            bout.GinvR[o + q] = GRi[q];
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
        pose_load(bout.X + 12 * (pd.pose_base + k), X);   // this thread's own stores: program order
        pose_inv(X, Xi);
        pose_mul(G, Xi, tmp);
#pragma unroll
        for (int q = 0; q < 12; ++q) G[q] = tmp[q];
        emit(k + 1, G);
    }
}

template <typename T>
void launch_pose_update(const Layout& lay, const Camera& cam, StateBuf<T>* bufs, const Control* ctl,
                        const double* delta, double step_size, int init, hipStream_t s) {
    hipLaunchKernelGGL(k_pose_update<T>, dim3(lay.n_probs), dim3(WG), 0, s, lay, cam, bufs[0], bufs[1], ctl,
                       delta, step_size, init);
}
template void launch_pose_update<double>(const Layout&, const Camera&, StateBuf<double>*, const Control*,
                                         const double*, double, int, hipStream_t);
template void launch_pose_update<float>(const Layout&, const Camera&, StateBuf<float>*, const Control*,
                                        const double*, double, int, hipStream_t);

__global__ __launch_bounds__(WG) void k_accept(const double* c0, const double* c1, const ProbDesc* probs,
                                              Control* ctls, IterRecord* recs, int reference_mode, double lam_up,
                                              double lam_down, double lam_min, double lam_max, int init) {
    // one workgroup per problem: the costs of its groups, its control block and records
    __shared__ double red[WG];
    const int tid = threadIdx.x;
    const ProbDesc pd = probs[blockIdx.x];
    Control* ctl = ctls + blockIdx.x;
    IterRecord* rec = recs + pd.rec_off;
    const int cur = ctl->cur;
    const int out = init ? cur : cur ^ 1;
    const double* cp = out ? c1 : c0;
    double s = 0.0;
    for (int g = pd.g_begin + tid; g < pd.g_end; g += WG) s += cp[g];
    red[tid] = s;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid != 0) return;
    const double cand = red[0];
    if (init) {
        ctl->cost = cand;
        ctl->cost0 = cand;
        ctl->iter = 0;
        return;
    }
    const int ok = ctl->chol_ok;
    const double lam_used = ctl->lambda;
    int accept;
    if (reference_mode) {
        accept = 1;
    } else {
        accept = (cand < ctl->cost) ? 1 : 0;
        double l = accept ? lam_used * lam_down : lam_used * lam_up;
        ctl->lambda = fmin(fmax(l, lam_min), lam_max);
    }
    if (accept) {
        ctl->cur = out;
        ctl->cost = cand;
    }
    IterRecord r;
    r.cost_candidate = cand;
    r.cost = ctl->cost;
    r.lambda = lam_used;
    r.accepted = accept;
    r.chol_ok = ok;
    r.pcg_iters = ctl->pcg_iters;
    r.pcg_fail = ctl->pcg_fail;
    rec[ctl->iter] = r;
    ctl->iter = ctl->iter + 1;
}

void launch_accept(const double* cost_buf0, const double* cost_buf1, const Layout& lay, Control* ctl,
                   IterRecord* rec, int reference_mode, double lam_up, double lam_down, double lam_min, double lam_max,
                   int init, hipStream_t s) {
    hipLaunchKernelGGL(k_accept, dim3(lay.n_probs), dim3(WG), 0, s, cost_buf0, cost_buf1, lay.probs, ctl, rec,
                       reference_mode, lam_up, lam_down, lam_min, lam_max, init);
}

// Start of a step call: restart the iteration records, mark the factorisation
// usable, optionally reset the damping, and (when the last linearisation is
// kept) take its cost as the step's initial cost.  One problem's control block.
__global__ void k_begin_step(Control* ctl, int reset_lambda, double lambda0, int keep_lin) {
    if (threadIdx.x != 0) return;
    if (reset_lambda) ctl->lambda = lambda0;
    ctl->iter = 0;
    ctl->chol_ok = 1;
    if (keep_lin) ctl->cost0 = ctl->cost;
}

void launch_begin_step(Control* ctl, int reset_lambda, double lambda0, int keep_lin, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(64), 0, s, ctl, reset_lambda, lambda0, keep_lin);
}

__global__ void k_fill(double* p, double v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_set_diag(double* H, int Pp, int from, double v) {
    const int i = from + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Pp) H[(long long)i * Pp + i] = v;
}

void launch_fill(double* p, double v, long long n, hipStream_t s) {
    if (n <= 0) return;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

void launch_set_diag(double* H, int Pp, int from, double v, hipStream_t s) {
    if (from >= Pp) return;
    hipLaunchKernelGGL(k_set_diag, dim3((Pp - from + 255) / 256), dim3(256), 0, s, H, Pp, from, v);
}

}  // namespace spgo
