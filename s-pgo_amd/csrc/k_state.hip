// State update, global pose chain and the LM accept/reject decision.
//
// k_pose_update: X_k <- X_k Exp(s delta_k) and the global pose chain, one
// workgroup per problem (spgo_pose.h).
// k_accept: robust cost of the candidate (sum of the group partials in a fixed
// order), accept if it decreased, lambda update, per-iteration record.
#include "spgo_device.h"
#include "spgo_kernels.h"
#include "spgo_pose.h"

namespace spgo {

template <typename T>
__global__ __launch_bounds__(WG) void k_pose_update(Layout lay, Camera cam, StateBuf<T> B0, StateBuf<T> B1,
                                                    const Control* ctls, const double* delta, double step,
                                                    int init) {
    // one workgroup per problem: its N poses start at pose_base, its frames at frame_base
    pose_update_block<false>(lay.probs[blockIdx.x], lay.frame_mode, cam, pose_bufs(B0, B1), ctls + blockIdx.x,
                             delta, step, init);
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
