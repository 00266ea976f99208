// Learned-preconditioner PCG: sizes, the device view of a set of block-banded
// systems, and the kernel launchers (k_lpcg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "spgo_internal.h"

struct spgo_lpcg;

namespace spgo {

constexpr int LP_HID = 64;     // hidden width
constexpr int LP_NIN = 40;     // node input: normalised (i, i-1) block, 2 norms, degree, 1
constexpr int LP_NOUT = 57;    // 21 lower-triangle corrections of G_ii, 36 entries of G_{i,i-1}
constexpr int LP_NLAYER = 3;   // message-passing layers
constexpr int LP_WMAX = 16;    // widest band (blocks below the diagonal)

// Parameter layout (FP32, flat): W_in (HID x NIN), b_in, then per layer W_s,
// W_m (HID x HID) and b, then W_o (NOUT x HID), W_k (NOUT x NIN), b_o.
constexpr int LP_OFF_WIN = 0;
constexpr int LP_OFF_BIN = LP_OFF_WIN + LP_HID * LP_NIN;
constexpr int LP_OFF_L0 = LP_OFF_BIN + LP_HID;
constexpr int LP_LAYER_SIZE = 2 * LP_HID * LP_HID + LP_HID;
constexpr int LP_OFF_WO = LP_OFF_L0 + LP_NLAYER * LP_LAYER_SIZE;
constexpr int LP_OFF_WK = LP_OFF_WO + LP_NOUT * LP_HID;
constexpr int LP_OFF_BO = LP_OFF_WK + LP_NOUT * LP_NIN;
constexpr int LP_NPARAM = LP_OFF_BO + LP_NOUT;

// Systems on the device: T nodes (6x6 blocks) of S systems, band W.
struct LpSet {
    int T, S, W;
    const double* Hb;       // T x (W + 1) x 36: Hb[i][d] = H_{i,i-d}
    const double* g;        // T x 6
    const int* node_sys;    // T
    const int* sys_begin;   // S + 1
};

void lp_launch_features(const LpSet& s, double* C, double* sfeat, float* X, double* N1, double* wsum, int* bad,
                        hipStream_t st);
void lp_launch_linear(int T, int O, int I1, const float* in1, const float* W1, int I2, const float* in2,
                      const float* W2, const float* b, float* z, float* h, hipStream_t st);
void lp_launch_linearT(int T, int O, int I, const float* dy, const float* Wt, float* dx, int acc, hipStream_t st);
void lp_launch_agg(const LpSet& s, const double* sfeat, const double* wsum, const float* hin, float* m, int transpose,
                   hipStream_t st);
void lp_launch_reluback(int n, const float* dh, const float* z, float* dz, hipStream_t st);
// gW = dy^T x (O x I), gb = column sums of dy (optional); part holds
// lp_wgrad_blocks(T) x (O I + O) floats of partial sums
void lp_launch_wgrad(int T, int O, int I, const float* dy, const float* x, float* gW, float* gb, float* part,
                     hipStream_t st);
int lp_wgrad_blocks(int T);
void lp_launch_buildG(const LpSet& s, const float* out, const double* C, double* Gd, double* Go, hipStream_t st);
// Kaporin loss: per-node terms, per-system T and loss; gradient wrt the head output
void lp_launch_kap(const LpSet& s, const float* out, const double* N1, double* tnode, double* ldnode, double* Tsys,
                   double* loss, hipStream_t st);
void lp_launch_kap_dout(const LpSet& s, const float* out, const double* N1, const double* Tsys, float* dout,
                        hipStream_t st);
void lp_launch_adam(int n, float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                    int step, hipStream_t st);
// PCG from x = 0 on every system; with ctl, ctl[s].pcg_iters / pcg_fail are set;
// with bad (device flag of a failed diagonal factor), nothing is solved when set
void lp_launch_pcg(const LpSet& s, const double* Gd, const double* Go, double* x, double* r, double* z, double* p,
                   double* q, double* tmp, double tol, int maxit, int* iters, double* relres, hipStream_t st,
                   Control* ctl = nullptr, const int* bad = nullptr);
void lp_launch_band_from_dense(int N, int W, const double* H, int Pp, double* Hb, hipStream_t st);
void lp_launch_zero_flag(int* p, hipStream_t st);   // a kernel, so a captured graph clears it too
// PCG on one system with a cooperative launch of `blocks` workgroups (at most
// lp_coop_max_blocks(), all resident); part holds 4 x blocks doubles.  Returns
// the launch status.  Sets ctl->pcg_iters and ctl->pcg_fail.
int lp_coop_max_blocks();
// Blocks of H outside the band (loop revisits in frame coordinates) join the
// product H v: far_beg (T + 1) ranges per block row into far_col / far_val
// (36 values each, H_{i,j} row-major, both triangles listed); far_beg null: none.
int lp_launch_pcg_coop(int T, int W, const double* Hb, const double* Gd, const double* Go, const double* g, double* x,
                       double* r, double* z, double* p, double* q, double* tmp, double* part, int blocks, double tol,
                       int maxit, Control* ctl, const int* bad, const int* far_beg, const int* far_col,
                       const double* far_val, hipStream_t st);
// far_val[k] = H_{far_row[k], far_col[k]} from the lower triangle of a dense H
void lp_launch_far_from_dense(int nfar, const int* far_row, const int* far_col, const double* H, int Pp,
                              double* far_val, hipStream_t st);

// per-node buffers of one network forward pass
struct LpNet {
    float* X;
    float* Z[LP_NLAYER + 1];
    float* Hh[LP_NLAYER + 1];
    float* M[LP_NLAYER];
    float* OUT;
    const double* sfeat;
    const double* wsum;
};
void lp_forward(const LpSet& s, const float* P, const LpNet& b, hipStream_t st);
// the network's parameters on the device, pushed if the host copy is newer
// (call outside graph capture); nullptr on failure
const float* lpcg_device_params(spgo_lpcg* h);
int lpcg_band_blocks(const spgo_lpcg* h);

// One CG solve of a solver's reduced system inside an LM iteration: the block
// band of the dense H, features, the network (or block Jacobi), PCG into x.
struct LpSolve {
    int N = 0, W = 0;
    LpSet set{};
    LpNet net{};
    double *Hb = nullptr, *C = nullptr, *N1 = nullptr, *sfeat = nullptr, *wsum = nullptr, *Gd = nullptr, *Go = nullptr;
    double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *tmp = nullptr;
    int *bad = nullptr, *node_sys = nullptr, *sys_begin = nullptr;
    double* part = nullptr;   // cooperative PCG: 4 x coop_blocks partial sums
    int coop_blocks = 0;
    // blocks outside the band: per block row (N + 1) ranges, their columns and rows
    int nfar = 0;
    int *far_beg = nullptr, *far_col = nullptr, *far_row = nullptr;
    double* far_val = nullptr;
    std::vector<void*> ptrs;
    // far_beg / far_col: host CSR of the blocks outside the band (both
    // triangles), empty for a banded system
    int setup(int n_poses, int band_blocks, const std::vector<int>& far_beg_h = {},
              const std::vector<int>& far_col_h = {});   // 0 or a HIP error
    void release();
};
// x = CG solution of the leading 6N x 6N block of the dense H (lower triangle,
// row-major Pp x Pp) with right-hand side g; P: network parameters or nullptr
// (block Jacobi).  Writes ctl->pcg_iters and ctl->pcg_fail.
int lp_solve_launch(const LpSolve& b, const double* H, int Pp, const double* g, double* x, const float* P,
                    double tol, int maxit, Control* ctl, hipStream_t st);   // 0 or a HIP error

}  // namespace spgo
