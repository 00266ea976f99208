// Host-side launchers of the gfx950 kernels (defined in k_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "spgo_internal.h"

namespace spgo {

struct AsmArgs {
    Layout lay;
    Camera cam;
    Robust rb;
    double eps, alpha_min, alpha_max;
    const Control* ctl;       // per problem (one alone), indexed by GroupDesc::prob
    int init;                 // 1: linearise the current state in place
    const double* delta;      // reduced-system solution (candidate mode)
};

template <typename T>
void launch_assemble(const AsmArgs& a, const StateBuf<T>* bufs, hipStream_t s);
int assemble_resident_blocks(int kmax, int fp32, int multi_chunk);

// the linearisation in frame coordinates (layouts with lay.frame_mode set)
template <typename T>
void launch_assemble_frames(const AsmArgs& a, const StateBuf<T>* bufs, hipStream_t s);
// landmarks per group of the frame-coordinate kernel for groups of kt target frames
int frames_landmarks_per_group(int kt, int fp64);

template <typename T>
void launch_form_system(const Layout& lay, const StateBuf<T>* bufs, Control* ctls, double* H, double* g,
                        hipStream_t s);

// dense blocked Cholesky within a block profile (NB-wide blocks): per step k the
// active block rows act[act_off[k] ..] below k whose profile reaches column
// block k, nact[k] of them (host array: launch sizes)
void launch_cholesky(double* H, double* Linv, int Pp, const int* act_off, const int* act, const int* nact,
                     Control* ctl, hipStream_t s);
// the NB-block profile and active lists of the dense Cholesky from the
// 16-row block profile fb16 (first nonzero 16-column block of each 16-row block)
void dense_profile(int Pp, const std::vector<int>& fb16, std::vector<int>& first, std::vector<int>& act_off,
                   std::vector<int>& act, std::vector<int>& nact);
// envelope Cholesky and both solves, one workgroup per problem of the table;
// Dmax the largest profile depth, lds the largest chol_band_lds of the problems
void launch_chol_band(const ProbDesc* probs, int n_probs, int Dmax, size_t lds, double* H, const int* first,
                      const double* g, double* x, double* winv, double* Lp, Control* ctl, hipStream_t s);
void launch_chol_split(double* H, int Pp, int D, int s0, int s1, const int* firstT, const int* firstB,
                       const double* g, double* x, double* winv, double* Lp, double* S, double* Sm, double* winvS,
                       double* LpS, Control* ctl, hipStream_t s);
int prepare_chol_split(int Pp, int D);
size_t chol_split_lds(int Pp, int D, int part);
int chol_band_max_blocks();
size_t chol_panel_doubles(int Pp);
int prepare_chol_band(int Dmax, size_t lds);
int chol_band_block();
size_t chol_band_lds(int Pp, int D);
// block cyclic reduction for profiles at most cr_max_depth() blocks of 16 deep:
// the factorisation and both solves of H x = g, 2 log2(super-rows) + 2 launches;
// work holds cr_work_doubles(Pp, D) doubles; gated: skipped when ctl->pcg_fail
// is 0 (the iteration's CG solve converged).  With ps, the last workgroup of
// the back-substitution also runs that pose update (x is its step); returns
// whether it did (not when the system is solved by the tail kernel alone).
bool launch_cr(const double* H, const double* g, double* x, int Pp, int D, double* work, Control* ctl, int gated,
               const PoseStep* ps, hipStream_t s);
size_t cr_work_doubles(int Pp, int D);
int cr_max_depth();
void launch_chol_solve(const double* H, const double* Linv, const double* g, double* delta, int Pp,
                       const int* first, const int* act_off, const int* act, double* ywork, int* sync, Control* ctl,
                       hipStream_t s);

template <typename T>
void launch_pose_update(const Layout& lay, const Camera& cam, StateBuf<T>* bufs, const Control* ctl,
                        const double* delta, double step_size, int init, hipStream_t s);

void launch_accept(const double* cost_buf0, const double* cost_buf1, const Layout& lay, Control* ctl,
                   IterRecord* rec, int reference_mode, double lam_up, double lam_down, double lam_min, double lam_max,
                   int init, hipStream_t s);

void launch_begin_step(Control* ctl, int reset_lambda, double lambda0, int keep_lin, hipStream_t s);
void launch_fill(double* p, double v, long long n, hipStream_t s);
void launch_set_diag(double* H, int Pp, int from, double v, hipStream_t s);

}  // namespace spgo
