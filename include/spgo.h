/*
 * spgo.h -- C ABI of the MI355X (gfx950) stereo bundle-adjustment LM step.
 *
 * One opaque solver handle replaces the reference's C++ facade `struct Solver`
 * (include/Solver.hpp:18-85) and its Cython wrapper `CudaSolver`
 * (src/CythonAPI/Solver.pyx:16-107).  Every entry point returns an int status:
 * 0 on success, a negative SPGO_E* code on failure, with a message available
 * from spgo_last_error().  No torch, HIP or C++ types cross this boundary.
 *
 * Conventions (as in the reference):
 *   N incremental poses X_k = T_curr_to_next (N+1 frames, frame 0 fixed at I),
 *   N*M landmark slots holding inverse depths, 4x4 poses row-major float32,
 *   intrinsics packed [fxL, fyL, cxL, cyL, fxR, fyR, cxR, cyR],
 *   T_r_to_l = pose of the right camera in the left camera frame.
 */
#ifndef SPGO_H
#define SPGO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPGO_OK 0
#define SPGO_E_ARG -1        /* invalid argument or id out of range           */
#define SPGO_E_CAPACITY -2   /* more measurements than N(N+1)M                 */
#define SPGO_E_LAYOUT -3     /* problem structure this build does not support  */
#define SPGO_E_DEVICE -4     /* no gfx950 device / HIP runtime error           */
#define SPGO_E_STATE -5      /* call not valid in the solver's current state   */

/* Robust kernels.  CAUCHY with param c^2 = 9 and the x4 near-frame boost is
 * the reference weighting (Solver_cuda.cu:215-228, include/LMutils.h:48). */
#define SPGO_ROBUST_CAUCHY 0
#define SPGO_ROBUST_HUBER 1
#define SPGO_ROBUST_NONE 2

/* REFERENCE: lambda = 0, every step accepted (Solver_cuda.cu:594-782).
 * LM: lambda on the reduced pose diagonal, cost-based accept/reject. */
#define SPGO_MODE_REFERENCE 0
#define SPGO_MODE_LM 1

/* FP64: every phase in double.  MIXED: residuals, Jacobians, assembly and the
 * landmark Schur contraction in float (FP32 MFMA); the reduced pose system is
 * summed, factored and solved in double. */
#define SPGO_PREC_FP64 0
#define SPGO_PREC_MIXED 1

/* Reduced-system factorisation.  ENVELOPE: the profile Cholesky in one
 * workgroup; ENVELOPE_SPLIT: the same, split at a separator in the middle of
 * the profile into a top and a bottom part factored concurrently; DENSE: the
 * multi-kernel blocked Cholesky; CR: block cyclic reduction for profiles at
 * most 4 blocks deep (the band is block tridiagonal in super-rows of 16 D
 * rows, eliminated every other one per level, log2 of their number levels:
 * the Cholesky factorisation of H with its super-rows reordered).  AUTO picks ENVELOPE_SPLIT for long narrow
 * profiles (at least 2 D + 6 blocks of 16 rows, D the profile depth),
 * ENVELOPE for short narrow ones and DENSE for wide ones.  PCG: conjugate
 * gradients preconditioned by the network set with spgo_set_preconditioner
 * (block Jacobi without one) to pcg_tol; when CG does not converge in
 * pcg_max_iter iterations the Cholesky AUTO would pick computes the step
 * instead. */
#define SPGO_FACTOR_AUTO 0
#define SPGO_FACTOR_DENSE 1
#define SPGO_FACTOR_ENVELOPE 2
#define SPGO_FACTOR_ENVELOPE_SPLIT 3
#define SPGO_FACTOR_PCG 4
#define SPGO_FACTOR_CR 5

/* Parameters of the reduced pose system.  INCREMENTAL: the increments of the
 * incremental poses X_k themselves (the reference's system, Solver_cuda.cu:
 * 636-736); every landmark couples the consecutive poses its anchor..target
 * range covers, so the system is block-banded.  FRAMES: the exact change of
 * variables z_k = sum_{i >= k} Ad(G_i) delta_i (z_N = 0), in which a
 * measurement couples only its anchor and target frames: landmarks observed
 * far apart (loop revisits) add sparse blocks instead of a dense square.  The
 * damping lambda ||delta||^2 is carried over exactly, so both give the same
 * step.  AUTO takes INCREMENTAL when every landmark spans at most 16 poses
 * (and at most 512 landmark groups reach one pose), FRAMES otherwise. */
#define SPGO_COORD_AUTO 0
#define SPGO_COORD_INCREMENTAL 1
#define SPGO_COORD_FRAMES 2

typedef struct spgo_options {
    int32_t mode;            /* SPGO_MODE_*                                    */
    int32_t robust;          /* SPGO_ROBUST_*                                  */
    int32_t near_boost;      /* 1: weight x4 when |target - anchor| < 2        */
    int32_t precision;       /* SPGO_PREC_*                                    */
    double robust_param;     /* c^2 (Cauchy) or delta in px (Huber)            */
    double eps;              /* landmark diagonal regulariser (LMutils.h:50)   */
    double lambda0;          /* initial LM damping                             */
    double lam_up, lam_down, lam_min, lam_max;
    double alpha_min, alpha_max;  /* inverse-depth clamp (LMutils.h:54-55)     */
    int32_t reserved0;       /* unused, zero                                    */
    int32_t use_graph;       /* replay the LM iterations from captured graphs  */
    int32_t chunk_landmarks; /* tile shape: landmarks per LDS chunk (<= 64)    */
    int32_t group_chunks;    /* tile shape: chunks per workgroup; 0 (default):  */
                             /* the count that best fills the resident workgroups */
    int32_t factorization;   /* SPGO_FACTOR_*                                   */
    int32_t pcg_max_iter;    /* CG iterations before the Cholesky fallback (PCG) */
    int32_t coordinates;     /* SPGO_COORD_*                                    */
    int32_t reserved[5];
    double pcg_tol;          /* CG stops at ||r|| <= pcg_tol ||g|| (PCG)         */
} spgo_options;

typedef struct spgo_iter_stats {
    double cost_candidate;   /* robust cost of the candidate state             */
    double cost;             /* cost of the state kept after accept/reject     */
    double lambda;           /* damping used by this iteration's solve         */
    int32_t accepted;
    int32_t chol_ok;         /* 0 if the reduced system was not positive definite */
    int32_t pcg_iters;       /* CG iterations of the step (SPGO_FACTOR_PCG)      */
    int32_t pcg_fallback;    /* 1: the Cholesky computed the step               */
} spgo_iter_stats;

typedef struct spgo_layout_info {
    int32_t n_poses, n_slots, n_measurements, n_landmarks;
    int32_t n_groups, n_chunks, max_span, n_anchors_used;
    int32_t factorization;   /* SPGO_FACTOR_DENSE, _ENVELOPE, _ENVELOPE_SPLIT, _CR or _PCG: the path spgo_step takes */
    int32_t profile_depth;   /* blocks of 16 below the diagonal in the reduced system's profile */
    int32_t reduced_size;    /* rows of the padded reduced pose system                          */
    int32_t coordinates;     /* SPGO_COORD_INCREMENTAL or _FRAMES: the system spgo_step solves */
} spgo_layout_info;

typedef struct spgo_solver spgo_solver;

/* Defaults: LM mode, Cauchy c^2 = 9 with boost, FP64, lambda0 = 1e-4. */
int spgo_default_options(spgo_options* out);

/* Solver(int N, int M) (include/Solver.hpp:20-57; Solver.pyx:19-20).  Device
 * memory is taken on the first spgo_step, so every host-side call works
 * without a GPU. */
int spgo_create(int32_t n_poses, int32_t m_per_frame, const spgo_options* opts, spgo_solver** out);
int spgo_destroy(spgo_solver* h);
int spgo_set_options(spgo_solver* h, const spgo_options* opts);
int spgo_get_options(spgo_solver* h, spgo_options* out);

/* Solver::loadCalibration (Solver.cpp:29-38; Solver.pyx:28-32). */
int spgo_load_calibration(spgo_solver* h, const float intrinsics[8], const float T_r_to_l[16]);
/* Solver::getCalibration (Solver.cpp:101-109; Solver.pyx:75-88). */
int spgo_get_calibration(spgo_solver* h, float intrinsics[8], float T_r_to_l[16]);

/* Solver::writeObservations (Solver.cpp:40-56; Solver.pyx:34-38): writes the
 * (target, feat) cell and appends one measurement. */
int spgo_write_observation(spgo_solver* h, int32_t anchor, int32_t target, int32_t feat,
                           const float left[3], const float right[3]);
/* Bulk form of the same: count measurements, uvLR = count x [uL, vL, uR, vR]
 * (homogeneous third coordinate 1).  No reference counterpart: replaces
 * count calls of writeObservations. */
int spgo_write_observations_bulk(spgo_solver* h, int32_t count, const int32_t* anchor,
                                 const int32_t* target, const int32_t* feat, const float* uvLR);
/* Solver::getObservation (Solver.cpp:86-98; Solver.pyx:59-72). */
int spgo_get_observation(spgo_solver* h, int32_t frame, int32_t feat, float left[3], float right[3]);

/* Solver::step -> updateState (Solver.cpp:5-27; Solver_cuda.cu:557-795). */
int spgo_step(spgo_solver* h, int32_t iterations);

/* Solver::getIncrementalPose (Solver.cpp:59-67; Solver.pyx:41-52). */
int spgo_get_incremental_pose(spgo_solver* h, int32_t k, float T[16]);
/* Solver::writeIncrementalPose, rows 0-2 only (Solver.cpp:69-84). */
int spgo_write_incremental_pose(spgo_solver* h, int32_t k, const float T[16]);
/* Solver::loadInverseDepths / getInverseDepths (Solver.cpp:112-128), N*M values. */
int spgo_load_inverse_depths(spgo_solver* h, const float* alpha);
int spgo_get_inverse_depths(spgo_solver* h, float* alpha);
/* Optimizer.initalize_depth_with_disparity (PythonUtils/Optimizer.py:87-114):
 * every landmark whose anchor-frame cell holds both pixels gets the inverse of
 * the least-squares stereo depth along its left ray,
 *   d = ((v x b_l) . (v x t)) / |v x b_l|^2,  v = R b_r,
 * with b_c = K_c^-1 p_c and (R, t) = T_r_to_l.  Landmarks without that cell or
 * with a non-positive depth keep their inverse depth.  n_set (optional)
 * receives the number of landmarks set. */
int spgo_init_inverse_depths_from_disparity(spgo_solver* h, int32_t* n_set);

/* CudaSolver.setStepSize (Solver.pyx:106-107): scales the pose update only. */
int spgo_set_step_size(spgo_solver* h, float step_size);

/* Double-precision state access: poses N x 16 row-major, alpha N*M.  No
 * reference counterpart (the reference state is float32). */
int spgo_get_state_f64(spgo_solver* h, double* poses, double* alpha);
int spgo_set_state_f64(spgo_solver* h, const double* poses, const double* alpha);

/* Per-iteration record of the last spgo_step: up to max_records entries,
 * n_out set to the number of iterations run, initial_cost to the cost of the
 * state the step started from. */
int spgo_get_stats(spgo_solver* h, spgo_iter_stats* out, int32_t max_records, int32_t* n_out,
                   double* initial_cost);
/* Current LM damping (persists across spgo_step calls; reset to lambda0 when
 * the state is written). */
int spgo_get_lambda(spgo_solver* h, double* lambda);
int spgo_set_lambda(spgo_solver* h, double lambda);

/* Linearise the current state and return the reduced pose system
 * H = A - B diag(1/(C+eps)) B^T + lambda I (P x P row-major, P = 6N, both
 * triangles filled), its right-hand side g = g_T - B (g_alpha/(C+eps)) and the
 * robust cost.  The products of Solver_cuda.cu:636-730, exposed for
 * verification; no reference entry point returns them (the declared
 * get_H_and_g_* of include/Solver.hpp:68-69 are never defined). */
int spgo_linearize(spgo_solver* h, double lambda, double* H, double* g, double* cost);

/* The block band of the reduced system spgo_step factors, in the solver's own
 * coordinates (spgo_layout_info.coordinates), after linearising the current
 * state with damping lambda: Hband[i][d] = H_{i,i-d} (6x6 row-major,
 * d = 0..band_blocks, zero where i - d < 0), g (N x 6) and the robust cost.
 * Blocks farther from the diagonal (loop revisits) are left out.  These are
 * the systems spgo_lpcg_load_systems takes.  No reference counterpart. */
int spgo_reduced_band(spgo_solver* h, double lambda, int32_t band_blocks, double* Hband, double* g, double* cost);

/* spgo_reduced_band without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, serving loops with a fixed memory budget).  The band is staged
 * in a per-device scratch array of 2^20 doubles (8 MiB, held until the process
 * ends): N (band_blocks + 1) 36 must not exceed that (SPGO_E_ARG otherwise).
 * The solver must already hold its device state -- stepped or linearised once
 * since its measurements or options last changed (SPGO_E_STATE otherwise).
 * Calls that overlap on one GPU share the scratch array and race. */
int spgo_reduced_band_noalloc(spgo_solver* h, double lambda, int32_t band_blocks, double* Hband, double* g,
                              double* cost);

/* Host-side layout summary of the measurements written so far (builds the
 * layout without a GPU). */
int spgo_problem_info(spgo_solver* h, spgo_layout_info* out);

/* Elapsed device time (ms) of the iterations of the last spgo_step, measured
 * with events on the solver's stream. */
int spgo_last_step_ms(spgo_solver* h, double* ms);

/* Run `iterations` LM iterations with one launch per phase (no graph) and
 * HIP events around each phase on the solver's stream; ms[p] receives the
 * mean device time of phase p per iteration:
 * 0 reduced-system formation, 1 Cholesky factorisation, 2 triangular solves,
 * 3 pose update + global chain, 4 linearisation (assembly + landmark Schur),
 * 5 accept/reject.  No reference counterpart (instrumentation). */
#define SPGO_N_PHASES 6
int spgo_profile_phases(spgo_solver* h, int32_t iterations, double ms[SPGO_N_PHASES]);

/* Message of the last failed call on h (h == NULL: last failed spgo_create). */
const char* spgo_last_error(spgo_solver* h);

/* Wait for all work the library queued on the device (every solver, batch and
 * network) to finish; SPGO_E_DEVICE without a device.  No reference
 * counterpart (the reference's step is synchronous, Solver_cuda.cu:557-795). */
int spgo_device_synchronize(void);

struct spgo_lpcg;
/* The learned preconditioner of SPGO_FACTOR_PCG (NULL: block Jacobi).  The
 * solver reads the network's parameters at every step, so training it later
 * takes effect; it must outlive its use.  Its band (spgo_lpcg_create) must
 * cover the problem's: landmark spans of K poses need K - 1 blocks.  No
 * reference counterpart. */
int spgo_set_preconditioner(spgo_solver* h, struct spgo_lpcg* net);

/* ---- batches of independent problems ----------------------------------------
 * No reference counterpart: the reference steps one Solver per problem
 * (include/Solver.hpp:18-85, one launch sequence per step in
 * Solver_cuda.cu:557-795).  A batch steps many solver handles together: their
 * measurements, calibration and state are written through the calls above,
 * then each LM phase runs as one launch over all of them (the reduced systems
 * factored by the envelope Cholesky, one workgroup per problem).  Every
 * problem keeps its own damping and accept/reject, so it takes the steps it
 * would take alone; the options (precision, robust kernel, tile shape) are the
 * batch's and apply to all problems.  The problems must share one stereo
 * calibration; the pose update is applied with step size 1. */
typedef struct spgo_batch spgo_batch;

typedef struct spgo_batch_info {
    int32_t n_problems;
    int32_t n_poses, n_landmarks, n_measurements, n_groups;   /* totals; 0 before the first step */
    int32_t max_depth;        /* deepest reduced-system profile, blocks of 16 */
    int32_t max_span;         /* longest landmark pose span                   */
    int32_t chunk_landmarks, group_chunks;   /* tile shape of the options    */
    int32_t reserved[7];
} spgo_batch_info;

/* The handles must outlive the batch.  NULL opts = spgo_default_options. */
int spgo_batch_create(int32_t n_problems, spgo_solver* const* problems, const spgo_options* opts, spgo_batch** out);
int spgo_batch_destroy(spgo_batch* b);
int spgo_batch_set_options(spgo_batch* b, const spgo_options* opts);
int spgo_batch_get_options(spgo_batch* b, spgo_options* out);
/* The first step (and the first after spgo_batch_reload, or after a change of
 * precision or tile shape) reads every problem's measurements and current
 * state; later steps continue from the batch's own device state. */
int spgo_batch_reload(spgo_batch* b);
/* Re-read only the problems' states (not their measurements) and reset every
 * problem's damping to lambda0; the next step linearises them first. */
int spgo_batch_reset_states(spgo_batch* b);
int spgo_batch_step(spgo_batch* b, int32_t iterations);
/* State of one problem, as spgo_get_state_f64 (the batch's state once stepped). */
int spgo_batch_get_state_f64(spgo_batch* b, int32_t problem, double* poses, double* alpha);
/* Write every problem's batch state into its solver, as spgo_set_state_f64. */
int spgo_batch_store(spgo_batch* b);
/* Per-iteration records of one problem from the last spgo_batch_step. */
int spgo_batch_get_stats(spgo_batch* b, int32_t problem, spgo_iter_stats* out, int32_t max_records,
                         int32_t* n_out, double* initial_cost);
/* Device time (ms) of the iterations of the last spgo_batch_step. */
int spgo_batch_last_step_ms(spgo_batch* b, double* ms);
int spgo_batch_get_info(spgo_batch* b, spgo_batch_info* out);
/* As spgo_profile_phases, over the whole batch (phase 2, the solves, is part of
 * phase 1: the envelope kernel factors and solves). */
int spgo_batch_profile_phases(spgo_batch* b, int32_t iterations, double ms[SPGO_N_PHASES]);
/* Tile-shape search over chunk_landmarks in {16, 32, 48, 64} and group_chunks
 * in {1, 2, 4}: each shape is timed over `iterations` LM iterations from the
 * problems' current states; the fastest is kept in the batch's options and the
 * batch is left to reload the problems on its next step. */
int spgo_batch_autotune(spgo_batch* b, int32_t iterations, int32_t* chunk_landmarks, int32_t* group_chunks,
                        double* best_ms);
/* Message of the last failed call on b (b == NULL: last failed spgo_batch_create). */
const char* spgo_batch_last_error(spgo_batch* b);

/* ---- learned preconditioner for CG on the reduced system ---------------------
 * No reference counterpart: the reference factors the reduced system with a
 * dense Cholesky (linear_solvers.cu:195-218).  A message-passing network over
 * the pose covisibility band (three layers, hidden width 64, FP32) predicts a
 * factored approximate inverse P = G^T G of a block-banded reduced system H,
 *   G_ii = (I + tril(A_i)) C_i,  G_{i,i-1} = B_i C_{i-1},
 * C_i the inverse Cholesky factor of H_ii; an untrained network (head at zero)
 * gives block Jacobi.  Systems are block bands: Hband[i][d] = H_{i,i-d}
 * (6x6 row-major) for d = 0..band_blocks (1..16), nodes of all systems
 * concatenated; g is T x 6.  Training minimises the mean over systems of
 * Kaporin's condition measure of P^1/2 H P^1/2 (relative to block Jacobi),
 * log(tr(G H G^T) / 6n) - log det(G^T G) / 6n, with Adam; forward, backward,
 * loss and PCG run in this library's kernels. */
typedef struct spgo_lpcg spgo_lpcg;   /* declared above with spgo_set_preconditioner */
int spgo_lpcg_create(int32_t band_blocks, uint64_t seed, spgo_lpcg** out);
int spgo_lpcg_destroy(spgo_lpcg* h);
int spgo_lpcg_n_params(spgo_lpcg* h, int32_t* n);
/* Parameters, FP32 flat: W_in (64 x 40), b_in, per layer W_s, W_m (64 x 64), b,
 * then W_o (57 x 64), W_k (57 x 40), b_o.  Setting them resets Adam. */
int spgo_lpcg_get_params(spgo_lpcg* h, float* out);
int spgo_lpcg_set_params(spgo_lpcg* h, const float* in);
int spgo_lpcg_load_systems(spgo_lpcg* h, int32_t n_systems, const int32_t* n_nodes, const double* Hband,
                           const double* g);
/* Mean loss and its gradient (FP32, flat like the parameters) over the loaded
 * systems; block Jacobi scores 0. */
int spgo_lpcg_loss_grad(spgo_lpcg* h, double* loss, float* grad);
/* `steps` Adam steps (beta 0.9 / 0.999) over all loaded systems; losses
 * (optional, steps entries) receives the loss before each step. */
int spgo_lpcg_train(spgo_lpcg* h, int32_t steps, float lr, float* losses);
/* PCG on every loaded system from x = 0 to ||r|| <= tol ||g||: learned != 0
 * uses the network's preconditioner, 0 block Jacobi.  iters, relres (per
 * system) and x (T x 6) are optional outputs. */
int spgo_lpcg_solve(spgo_lpcg* h, int32_t learned, double tol, int32_t max_iter, int32_t* iters, double* relres,
                    double* x);
/* Device ms of the last solve: [0] network + preconditioner blocks, [1] PCG. */
int spgo_lpcg_last_ms(spgo_lpcg* h, double ms[2]);
/* Node inputs (T x 40, FP32) and inverse Cholesky factors C_i (T x 36). */
int spgo_lpcg_get_features(spgo_lpcg* h, float* X, double* C);
const char* spgo_lpcg_last_error(spgo_lpcg* h);

#ifdef __cplusplus
}
#endif
#endif /* SPGO_H */
