// Internal types shared by the host runtime and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spgo {

// Upper bound on the pose span (max target - anchor) of one landmark group.
// The per-group dense block is 6*KMAX square.
constexpr int KMAX = 16;
constexpr int CHUNK_LM = 64;        // landmarks per LDS chunk (upper bound)
constexpr int CHUNK_OBS = 512;      // measurements per LDS chunk (upper bound)
constexpr int NB = 32;              // Cholesky block size
constexpr int WG = 256;             // threads per workgroup of the wide kernels
// Frame coordinates: most distinct target frames of one landmark group (and so
// of one landmark); a group's block covers its anchor and these frames.
constexpr int KTF = 64;

// Everything the linearisation needs to start a landmark group, in one record
// (one memory round trip): the group's fields and those of its first chunk.
struct alignas(16) GroupDesc {
    int anchor, K, chunk_begin, chunk_end;   // anchor: frame index in the state arrays
    int lm_begin, d_off, prob, a_loc;        // prob: problem of a batch (0 alone); a_loc: anchor inside it
    long long v_off, D_off;
    long long delta_off;                     // first entry of the problem's reduced-system solution
    long long r_off;                         // frame coordinates: the group's anchor blocks in Rz
    int4 chunk0;                  // (first active landmark, n landmarks, first obs, n obs)
    int soff0[KMAX + 2];          // slot offsets of the first chunk
    int gid;                      // index of the group's cost partial (its place in the layout)
};

// One landmark group whose pose span covers a given pose row block (the
// reduced system's row sums run over these; built once per layout).
struct FsEntry {
    int anchor, K6, d_off, pad;
    long long D_off;
};

// Frame coordinates: one contribution to a block (row frame r, column frame c)
// of the reduced system from a landmark group.  K6 > 0: a block of the group's
// target-frame block B (StateBuf::Dw, world axes), element (e, j) at
// off + e K6 + j (a diagonal block is read from its lower triangle); K6 = -1:
// minus the row sum R(s) of the target frame r against the anchor c
// (StateBuf::Rz, off + 6e + j); K6 = -2: the anchor's own block sum_s R(s).
// doff, on the diagonal only, indexes the gradient in Rz: minus b(s) (K6 > 0)
// or plus sum_s b(s) (K6 = -2); -1 off the diagonal.
struct ZCon {
    long long off;
    int K6, doff;
};
// k_form_frames reads an entry as one int4: off in .x/.y, K6 in .z, doff in .w
static_assert(sizeof(ZCon) == 16 && offsetof(ZCon, K6) == 8 && offsetof(ZCon, doff) == 12, "ZCon layout");

// One problem of a batch (a lone solver is a batch of one).  The per-problem
// kernels run one workgroup per problem.
struct ProbDesc {
    int N, pose_base, frame_base, lm_base;   // poses; first pose, frame and landmark slot in the state arrays
    int g_begin, g_end;                      // its landmark groups
    int Pp, D;                               // padded reduced-system size, profile depth in blocks of 16
    long long H_off, vec_off;                // its reduced system (Pp x Pp, row-major) and vectors (Pp)
    long long winv_off, Lp_off;              // inverse diagonal blocks and factored panels
    int first_off, rec_off;                  // block profile (Pp / 16 entries); first iteration record
};

// Device control block: which state buffer is current, damping, cost.
struct Control {
    int cur;                // index (0/1) of the accepted state buffers
    int iter;               // records written in this step
    int chol_ok;            // set by the factorisation of this iteration
    int pcg_fail;           // 0: this iteration's PCG converged and the Cholesky is skipped
    double lambda;
    double cost;            // cost of the current state
    double cost0;           // cost at the start of the step
    int pcg_iters;          // CG iterations of this iteration's step (0 without CG)
    int pad;
};

struct IterRecord {
    double cost_candidate, cost, lambda;
    int accepted, chol_ok;
    int pcg_iters, pcg_fail;   // CG iterations of the step; 1: the Cholesky computed it
};

struct Camera {
    double fx[2], fy[2], cx[2], cy[2];   // left, right
    double Trl[12];                       // right camera in left frame, 3x4
};

struct Robust {
    int kind;            // 0 cauchy, 1 huber, 2 none
    int near_boost;
    double param;        // c^2 or delta
};

// Problem layout on the device (built by the host, immutable during a step).
struct Layout {
    int N, F, L;                  // poses, frames, landmark slots
    int n_act;                    // active landmarks (with measurements)
    int n_groups, n_chunks, n_obs, kmax;
    int multi_chunk;              // some group holds more than one chunk
    int form_rows;                // reduced-system rows per workgroup of k_form_system (1 or 6)
    // measurements, chunk/slot-major order inside each group
    const int* ob_t;              // target frame
    const uint32_t* ob_meta;      // bits 0-15: slot-major -> landmark-major position in chunk,
                                  // bits 16-31: landmark index inside chunk
    const float4* ob_uv;          // uL vL uR vR
    // active landmarks, sorted by (anchor, slot id)
    const int* lm_id;             // landmark slot
    const float2* lm_ray;         // anchor-left pixel
    const int2* lm_pos;           // (first landmark-major position in chunk, count)
    // chunks
    const int4* chunk;            // (first active landmark, n landmarks, first obs, n obs)
    const int* chunk_soff;        // per chunk KMAX + 2 entries: first position of slot s (slot-major order)
    // groups
    const GroupDesc* grp_desc;    // per group, for k_assemble
    const int* fs_begin;          // per pose row block (N + 1): range in fs_list
    const FsEntry* fs_list;       // covering groups per pose row block, anchors ascending
    // frame coordinates (frame_mode = 1): the groups' frames and the plan of the
    // reduced system's blocks
    int frame_mode;
    int soff_stride;              // entries per chunk in chunk_soff
    const int* grp_frames;        // per group KTF + 1 frames: slot 0 the anchor, then its target frames
    const int* zp_row;            // per frame row r (N + 1 entries): its block entries zp_row[r] .. zp_row[r+1]
    const int* zp_col;            // column frame of each block entry (ascending within a row)
    const int* zp_cbeg;           // per block entry (+1): its contributions zp_cbeg[k] .. zp_cbeg[k+1]
    const ZCon* zp_con;
    int n_probs;                  // problems (1 for a lone solver)
    const ProbDesc* probs;
    const int* pose_prob;         // problem of every pose (N entries)
};

// One of the two state/linearisation buffer sets (current and candidate).
template <typename T>
struct StateBuf {
    double* X;        // N x 12 incremental poses (3x4 row-major), always double
    double* alpha;    // L inverse depths
    double* G;        // F x 12 global left poses
    double* GinvL;    // F x 12 inverses
    double* GinvR;    // F x 12 inverses of the right camera poses
    T* v;             // per landmark 6K_g values (B columns in the global frame)
    T* Ct;            // per active landmark C + eps
    T* g;             // per active landmark g_alpha
    T* D;             // per group (6K)^2 reduced block
    T* d;             // per group 6K reduced gradient
    double* cost;     // per group robust cost
    double* Rz;       // frame coordinates: per group the anchor's blocks and the gradient (see ZCon)
    double* Dw;       // frame coordinates: per group B in world axes (D holds it in the anchor's)
};

// The pose arrays of both state buffers (double in either precision).
struct PoseBufs {
    double* X[2];
    double* G[2];
    double* GinvL[2];
    double* GinvR[2];
};
template <typename T>
inline __host__ __device__ PoseBufs pose_bufs(const StateBuf<T>& b0, const StateBuf<T>& b1) {
    return PoseBufs{{b0.X, b1.X}, {b0.G, b1.G}, {b0.GinvL, b1.GinvL}, {b0.GinvR, b1.GinvR}};
}

// A pose update run by the last workgroup of another kernel (one problem).
struct PoseStep {
    ProbDesc pd;
    Camera cam;
    PoseBufs b;
    double step;
    int frame_mode;
};

}  // namespace spgo
