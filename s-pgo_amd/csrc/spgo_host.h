// Host-side structures shared by the single-problem runtime and the batch
// runtime: device allocations, the host image of a problem's device layout
// and the solver handle behind include/spgo.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/spgo.h"
#include "spgo_internal.h"
#include "spgo_lpcg.h"

namespace spgo {

// Iterations captured in one hipGraph.  Measured on MI355X at 64 keyframes:
// with one iteration per graph the device idles 8 us between graph launches;
// two per graph halve that count, while graphs of eight or more iterations
// stall for milliseconds at a time and run slower overall.
constexpr int GRAPH_UNROLL = 2;

struct DevMem {
    std::vector<void*> ptrs;
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

struct HostLayout {
    std::vector<int> ob_t;
    std::vector<uint32_t> ob_meta;
    std::vector<float4> ob_uv;
    std::vector<int> lm_id;
    std::vector<float2> lm_ray;
    std::vector<int2> lm_pos;
    std::vector<int4> chunk;
    std::vector<int> chunk_soff;    // per chunk KMAX + 2 offsets: first position of each slot
    std::vector<int> grp_anchor, grp_K, grp_lm_begin, grp_d_off;
    std::vector<int2> grp_chunks;
    std::vector<long long> grp_v_off, grp_D_off;
    std::vector<int> anchor_grp;
    std::vector<GroupDesc> grp_desc;
    std::vector<int> fs_begin;
    std::vector<FsEntry> fs_list;
    int kmax = 0;
    long long v_total = 0, D_total = 0, d_total = 0;
    int anchors_used = 0;
    // every layout: first pose each pose row couples to in the reduced system
    std::vector<int> fpose;
    int band_need = 0;              // largest row - column distance, in poses
    // frame coordinates
    int frame_mode = 0;
    int kt = 0, cl = 0;             // kernel shape: target frames per group, landmarks per chunk
    std::vector<int> grp_frames;    // per group KTF + 1
    std::vector<long long> grp_r_off;   // per group: offset of its anchor blocks in Rz
    long long Rz_total = 0;
    std::vector<int> zp_row, zp_col, zp_cbeg;
    std::vector<ZCon> zp_con;
};

}  // namespace spgo

struct spgo_solver {
    using DevMem = spgo::DevMem;
    using HostLayout = spgo::HostLayout;
    using Layout = spgo::Layout;
    using Control = spgo::Control;
    using IterRecord = spgo::IterRecord;
    template <typename T>
    using StateBuf = spgo::StateBuf<T>;

    int N = 0, M = 0, L = 0, F = 0;
    long long max_meas = 0;
    spgo_options opt{};
    float step_size = 1.0f;
    float intr[8] = {0};
    float Trl[16] = {0};

    std::unordered_map<long long, int> cell_of;   // frame * L + feat -> cell
    std::vector<float> cells;                     // 6 floats per cell
    std::vector<int> ma, mt, ml;

    std::vector<double> X;        // N x 12
    std::vector<double> alpha;    // L
    bool host_newer = true;       // host mirror must be uploaded
    bool dev_newer = false;       // device state newer than host mirror
    bool layout_dirty = true;
    bool lin_valid = false;
    bool lambda_reset = true;
    bool graph_dirty = true;

    std::string err;

    // device
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevMem mem;          // problem + buffers (rebuilt with the layout)
    HostLayout hl;
    Layout lay{};
    spgo::ProbDesc pd1{};   // the problem table's one entry (host copy)
    int prec = SPGO_PREC_FP64;
    StateBuf<double> bd[2]{};
    StateBuf<float> bf[2]{};
    double *H = nullptr, *gvec = nullptr, *delta = nullptr, *Linv = nullptr;
    int Pp = 0;
    bool band = false;          // envelope Cholesky in one workgroup (narrow profile)
    const int* first = nullptr; // first nonzero block column of every block row
    int band_depth = 0;         // max over block rows of (row - first)
    int path = SPGO_FACTOR_DENSE;   // factorisation taken: SPGO_FACTOR_DENSE, _ENVELOPE, _ENVELOPE_SPLIT or _CR
    int s0 = 0, s1 = 0;         // split: separator block rows [s0, s1)
    const int* firstB = nullptr;   // split: profile of the reversed bottom part
    double *Ssep = nullptr, *Smrg = nullptr, *winvS = nullptr, *LpS = nullptr;   // split: separator scratch
    double* solve_y = nullptr;  // dense path: forward solution of the pipelined solves (Pp)
    int* solve_sync = nullptr;  // dense path: their tickets and ready flags (2 + 2 Pp / NB)
    double* Lpan = nullptr;     // envelope: factored panels for the backward solve
    double* crw = nullptr;      // cyclic reduction: per super-row blocks and vectors
    // dense path: the factor's block profile (NB-wide blocks) and per step the
    // block rows it reaches (nact on the host for the launch sizes)
    const int *dense_first = nullptr, *dense_act_off = nullptr, *dense_act = nullptr;
    std::vector<int> dense_nact;
    bool pcg = false;           // SPGO_FACTOR_PCG: CG first, the envelope path above as fallback
    spgo_lpcg* net = nullptr;   // its preconditioner network (nullptr: block Jacobi)
    const float* net_params = nullptr;   // the network's device parameters for this step
    spgo::LpSolve lps;          // CG buffers
    int* pcg_flag_host = nullptr;   // pinned: CG convergence flag read back each iteration
    Control* ctl_host = nullptr;    // pinned: the control block and records read back after a step
    IterRecord* rec_host = nullptr; // rec_host_cap entries
    int rec_host_cap = 0;
    Control* ctl = nullptr;
    IterRecord* rec = nullptr;
    int rec_cap = 0;
    // captured iterations: [0] holds graph_iters[0] iterations and is launched
    // repeatedly, [1] the remainder of a call
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_iters[2] = {0, 0};

    std::vector<IterRecord> last_rec;
    double last_cost0 = 0.0;
    double last_ms = 0.0;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace spgo {

int check_hip(spgo_solver* h, hipError_t e, const char* what);
// rotation part of a 3x4 pose replaced by its polar factor
void project_rotation(double* P);
// device layout of one problem; tile shape from the solver's options, or the
// given chunk_landmarks / group_chunks (0: default)
int build_layout(spgo_solver* h, HostLayout& out);
int build_layout(spgo_solver* h, int chunk_landmarks, int group_chunks, HostLayout& out);
// the layout in frame coordinates (spgo_frames.hip)
int build_layout_frames(spgo_solver* h, HostLayout& out);
// layout, state upload and `iterations` LM iterations (0: linearise only)
int run_step(spgo_solver* h, int iterations);
// the layout in the coordinates the solver's options ask for
int build_any_layout(spgo_solver* h, HostLayout& out);
// global left poses G (N + 1 of them, 3x4) of the incremental poses X (N x 12)
void global_chain(int N, const double* X, double* G);
// incremental-pose system H = U^T Z U, g = U^T gz from the frame-coordinate
// system (dense lower triangle of Z, row stride ld), U the map delta -> z at
// the global poses G (F x 12); H, g of size 6N
void frames_to_incremental(int N, const double* Z, int ld, const double* gz, const double* G, double* H, double* g);
// rows of the padded reduced pose system of n_poses poses
int reduced_size(int n_poses);
// the assembly groups in launch order, each with gid = its place in `in`
// (see spgo_runtime.hip)
std::vector<GroupDesc> launch_order(const std::vector<GroupDesc>& in);
// block profile of the reduced system: fb[i] first nonzero block column of
// block row i; returns the depth max(i - fb[i])
int profile_blocks(int n_poses, const HostLayout& L, std::vector<int>& fb);
// device state into the host mirror if the device is newer
int pull_state(spgo_solver* h);
Camera make_camera(const spgo_solver* h);
// options within the ranges the kernels support; why names the first bad field
bool valid_opts(const spgo_options* o, std::string& why);

}  // namespace spgo
