// Batch runtime: many independent stereo bundle-adjustment problems stepped
// together, one launch per LM phase for all of them.
//
// The reference runs one Solver object per problem (include/Solver.hpp:18-85),
// each step a full launch sequence of its own (Solver_cuda.cu:557-795).  Here
// the problems stay ordinary solver handles on the host (measurements,
// calibration and state are written through the single-problem API); a batch
// concatenates their device layouts into one:
//   * poses, frames and landmark slots of problem p follow those of problems
//     0..p-1 in the state arrays, so k_pose_update and k_assemble index the
//     concatenated arrays directly (frame ids in the measurement list and the
//     group descriptors are global; the pose-update delta is read per problem
//     through GroupDesc::delta_off / a_loc);
//   * each problem keeps its own reduced system (Pp x Pp), profile, control
//     block (current buffer, damping, cost) and iteration records, described by
//     one ProbDesc entry;
//   * per iteration: k_form_system over every problem's pose rows, the envelope
//     Cholesky with one workgroup per problem, k_pose_update and k_accept with
//     one workgroup per problem, k_assemble over every landmark group -- five
//     launches for the whole batch, captured once in a hipGraph.
// Accept/reject and damping are per problem, so every problem takes exactly the
// steps it would take alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spgo.h"
#include "spgo_host.h"
#include "spgo_internal.h"
#include "spgo_kernels.h"

using namespace spgo;

namespace {
thread_local std::string g_batch_create_error;
constexpr int FS_LIMIT = 512;   // covering groups k_form_system stages per pose row
}  // namespace

struct spgo_batch {
    std::vector<spgo_solver*> pr;
    spgo_options opt{};
    std::string err;

    bool dev_ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // built from the problems by the first step after a (re)load
    bool dirty = true;
    bool lin_valid = false;
    int prec = SPGO_PREC_FP64;
    int chunk_lm = 0, group_chunks = 0;   // tile shape in use (0: default)
    DevMem mem;
    Layout lay{};
    std::vector<ProbDesc> pd;   // host copy of the problem table
    ProbDesc* pd_dev = nullptr;
    StateBuf<double> bd[2]{};
    StateBuf<float> bf[2]{};
    double *H = nullptr, *gvec = nullptr, *delta = nullptr, *winv = nullptr, *Lp = nullptr;
    const int* first = nullptr;
    Control* ctl = nullptr;
    IterRecord* rec = nullptr;
    int rec_cap = 0;
    int Dmax = 0;
    size_t chol_lds = 0;
    long long tot_poses = 0, tot_frames = 0, tot_slots = 0, tot_meas = 0, tot_act = 0;
    int tot_groups = 0, max_span = 0;
    Camera cam{};
    // captured iterations: [0] holds graph_iters[0] iterations and is launched
    // repeatedly, [1] the remainder of a call
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_iters[2] = {0, 0};

    std::vector<IterRecord> last_rec;   // n_probs x last_iters
    std::vector<double> cost0;
    int last_iters = 0;
    double last_ms = 0.0;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace {

int bcheck(spgo_batch* b, hipError_t e, const char* what) {
    if (e == hipSuccess) return SPGO_OK;
    return b->fail(SPGO_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename V>
int bupload(spgo_batch* b, const std::vector<V>& v, const V** dst) {
    V* p = b->mem.alloc<V>(v.size());
    if (!p) return b->fail(SPGO_E_DEVICE, "device allocation failed");
    if (!v.empty()) {
        const int rc = bcheck(b, hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice), "hipMemcpy");
        if (rc) return rc;
    }
    *dst = p;
    return SPGO_OK;
}

void destroy_graph_slot(spgo_batch* b, int k) {
    if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    b->gexec[k] = nullptr;
    b->graph[k] = nullptr;
    b->graph_iters[k] = 0;
}

void destroy_graph(spgo_batch* b) {
    destroy_graph_slot(b, 0);
    destroy_graph_slot(b, 1);
}

int ensure_device(spgo_batch* b) {
    if (b->dev_ready) return SPGO_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return b->fail(SPGO_E_DEVICE, "no HIP device available");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) return bcheck(b, e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return b->fail(SPGO_E_DEVICE, std::string("device 0 is ") + prop.gcnArchName +
                                          ", this library is built for gfx950");
    e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bcheck(b, e, "hipStreamCreate");
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
        return b->fail(SPGO_E_DEVICE, "hipEventCreate failed");
    b->dev_ready = true;
    return SPGO_OK;
}

template <typename T>
int alloc_state(spgo_batch* b, StateBuf<T>* s, long long v_total, long long D_total, long long d_total) {
    for (int i = 0; i < 2; ++i) {
        s[i].X = b->mem.alloc<double>((size_t)b->tot_poses * 12);
        s[i].alpha = b->mem.alloc<double>((size_t)b->tot_slots);
        s[i].G = b->mem.alloc<double>((size_t)b->tot_frames * 12);
        s[i].GinvL = b->mem.alloc<double>((size_t)b->tot_frames * 12);
        s[i].GinvR = b->mem.alloc<double>((size_t)b->tot_frames * 12);
        s[i].v = b->mem.alloc<T>((size_t)v_total);
        s[i].Ct = b->mem.alloc<T>((size_t)b->tot_act);
        s[i].g = b->mem.alloc<T>((size_t)b->tot_act);
        s[i].D = b->mem.alloc<T>((size_t)D_total);
        s[i].d = b->mem.alloc<T>((size_t)d_total);
        s[i].cost = b->mem.alloc<double>((size_t)std::max(b->tot_groups, 1));
        if (!s[i].X || !s[i].alpha || !s[i].G || !s[i].GinvL || !s[i].GinvR || !s[i].v || !s[i].Ct || !s[i].g ||
            !s[i].D || !s[i].d || !s[i].cost)
            return b->fail(SPGO_E_DEVICE, "device allocation failed");
    }
    return SPGO_OK;
}

// Upload every problem's host state (rotations projected) into both buffer sets.
int push_states(spgo_batch* b) {
    std::vector<double> X((size_t)b->tot_poses * 12), al((size_t)b->tot_slots);
    for (size_t p = 0; p < b->pr.size(); ++p) {
        spgo_solver* h = b->pr[p];
        const ProbDesc& d = b->pd[p];
        int rc = pull_state(h);
        if (rc) return b->fail(rc, "problem " + std::to_string(p) + ": " + h->err);
        for (int k = 0; k < h->N; ++k) {
            double* dst = &X[12 * ((size_t)d.pose_base + k)];
            std::memcpy(dst, &h->X[12 * (size_t)k], 12 * sizeof(double));
            project_rotation(dst);
        }
        std::memcpy(&al[d.lm_base], h->alpha.data(), h->alpha.size() * sizeof(double));
    }
    for (int i = 0; i < 2; ++i) {
        double* Xd = b->prec == SPGO_PREC_MIXED ? b->bf[i].X : b->bd[i].X;
        double* ad = b->prec == SPGO_PREC_MIXED ? b->bf[i].alpha : b->bd[i].alpha;
        int rc = bcheck(b, hipMemcpy(Xd, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        if (rc) return rc;
        rc = bcheck(b, hipMemcpy(ad, al.data(), al.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        if (rc) return rc;
    }
    b->lin_valid = false;
    return SPGO_OK;
}

// Concatenate the problems' layouts and allocate the batch's device buffers.
int setup(spgo_batch* b) {
    destroy_graph(b);
    b->mem.release();
    b->rec = nullptr;
    b->rec_cap = 0;
    const int np = (int)b->pr.size();
    b->prec = b->opt.precision;
    b->chunk_lm = b->opt.chunk_landmarks;
    b->group_chunks = b->opt.group_chunks;
    b->cam = make_camera(b->pr[0]);

    HostLayout C;   // concatenation
    std::vector<int> first_all, pose_prob;
    b->pd.assign(np, ProbDesc{});
    long long pose_base = 0, frame_base = 0, lm_base = 0, act_base = 0, obs_base = 0, chunk_base = 0;
    long long v_base = 0, D_base = 0, d_base = 0, H_off = 0, vec_off = 0, winv_off = 0, Lp_off = 0;
    int g_base = 0, fs_base = 0;
    b->Dmax = 0;
    b->chol_lds = 0;
    b->max_span = 0;
    const size_t lds_lim = 160 * 1024 - 1024;
    C.fs_begin.clear();
    for (int p = 0; p < np; ++p) {
        spgo_solver* h = b->pr[p];
        HostLayout L;
        int rc = build_layout(h, b->chunk_lm, b->group_chunks, L);
        if (rc) return b->fail(rc, "problem " + std::to_string(p) + ": " + h->err);
        std::vector<int> fb;
        const int D = profile_blocks(h->N, L, fb);
        const int Pp = reduced_size(h->N);
        if (D > chol_band_max_blocks() || chol_band_lds(Pp, D) > lds_lim)
            return b->fail(SPGO_E_LAYOUT, "problem " + std::to_string(p) +
                                              ": reduced-system profile too deep for the batched envelope Cholesky");
        for (int j = 0; j < h->N; ++j) {
            const int amin = std::max(0, j - L.kmax + 1);
            if (L.anchor_grp[j + 1] - L.anchor_grp[amin] > FS_LIMIT)
                return b->fail(SPGO_E_LAYOUT, "problem " + std::to_string(p) + ": too many landmark groups reach pose " +
                                                  std::to_string(j) + "; raise chunk_landmarks or group_chunks");
        }
        ProbDesc& d = b->pd[p];
        d.N = h->N;
        d.pose_base = (int)pose_base;
        d.frame_base = (int)frame_base;
        d.lm_base = (int)lm_base;
        d.g_begin = g_base;
        d.g_end = g_base + (int)L.grp_desc.size();
        d.Pp = Pp;
        d.D = D;
        d.H_off = H_off;
        d.vec_off = vec_off;
        d.winv_off = winv_off;
        d.Lp_off = Lp_off;
        d.first_off = (int)first_all.size();
        d.rec_off = 0;
        b->Dmax = std::max(b->Dmax, D);
        b->chol_lds = std::max(b->chol_lds, chol_band_lds(Pp, D));
        b->max_span = std::max(b->max_span, L.kmax);

        // measurements, landmarks, chunks: frame ids and landmark slots made global
        for (int t : L.ob_t) C.ob_t.push_back(t + (int)frame_base);
        C.ob_meta.insert(C.ob_meta.end(), L.ob_meta.begin(), L.ob_meta.end());
        C.ob_uv.insert(C.ob_uv.end(), L.ob_uv.begin(), L.ob_uv.end());
        for (int id : L.lm_id) C.lm_id.push_back(id + (int)lm_base);
        C.lm_ray.insert(C.lm_ray.end(), L.lm_ray.begin(), L.lm_ray.end());
        C.lm_pos.insert(C.lm_pos.end(), L.lm_pos.begin(), L.lm_pos.end());
        for (int4 c : L.chunk) C.chunk.push_back(make_int4(c.x + (int)act_base, c.y, c.z + (int)obs_base, c.w));
        C.chunk_soff.insert(C.chunk_soff.end(), L.chunk_soff.begin(), L.chunk_soff.end());
        for (GroupDesc g : L.grp_desc) {
            g.a_loc = g.anchor;
            g.anchor += (int)frame_base;
            g.prob = p;
            g.delta_off = vec_off;
            g.chunk_begin += (int)chunk_base;
            g.chunk_end += (int)chunk_base;
            g.lm_begin += (int)act_base;
            g.d_off += (int)d_base;
            g.v_off += v_base;
            g.D_off += D_base;
            g.chunk0.x += (int)act_base;
            g.chunk0.z += (int)obs_base;
            C.grp_desc.push_back(g);
        }
        // covering lists per pose row block, anchors in global pose numbering
        for (int j = 0; j < h->N; ++j) C.fs_begin.push_back(L.fs_begin[j] + fs_base);
        for (FsEntry f : L.fs_list) {
            f.anchor += (int)pose_base;
            f.d_off += (int)d_base;
            f.D_off += D_base;
            C.fs_list.push_back(f);
        }
        for (int j = 0; j < h->N; ++j) pose_prob.push_back(p);
        first_all.insert(first_all.end(), fb.begin(), fb.end());

        if (L.d_total + d_base > INT32_MAX || (long long)C.ob_t.size() > INT32_MAX)
            return b->fail(SPGO_E_LAYOUT, "batch too large for 32-bit offsets");
        pose_base += h->N;
        frame_base += h->F;
        lm_base += h->L;
        act_base += (long long)L.lm_id.size();
        obs_base += (long long)L.ob_t.size();
        chunk_base += (long long)L.chunk.size();
        v_base += L.v_total;
        D_base += L.D_total;
        d_base += L.d_total;
        g_base += (int)L.grp_desc.size();
        fs_base += (int)L.fs_list.size();
        H_off += (long long)Pp * Pp;
        vec_off += Pp;
        winv_off += (long long)Pp * chol_band_block();
        Lp_off += (long long)chol_panel_doubles(Pp);
    }
    C.fs_begin.push_back(fs_base);
    b->tot_poses = pose_base;
    b->tot_frames = frame_base;
    b->tot_slots = lm_base;
    b->tot_act = act_base;
    b->tot_meas = obs_base;
    b->tot_groups = g_base;
    if (pose_base > INT32_MAX / 6 || lm_base > INT32_MAX) return b->fail(SPGO_E_LAYOUT, "batch too large");

    Layout& lay = b->lay;
    lay = Layout{};
    lay.N = (int)pose_base;
    lay.F = (int)frame_base;
    lay.L = (int)lm_base;
    lay.n_act = (int)act_base;
    lay.n_groups = g_base;
    lay.n_chunks = (int)chunk_base;
    lay.n_obs = (int)obs_base;
    lay.kmax = std::max(b->max_span, 1);
    lay.multi_chunk = 0;
    for (const GroupDesc& d : C.grp_desc) lay.multi_chunk |= d.chunk_end - d.chunk_begin > 1;
    // many short rows: one pose (6 rows sharing a covering list) per workgroup
    int max_pp = 0;
    for (const ProbDesc& d : b->pd) max_pp = std::max(max_pp, d.Pp);
    lay.form_rows = (6 * pose_base > 4096 && max_pp <= 512) ? 6 : 1;
    lay.n_probs = np;
    int rc;
    const ProbDesc* probs = nullptr;
    if ((rc = bupload(b, C.ob_t, &lay.ob_t)) || (rc = bupload(b, C.ob_meta, &lay.ob_meta)) ||
        (rc = bupload(b, C.ob_uv, &lay.ob_uv)) || (rc = bupload(b, C.lm_id, &lay.lm_id)) ||
        (rc = bupload(b, C.lm_ray, &lay.lm_ray)) || (rc = bupload(b, C.lm_pos, &lay.lm_pos)) ||
        (rc = bupload(b, C.chunk, &lay.chunk)) || (rc = bupload(b, C.chunk_soff, &lay.chunk_soff)) ||
        (rc = bupload(b, launch_order(C.grp_desc), &lay.grp_desc)) || (rc = bupload(b, C.fs_begin, &lay.fs_begin)) ||
        (rc = bupload(b, C.fs_list, &lay.fs_list)) || (rc = bupload(b, pose_prob, &lay.pose_prob)) ||
        (rc = bupload(b, first_all, &b->first)) || (rc = bupload(b, b->pd, &probs)))
        return rc;
    lay.probs = probs;
    b->pd_dev = const_cast<ProbDesc*>(probs);
    rc = b->prec == SPGO_PREC_MIXED ? alloc_state<float>(b, b->bf, v_base, D_base, d_base)
                                    : alloc_state<double>(b, b->bd, v_base, D_base, d_base);
    if (rc) return rc;
    b->H = b->mem.alloc<double>((size_t)H_off);
    b->gvec = b->mem.alloc<double>((size_t)vec_off);
    b->delta = b->mem.alloc<double>((size_t)vec_off);
    b->winv = b->mem.alloc<double>((size_t)winv_off);
    b->Lp = b->mem.alloc<double>((size_t)Lp_off);
    b->ctl = b->mem.alloc<Control>(np);
    if (!b->H || !b->gvec || !b->delta || !b->winv || !b->Lp || !b->ctl)
        return b->fail(SPGO_E_DEVICE, "device allocation failed");
    if (prepare_chol_band(b->Dmax, b->chol_lds) != 0)
        return b->fail(SPGO_E_DEVICE, "cannot reserve LDS for the envelope Cholesky");
    // reduced systems: zeros, identity on the padding rows beyond 6N
    launch_fill(b->H, 0.0, H_off, b->stream);
    for (int p = 0; p < np; ++p) launch_set_diag(b->H + b->pd[p].H_off, b->pd[p].Pp, 6 * b->pd[p].N, 1.0, b->stream);
    launch_fill(b->gvec, 0.0, vec_off, b->stream);
    launch_fill(b->delta, 0.0, vec_off, b->stream);
    std::vector<Control> c(np);
    for (int p = 0; p < np; ++p) {
        c[p] = Control{};
        c[p].chol_ok = 1;
        c[p].pcg_fail = 1;
        c[p].lambda = b->opt.mode == SPGO_MODE_REFERENCE ? 0.0 : b->opt.lambda0;
    }
    rc = bcheck(b, hipMemcpy(b->ctl, c.data(), np * sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    rc = push_states(b);
    if (rc) return rc;
    b->dirty = false;
    return SPGO_OK;
}

AsmArgs make_asm(const spgo_batch* b, int init) {
    AsmArgs a;
    a.lay = b->lay;
    a.cam = b->cam;
    a.rb.kind = b->opt.robust;
    a.rb.near_boost = b->opt.near_boost;
    a.rb.param = b->opt.robust_param;
    a.eps = b->opt.eps;
    a.alpha_min = b->opt.alpha_min;
    a.alpha_max = b->opt.alpha_max;
    a.ctl = b->ctl;
    a.init = init;
    a.delta = b->delta;
    return a;
}

template <typename T>
void launch_iteration(spgo_batch* b, StateBuf<T>* s) {
    hipStream_t st = b->stream;
    launch_form_system<T>(b->lay, s, b->ctl, b->H, b->gvec, st);
    launch_chol_band(b->pd_dev, b->lay.n_probs, b->Dmax, b->chol_lds, b->H, b->first, b->gvec, b->delta, b->winv,
                     b->Lp, b->ctl, st);
    launch_pose_update<T>(b->lay, b->cam, s, b->ctl, b->delta, 1.0, 0, st);
    launch_assemble<T>(make_asm(b, 0), s, st);
    launch_accept(s[0].cost, s[1].cost, b->lay, b->ctl, b->rec, b->opt.mode == SPGO_MODE_REFERENCE, b->opt.lam_up,
                  b->opt.lam_down, b->opt.lam_min, b->opt.lam_max, 0, st);
}

template <typename T>
void launch_linearise(spgo_batch* b, StateBuf<T>* s) {
    launch_pose_update<T>(b->lay, b->cam, s, b->ctl, b->delta, 0.0, 1, b->stream);
    launch_assemble<T>(make_asm(b, 1), s, b->stream);
    launch_accept(s[0].cost, s[1].cost, b->lay, b->ctl, b->rec, 0, 0, 0, 0, 0, 1, b->stream);
}

// Graph slot k holds `iters` consecutive iterations (captured again when the
// count differs); iters == 0 keeps what the slot holds for a later call.
int capture_iterations(spgo_batch* b, int k, int iters, bool mixed) {
    if (iters == 0 || (iters == b->graph_iters[k] && b->gexec[k])) return SPGO_OK;
    destroy_graph_slot(b, k);
    int rc = bcheck(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) return rc;
    for (int it = 0; it < iters; ++it) {
        if (mixed)
            launch_iteration<float>(b, b->bf);
        else
            launch_iteration<double>(b, b->bd);
    }
    rc = bcheck(b, hipStreamEndCapture(b->stream, &b->graph[k]), "hipStreamEndCapture");
    if (rc) return rc;
    rc = bcheck(b, hipGraphInstantiate(&b->gexec[k], b->graph[k], nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) return rc;
    b->graph_iters[k] = iters;
    return SPGO_OK;
}

int ensure_records(spgo_batch* b, int iterations) {
    if (iterations <= b->rec_cap) return SPGO_OK;
    destroy_graph(b);
    const int cap = std::max(iterations, 16);
    const int np = (int)b->pr.size();
    b->rec = b->mem.alloc<IterRecord>((size_t)cap * np);
    if (!b->rec) return b->fail(SPGO_E_DEVICE, "device allocation failed");
    b->rec_cap = cap;
    for (int p = 0; p < np; ++p) b->pd[p].rec_off = p * cap;
    return bcheck(b, hipMemcpy(b->pd_dev, b->pd.data(), np * sizeof(ProbDesc), hipMemcpyHostToDevice), "hipMemcpy");
}

int run_step(spgo_batch* b, int iterations) {
    int rc = ensure_device(b);
    if (rc) return rc;
    if (b->dirty || b->prec != b->opt.precision || b->chunk_lm != b->opt.chunk_landmarks ||
        b->group_chunks != b->opt.group_chunks) {
        rc = setup(b);
        if (rc) return rc;
    }
    rc = ensure_records(b, iterations);
    if (rc) return rc;
    const int np = (int)b->pr.size();
    const bool mixed = b->prec == SPGO_PREC_MIXED;
    std::vector<Control> c(np);
    rc = bcheck(b, hipMemcpy(c.data(), b->ctl, np * sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    for (Control& x : c) {
        x.iter = 0;
        x.chol_ok = 1;
        x.cost0 = x.cost;
    }
    rc = bcheck(b, hipMemcpy(b->ctl, c.data(), np * sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    if (!b->lin_valid) {
        if (mixed)
            launch_linearise<float>(b, b->bf);
        else
            launch_linearise<double>(b, b->bd);
        b->lin_valid = true;
    }
    rc = bcheck(b, hipGetLastError(), "kernel launch");
    if (rc) return rc;
    // up to GRAPH_UNROLL iterations per graph launch, the remainder in a second graph
    const bool graph = b->opt.use_graph && iterations > 0;
    const int unroll = std::min(iterations, GRAPH_UNROLL);
    const int n_full = graph ? iterations / unroll : 0;
    const int n_rem = graph ? iterations % unroll : 0;
    if (graph && ((rc = capture_iterations(b, 0, unroll, mixed)) || (rc = capture_iterations(b, 1, n_rem, mixed))))
        return rc;
    rc = bcheck(b, hipEventRecord(b->ev0, b->stream), "hipEventRecord");
    if (rc) return rc;
    if (graph) {
        for (int it = 0; it < n_full && !rc; ++it)
            rc = bcheck(b, hipGraphLaunch(b->gexec[0], b->stream), "hipGraphLaunch");
        if (!rc && n_rem > 0) rc = bcheck(b, hipGraphLaunch(b->gexec[1], b->stream), "hipGraphLaunch");
        if (rc) return rc;
    } else {
        for (int it = 0; it < iterations; ++it) {
            if (mixed)
                launch_iteration<float>(b, b->bf);
            else
                launch_iteration<double>(b, b->bd);
        }
    }
    rc = bcheck(b, hipEventRecord(b->ev1, b->stream), "hipEventRecord");
    if (rc) return rc;
    rc = bcheck(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (rc) return rc;
    rc = bcheck(b, hipGetLastError(), "kernel execution");
    if (rc) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, b->ev0, b->ev1);
    b->last_ms = ms;
    rc = bcheck(b, hipMemcpy(c.data(), b->ctl, np * sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    b->last_iters = iterations;
    b->cost0.resize(np);
    b->last_rec.assign((size_t)np * iterations, IterRecord{});
    if (iterations > 0) {
        std::vector<IterRecord> all((size_t)np * b->rec_cap);
        rc = bcheck(b, hipMemcpy(all.data(), b->rec, all.size() * sizeof(IterRecord), hipMemcpyDeviceToHost),
                    "hipMemcpy");
        if (rc) return rc;
        for (int p = 0; p < np; ++p)
            std::copy(all.begin() + (size_t)p * b->rec_cap, all.begin() + (size_t)p * b->rec_cap + iterations,
                      b->last_rec.begin() + (size_t)p * iterations);
    }
    for (int p = 0; p < np; ++p) b->cost0[p] = c[p].cost0;
    return SPGO_OK;
}

// State of problem p: from the batch's device buffers once they exist, else
// from the problem's own solver.
int get_state(spgo_batch* b, int p, double* poses, double* alpha) {
    spgo_solver* h = b->pr[p];
    if (b->dirty || !b->ctl) return spgo_get_state_f64(h, poses, alpha);
    const ProbDesc& d = b->pd[p];
    Control c;
    int rc = bcheck(b, hipMemcpy(&c, b->ctl + p, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    const bool mixed = b->prec == SPGO_PREC_MIXED;
    const double* X = mixed ? b->bf[c.cur].X : b->bd[c.cur].X;
    const double* al = mixed ? b->bf[c.cur].alpha : b->bd[c.cur].alpha;
    if (poses) {
        std::vector<double> x((size_t)12 * h->N);
        rc = bcheck(b, hipMemcpy(x.data(), X + 12LL * d.pose_base, x.size() * sizeof(double), hipMemcpyDeviceToHost),
                    "hipMemcpy");
        if (rc) return rc;
        for (int k = 0; k < h->N; ++k) {
            for (int i = 0; i < 12; ++i) poses[16 * k + i] = x[12 * (size_t)k + i];
            poses[16 * k + 12] = poses[16 * k + 13] = poses[16 * k + 14] = 0.0;
            poses[16 * k + 15] = 1.0;
        }
    }
    if (alpha) {
        rc = bcheck(b, hipMemcpy(alpha, al + d.lm_base, (size_t)h->L * sizeof(double), hipMemcpyDeviceToHost),
                    "hipMemcpy");
        if (rc) return rc;
    }
    return SPGO_OK;
}

}  // namespace

extern "C" {

int spgo_batch_create(int32_t n_problems, spgo_solver* const* problems, const spgo_options* opts, spgo_batch** out) {
    if (!out) return SPGO_E_ARG;
    *out = nullptr;
    if (n_problems <= 0 || !problems) {
        g_batch_create_error = "spgo_batch_create: need at least one problem";
        return SPGO_E_ARG;
    }
    spgo_options o;
    spgo_default_options(&o);
    if (opts) o = *opts;
    std::string why;
    if (!valid_opts(&o, why)) {
        g_batch_create_error = "spgo_batch_create: invalid option " + why;
        return SPGO_E_ARG;
    }
    for (int p = 0; p < n_problems; ++p) {
        const spgo_solver* h = problems[p];
        if (!h) {
            g_batch_create_error = "spgo_batch_create: problem " + std::to_string(p) + " is NULL";
            return SPGO_E_ARG;
        }
        for (int q = 0; q < p; ++q)
            if (problems[q] == h) {
                g_batch_create_error = "spgo_batch_create: problem " + std::to_string(p) + " repeats problem " +
                                       std::to_string(q);
                return SPGO_E_ARG;
            }
        if (std::memcmp(h->intr, problems[0]->intr, sizeof(h->intr)) != 0 ||
            std::memcmp(h->Trl, problems[0]->Trl, sizeof(h->Trl)) != 0) {
            g_batch_create_error = "spgo_batch_create: problem " + std::to_string(p) +
                                   " has a different stereo calibration from problem 0";
            return SPGO_E_ARG;
        }
    }
    spgo_batch* b = new spgo_batch();
    b->pr.assign(problems, problems + n_problems);
    b->opt = o;
    *out = b;
    return SPGO_OK;
}

int spgo_batch_destroy(spgo_batch* b) {
    if (!b) return SPGO_OK;
    destroy_graph(b);
    if (b->dev_ready) (void)hipStreamSynchronize(b->stream);
    b->mem.release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SPGO_OK;
}

int spgo_batch_set_options(spgo_batch* b, const spgo_options* opts) {
    if (!b || !opts) return SPGO_E_ARG;
    std::string why;
    if (!valid_opts(opts, why)) return b->fail(SPGO_E_ARG, "invalid option " + why);
    const bool rebuild = opts->precision != b->opt.precision || opts->chunk_landmarks != b->opt.chunk_landmarks ||
                         opts->group_chunks != b->opt.group_chunks;
    b->opt = *opts;
    destroy_graph(b);   // the captured launches carry the options
    if (rebuild) b->dirty = true;
    b->lin_valid = false;
    return SPGO_OK;
}

int spgo_batch_reload(spgo_batch* b) {
    if (!b) return SPGO_E_ARG;
    b->dirty = true;
    return SPGO_OK;
}

int spgo_batch_reset_states(spgo_batch* b) {
    if (!b) return SPGO_E_ARG;
    if (b->dirty || !b->ctl) return SPGO_OK;   // the next step reads them anyway
    const int np = (int)b->pr.size();
    std::vector<Control> c(np);
    int rc = bcheck(b, hipMemcpy(c.data(), b->ctl, np * sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    for (Control& x : c) x.lambda = b->opt.mode == SPGO_MODE_REFERENCE ? 0.0 : b->opt.lambda0;
    rc = bcheck(b, hipMemcpy(b->ctl, c.data(), np * sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    return push_states(b);
}

int spgo_batch_step(spgo_batch* b, int32_t iterations) {
    if (!b || iterations < 0) return SPGO_E_ARG;
    for (size_t p = 0; p < b->pr.size(); ++p)
        if (std::memcmp(b->pr[p]->intr, b->pr[0]->intr, sizeof(b->pr[0]->intr)) != 0 ||
            std::memcmp(b->pr[p]->Trl, b->pr[0]->Trl, sizeof(b->pr[0]->Trl)) != 0)
            return b->fail(SPGO_E_ARG, "problem " + std::to_string(p) +
                                           " has a different stereo calibration from problem 0");
    return run_step(b, iterations);
}

int spgo_batch_get_state_f64(spgo_batch* b, int32_t problem, double* poses, double* alpha) {
    if (!b || problem < 0 || problem >= (int)b->pr.size()) return SPGO_E_ARG;
    return get_state(b, problem, poses, alpha);
}

int spgo_batch_store(spgo_batch* b) {
    if (!b) return SPGO_E_ARG;
    if (b->dirty || !b->ctl) return SPGO_OK;
    for (size_t p = 0; p < b->pr.size(); ++p) {
        spgo_solver* h = b->pr[p];
        std::vector<double> poses((size_t)16 * h->N), alpha((size_t)h->L);
        int rc = get_state(b, (int)p, poses.data(), alpha.data());
        if (rc) return rc;
        rc = spgo_set_state_f64(h, poses.data(), alpha.data());
        if (rc) return b->fail(rc, "problem " + std::to_string(p) + ": " + h->err);
    }
    return SPGO_OK;
}

int spgo_batch_get_stats(spgo_batch* b, int32_t problem, spgo_iter_stats* out, int32_t max_records, int32_t* n_out,
                         double* initial_cost) {
    if (!b || problem < 0 || problem >= (int)b->pr.size()) return SPGO_E_ARG;
    const int n = b->cost0.empty() ? 0 : b->last_iters;
    if (n_out) *n_out = n;
    if (initial_cost) *initial_cost = b->cost0.empty() ? 0.0 : b->cost0[problem];
    if (out)
        for (int i = 0; i < n && i < max_records; ++i) {
            const IterRecord& r = b->last_rec[(size_t)problem * n + i];
            out[i].cost_candidate = r.cost_candidate;
            out[i].cost = r.cost;
            out[i].lambda = r.lambda;
            out[i].accepted = r.accepted;
            out[i].chol_ok = r.chol_ok;
        }
    return SPGO_OK;
}

int spgo_batch_last_step_ms(spgo_batch* b, double* ms) {
    if (!b || !ms) return SPGO_E_ARG;
    *ms = b->last_ms;
    return SPGO_OK;
}

int spgo_batch_get_info(spgo_batch* b, spgo_batch_info* out) {
    if (!b || !out) return SPGO_E_ARG;
    std::memset(out, 0, sizeof(*out));
    out->n_problems = (int)b->pr.size();
    if (!b->dirty) {
        out->n_poses = (int)b->tot_poses;
        out->n_landmarks = (int)b->tot_act;
        out->n_measurements = (int)b->tot_meas;
        out->n_groups = b->tot_groups;
        out->max_depth = b->Dmax;
        out->max_span = b->max_span;
    }
    out->chunk_landmarks = b->opt.chunk_landmarks;
    out->group_chunks = b->opt.group_chunks;
    return SPGO_OK;
}

// Tile-shape search: for each candidate (chunk_landmarks, group_chunks) the
// batch is rebuilt from the problems' current states and `iterations` LM
// iterations are timed on the device; the fastest shape is kept and the batch
// is left rebuilt from the problems' states, as if reloaded.
int spgo_batch_autotune(spgo_batch* b, int32_t iterations, int32_t* best_chunk_landmarks, int32_t* best_group_chunks,
                        double* best_ms) {
    if (!b || iterations <= 0) return SPGO_E_ARG;
    static const int cand_chunk[] = {16, 32, 48, 64};
    static const int cand_group[] = {1, 2, 4};
    const spgo_options keep = b->opt;
    double best = 1e300;
    int bc = keep.chunk_landmarks, bg = keep.group_chunks;
    for (int cl : cand_chunk)
        for (int gc : cand_group) {
            spgo_options o = keep;
            o.chunk_landmarks = cl;
            o.group_chunks = gc;
            b->opt = o;
            b->dirty = true;
            destroy_graph(b);
            int rc = run_step(b, iterations);
            if (rc == SPGO_E_LAYOUT) continue;   // shape not representable for some problem
            if (rc) {
                b->opt = keep;
                b->dirty = true;
                return rc;
            }
            if (b->last_ms < best) {
                best = b->last_ms;
                bc = cl;
                bg = gc;
            }
        }
    b->opt = keep;
    b->opt.chunk_landmarks = bc;
    b->opt.group_chunks = bg;
    b->dirty = true;
    destroy_graph(b);
    if (best_chunk_landmarks) *best_chunk_landmarks = bc;
    if (best_group_chunks) *best_group_chunks = bg;
    if (best_ms) *best_ms = best < 1e300 ? best : 0.0;
    if (best >= 1e300) return b->fail(SPGO_E_LAYOUT, "no candidate tile shape fits every problem");
    return SPGO_OK;
}

int spgo_batch_profile_phases(spgo_batch* b, int32_t iterations, double ms[SPGO_N_PHASES]) {
    if (!b || !ms || iterations < 1) return SPGO_E_ARG;
    int rc = run_step(b, 0);
    if (rc) return rc;
    rc = ensure_records(b, iterations);
    if (rc) return rc;
    hipEvent_t ev[SPGO_N_PHASES + 1];
    for (int i = 0; i <= SPGO_N_PHASES; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) return b->fail(SPGO_E_DEVICE, "hipEventCreate failed");
    double acc[SPGO_N_PHASES] = {0};
    hipStream_t s = b->stream;
    const bool mixed = b->prec == SPGO_PREC_MIXED;
    const bool refm = b->opt.mode == SPGO_MODE_REFERENCE;
    for (int it = 0; it < iterations && rc == SPGO_OK; ++it) {
        (void)hipEventRecord(ev[0], s);
        if (mixed) launch_form_system<float>(b->lay, b->bf, b->ctl, b->H, b->gvec, s);
        else launch_form_system<double>(b->lay, b->bd, b->ctl, b->H, b->gvec, s);
        (void)hipEventRecord(ev[1], s);
        launch_chol_band(b->pd_dev, b->lay.n_probs, b->Dmax, b->chol_lds, b->H, b->first, b->gvec, b->delta,
                         b->winv, b->Lp, b->ctl, s);
        (void)hipEventRecord(ev[2], s);
        (void)hipEventRecord(ev[3], s);   // the envelope kernel also solves
        if (mixed) launch_pose_update<float>(b->lay, b->cam, b->bf, b->ctl, b->delta, 1.0, 0, s);
        else launch_pose_update<double>(b->lay, b->cam, b->bd, b->ctl, b->delta, 1.0, 0, s);
        (void)hipEventRecord(ev[4], s);
        if (mixed) launch_assemble<float>(make_asm(b, 0), b->bf, s);
        else launch_assemble<double>(make_asm(b, 0), b->bd, s);
        (void)hipEventRecord(ev[5], s);
        if (mixed)
            launch_accept(b->bf[0].cost, b->bf[1].cost, b->lay, b->ctl, b->rec, refm, b->opt.lam_up,
                          b->opt.lam_down, b->opt.lam_min, b->opt.lam_max, 0, s);
        else
            launch_accept(b->bd[0].cost, b->bd[1].cost, b->lay, b->ctl, b->rec, refm, b->opt.lam_up,
                          b->opt.lam_down, b->opt.lam_min, b->opt.lam_max, 0, s);
        (void)hipEventRecord(ev[6], s);
        rc = bcheck(b, hipEventSynchronize(ev[6]), "phase timing");
        for (int p = 0; p < SPGO_N_PHASES && rc == SPGO_OK; ++p) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ev[p], ev[p + 1]);
            acc[p] += t;
        }
    }
    for (int i = 0; i <= SPGO_N_PHASES; ++i) (void)hipEventDestroy(ev[i]);
    if (rc) return rc;
    for (int p = 0; p < SPGO_N_PHASES; ++p) ms[p] = acc[p] / iterations;
    return SPGO_OK;
}

int spgo_batch_get_options(spgo_batch* b, spgo_options* out) {
    if (!b || !out) return SPGO_E_ARG;
    *out = b->opt;
    return SPGO_OK;
}

const char* spgo_batch_last_error(spgo_batch* b) {
    if (!b) return g_batch_create_error.c_str();
    return b->err.c_str();
}

}  // extern "C"
