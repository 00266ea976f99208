// Host runtime behind the C ABI of include/spgo.h.
//
// Replaces the reference's C++ facade (include/Solver.hpp:18-85,
// src/Optimization/Solver.cpp:1-128) and the orchestration of updateState
// (src/Optimization/Solver_cuda.cu:557-795):
//   * measurements are kept host-side and laid out for the device once, on the
//     first spgo_step after they change (grouped by anchor frame, chunked for
//     LDS, structure-of-arrays);
//   * device buffers persist across steps (the reference allocates and frees
//     every buffer on every step, LMutils.cu:5-185);
//   * one LM iteration is a fixed launch sequence that reads its buffer
//     selection and damping from a device control block, so it is captured
//     once in a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/spgo.h"
#include "spgo_host.h"
#include "spgo_internal.h"
#include "spgo_kernels.h"

using namespace spgo;

namespace {

thread_local std::string g_create_error;

}  // namespace

namespace spgo {

int check_hip(spgo_solver* h, hipError_t e, const char* what) {
    if (e == hipSuccess) return SPGO_OK;
    return h->fail(SPGO_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// G_0 = I, G_{k+1} = G_k X_k^-1 (3x4 poses, F = N + 1 of them)
void global_chain(int N, const double* X, double* G) {
    double P[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    std::memcpy(G, P, sizeof(P));
    for (int k = 0; k < N; ++k) {
        const double* A = X + 12 * k;
        double Ai[12], C[12];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Ai[4 * r + c] = A[4 * c + r];
        for (int r = 0; r < 3; ++r)
            Ai[4 * r + 3] = -(Ai[4 * r] * A[3] + Ai[4 * r + 1] * A[7] + Ai[4 * r + 2] * A[11]);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) {
                double v = P[4 * r] * Ai[c] + P[4 * r + 1] * Ai[4 + c] + P[4 * r + 2] * Ai[8 + c];
                if (c == 3) v += P[4 * r + 3];
                C[4 * r + c] = v;
            }
        std::memcpy(P, C, sizeof(P));
        std::memcpy(G + 12 * (k + 1), P, sizeof(P));
    }
}

void mat4_to_3x4(const float* T16, double* X12) {
    for (int i = 0; i < 12; ++i) X12[i] = (double)T16[i];
}

// Replace the rotation part of a 3x4 pose by its polar factor (the nearest
// rotation), in double: R <- (R + R^-T) / 2 until it stops changing.  The
// anchor-frame factorisation of the pose Jacobian relies on Ad(A B) =
// Ad(A) Ad(B), which holds for rotations only; float32 inputs are rotations to
// about 1e-7.  Matrices that are singular or reflections are left unchanged.
void project_rotation(double* P) {
    double R[9] = {P[0], P[1], P[2], P[4], P[5], P[6], P[8], P[9], P[10]};
    for (int it = 0; it < 30; ++it) {
        double C[9] = {R[4] * R[8] - R[5] * R[7], R[5] * R[6] - R[3] * R[8], R[3] * R[7] - R[4] * R[6],
                       R[2] * R[7] - R[1] * R[8], R[0] * R[8] - R[2] * R[6], R[1] * R[6] - R[0] * R[7],
                       R[1] * R[5] - R[2] * R[4], R[2] * R[3] - R[0] * R[5], R[0] * R[4] - R[1] * R[3]};
        const double det = R[0] * C[0] + R[1] * C[1] + R[2] * C[2];
        if (!(det > 1e-6) || !std::isfinite(det)) return;
        double change = 0.0;
        for (int i = 0; i < 9; ++i) {
            const double n = 0.5 * (R[i] + C[i] / det);   // C / det = R^-T
            change = std::max(change, std::fabs(n - R[i]));
            R[i] = n;
        }
        if (change < 1e-17) break;
    }
    P[0] = R[0], P[1] = R[1], P[2] = R[2];
    P[4] = R[3], P[5] = R[4], P[6] = R[5];
    P[8] = R[6], P[9] = R[7], P[10] = R[8];
}

static void destroy_graph_slot(spgo_solver* h, int k) {
    if (h->gexec[k]) (void)hipGraphExecDestroy(h->gexec[k]);
    if (h->graph[k]) (void)hipGraphDestroy(h->graph[k]);
    h->gexec[k] = nullptr;
    h->graph[k] = nullptr;
    h->graph_iters[k] = 0;
}

void destroy_graph(spgo_solver* h) {
    destroy_graph_slot(h, 0);
    destroy_graph_slot(h, 1);
    h->graph_dirty = true;
}

// Chunks per group when the options leave the choice to the library
// (group_chunks 0).  The assembly runs one workgroup per group, so a launch
// takes ceil(groups / resident workgroups) rounds, each as long as its largest
// group: about (chunks + 0.4) chunk times, the 0.4 for the group's set-up and
// write-back.  The choice minimises that product over 1..8 chunks per group,
// from the one-chunk layout's chunk count per anchor.
static int auto_group_chunks(const spgo_solver* h, const HostLayout& one) {
    const int F = (int)one.anchor_grp.size() - 1;
    int best = 1;
    double best_t = 0.0;
    for (int gc = 1; gc <= 8; ++gc) {
        long long groups = 0;
        int largest = 0;
        for (int a = 0; a < F; ++a) {
            const int n = one.anchor_grp[a + 1] - one.anchor_grp[a];
            groups += (n + gc - 1) / gc;
            largest = std::max(largest, std::min(n, gc));
        }
        const int slots = assemble_resident_blocks(one.kmax, h->opt.precision == SPGO_PREC_MIXED, gc > 1 && largest > 1);
        if (slots <= 0) return 1;   // no device to ask: one chunk per group
        const double t = (double)((groups + slots - 1) / slots) * (largest + 0.4);
        if (gc == 1 || t < best_t) {
            best = gc;
            best_t = t;
        }
        if (largest < gc) break;   // larger groups change nothing
    }
    return best;
}

int build_layout(spgo_solver* h, HostLayout& out) {
    if (h->opt.group_chunks > 0) return build_layout(h, h->opt.chunk_landmarks, h->opt.group_chunks, out);
    int rc = build_layout(h, h->opt.chunk_landmarks, 1, out);
    if (rc != SPGO_OK) return rc;
    const int gc = auto_group_chunks(h, out);
    if (gc == 1) return SPGO_OK;
    return build_layout(h, h->opt.chunk_landmarks, gc, out);
}

// Build the device layout from the measurement list.
int build_layout(spgo_solver* h, int chunk_landmarks, int group_chunks_opt, HostLayout& out) {
    const int O = (int)h->ma.size();
    const int chunk_lm = std::max(1, std::min(chunk_landmarks > 0 ? chunk_landmarks : CHUNK_LM, CHUNK_LM));
    const int group_chunks = std::max(1, group_chunks_opt > 0 ? group_chunks_opt : 1);
    HostLayout hl;
    // measurements of each landmark, targets ascending (stable)
    std::vector<int> order(O);
    for (int i = 0; i < O; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (h->ml[x] != h->ml[y]) return h->ml[x] < h->ml[y];
        return h->mt[x] < h->mt[y];
    });
    struct Lm {
        int id, anchor, begin, count, smax;
    };
    std::vector<Lm> lms;
    for (int i = 0; i < O;) {
        const int l = h->ml[order[i]];
        int j = i;
        const int a = h->ma[order[i]];
        int smax = 0;
        while (j < O && h->ml[order[j]] == l) {
            if (h->ma[order[j]] != a)
                return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) +
                                                  " has measurements with different anchor frames");
            smax = std::max(smax, h->mt[order[j]] - a);
            ++j;
        }
        if (smax > KMAX)
            return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) + " spans " + std::to_string(smax) +
                                              " poses; this build supports at most " + std::to_string(KMAX));
        if (j - i > CHUNK_OBS)
            return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) + " has more than " +
                                              std::to_string(CHUNK_OBS) + " measurements");
        lms.push_back({l, a, i, j - i, smax});
        i = j;
    }
    std::stable_sort(lms.begin(), lms.end(), [](const Lm& x, const Lm& y) {
        if (x.anchor != y.anchor) return x.anchor < y.anchor;
        return x.id < y.id;
    });
    const int n_act = (int)lms.size();
    hl.lm_id.resize(n_act);
    hl.lm_ray.resize(n_act);
    hl.lm_pos.resize(n_act);
    hl.ob_t.reserve(O);
    hl.ob_meta.reserve(O);
    hl.ob_uv.reserve(O);
    hl.anchor_grp.assign(h->F + 1, 0);

    std::vector<int> grp_of_anchor_count(h->F, 0);
    int li = 0;
    while (li < n_act) {
        // one group: landmarks of one anchor, up to group_chunks chunks
        const int a = lms[li].anchor;
        const int g = (int)hl.grp_anchor.size();
        hl.grp_anchor.push_back(a);
        hl.grp_lm_begin.push_back(li);
        const int chunk_begin = (int)hl.chunk.size();
        int K = 0;
        int nch = 0;
        while (li < n_act && lms[li].anchor == a && nch < group_chunks) {
            // one chunk
            const int lm0 = li;
            int nobs = 0;
            while (li < n_act && lms[li].anchor == a && li - lm0 < chunk_lm && nobs + lms[li].count <= CHUNK_OBS) {
                nobs += lms[li].count;
                ++li;
            }
            const int nl = li - lm0;
            const int ob0 = (int)hl.ob_t.size();
            // landmark-major positions, then slot-major processing order
            struct Ob {
                int slot, j, rank, m, pos;
            };
            std::vector<Ob> obs;
            obs.reserve(nobs);
            int pos = 0;
            for (int j = 0; j < nl; ++j) {
                const Lm& L = lms[lm0 + j];
                hl.lm_id[lm0 + j] = L.id;
                hl.lm_pos[lm0 + j] = make_int2(pos, L.count);
                K = std::max(K, L.smax);
                for (int r = 0; r < L.count; ++r) {
                    const int m = order[L.begin + r];
                    const int slot = std::max(h->mt[m] - a, 0);
                    obs.push_back({slot, j, r, m, pos});
                    ++pos;
                }
                // ray from the anchor cell's left pixel (Solver_cuda.cu:147); zero if never written
                auto it = h->cell_of.find((long long)a * h->L + L.id);
                if (it != h->cell_of.end())
                    hl.lm_ray[lm0 + j] = make_float2(h->cells[6 * it->second], h->cells[6 * it->second + 1]);
                else
                    hl.lm_ray[lm0 + j] = make_float2(0.f, 0.f);
            }
            std::stable_sort(obs.begin(), obs.end(), [](const Ob& x, const Ob& y) {
                if (x.slot != y.slot) return x.slot < y.slot;
                return x.j < y.j;
            });
            {
                // slot offsets inside the chunk (slot-major order), padded to KMAX + 2 entries
                const size_t base = hl.chunk_soff.size();
                hl.chunk_soff.resize(base + KMAX + 2, nobs);
                int q = 0;
                for (int sl = 0; sl <= KMAX + 1; ++sl) {
                    while (q < nobs && obs[q].slot < sl) ++q;
                    hl.chunk_soff[base + sl] = q;
                }
            }
            for (const Ob& o : obs) {
                const int m = o.m;
                hl.ob_t.push_back(h->mt[m]);
                hl.ob_meta.push_back((uint32_t)o.pos | ((uint32_t)o.j << 16));
                const int cell = h->cell_of.at((long long)h->mt[m] * h->L + h->ml[m]);
                const float* c = &h->cells[6 * cell];
                hl.ob_uv.push_back(make_float4(c[0], c[1], c[3], c[4]));
            }
            hl.chunk.push_back(make_int4(lm0, nl, ob0, nobs));
            ++nch;
        }
        hl.grp_chunks.push_back(make_int2(chunk_begin, (int)hl.chunk.size()));
        hl.grp_K.push_back(K);
        hl.kmax = std::max(hl.kmax, K);
        hl.grp_v_off.push_back(hl.v_total);
        hl.grp_D_off.push_back(hl.D_total);
        hl.grp_d_off.push_back((int)hl.d_total);
        const long long nlg = li - hl.grp_lm_begin[g];
        hl.v_total += nlg * 6 * K;
        hl.D_total += (long long)(6 * K) * (6 * K);
        hl.d_total += 6 * K;
        grp_of_anchor_count[a] += 1;
    }
    for (int a = 0; a < h->F; ++a) {
        hl.anchor_grp[a + 1] = hl.anchor_grp[a] + grp_of_anchor_count[a];
        if (grp_of_anchor_count[a]) hl.anchors_used++;
    }
    hl.kmax = std::max(hl.kmax, 1);
    hl.grp_desc.resize(hl.grp_anchor.size());
    for (size_t g = 0; g < hl.grp_anchor.size(); ++g) {
        GroupDesc& d = hl.grp_desc[g];
        d = GroupDesc{};
        d.anchor = hl.grp_anchor[g];
        d.a_loc = hl.grp_anchor[g];
        d.prob = 0;
        d.delta_off = 0;
        d.K = hl.grp_K[g];
        d.chunk_begin = hl.grp_chunks[g].x;
        d.chunk_end = hl.grp_chunks[g].y;
        d.lm_begin = hl.grp_lm_begin[g];
        d.d_off = hl.grp_d_off[g];
        d.v_off = hl.grp_v_off[g];
        d.D_off = hl.grp_D_off[g];
        d.chunk0 = hl.chunk[d.chunk_begin];
        for (int sl = 0; sl < KMAX + 2; ++sl) d.soff0[sl] = hl.chunk_soff[(size_t)d.chunk_begin * (KMAX + 2) + sl];
    }
    // covering groups of every pose row block, anchors ascending
    hl.fs_begin.assign(h->N + 1, 0);
    for (int bi = 0; bi < h->N; ++bi) {
        hl.fs_begin[bi] = (int)hl.fs_list.size();
        for (int a = std::max(0, bi - hl.kmax + 1); a <= bi; ++a)
            for (int g = hl.anchor_grp[a]; g < hl.anchor_grp[a + 1]; ++g)
                if (bi - a < hl.grp_K[g])
                    hl.fs_list.push_back(FsEntry{a, 6 * hl.grp_K[g], hl.grp_d_off[g], 0, hl.grp_D_off[g]});
    }
    hl.fs_begin[h->N] = (int)hl.fs_list.size();
    // pose j couples back to the smallest anchor among the groups covering it
    hl.fpose.resize(h->N);
    for (int j = 0; j < h->N; ++j) hl.fpose[j] = j;
    for (size_t g = 0; g < hl.grp_anchor.size(); ++g)
        for (int k = 0; k < hl.grp_K[g]; ++k) {
            const int j = hl.grp_anchor[g] + k;
            if (j < h->N) hl.fpose[j] = std::min(hl.fpose[j], hl.grp_anchor[g]);
        }
    for (int j = 0; j < h->N; ++j) hl.band_need = std::max(hl.band_need, j - hl.fpose[j]);
    out = std::move(hl);
    return SPGO_OK;
}

// Size of the reduced system with padding: a whole number of blocks for
// either factorisation.
int reduced_size(int n_poses) { return ((6 * n_poses + NB - 1) / NB) * NB; }

// Launch order of the assembly groups.  The device hands workgroup i to XCD
// i mod nx (round-robin dispatch), so the groups are cut into nx contiguous
// ranges of nearly equal size, which are interleaved: each range (neighbouring
// anchors; whole windows of a batch) runs on one XCD, whose L2 then serves the
// poses and updates its groups share instead of every XCD fetching them.  The
// kernels write their cost partial at gid, the group's place in the layout,
// so the results do not depend on the order.
std::vector<GroupDesc> launch_order(const std::vector<GroupDesc>& in) {
    int dev = 0, cus = 0;
    int nx = 1;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        nx = cus % 32 == 0 ? cus / 32 : (cus % 38 == 0 ? cus / 38 : 1);   // 32 CUs per XCD on gfx950
    const int G = (int)in.size();
    std::vector<GroupDesc> out(in);
    for (int g = 0; g < G; ++g) out[g].gid = g;
    if (nx <= 1 || G < 2 * nx) return out;
    // range x holds base + (x < rem) groups from start(x); workgroup i = nx s + x
    // takes element s of range x (the ranges that are one longer come first)
    const int base = G / nx, rem = G % nx;
    for (int i = 0; i < G; ++i) {
        const int x = i % nx, s = i / nx;
        const int src = x * base + std::min(x, rem) + s;
        out[i] = in[src];
        out[i].gid = src;
    }
    return out;
}

// Block profile of the reduced system for the envelope factorisation: pose j
// couples back to the smallest anchor among the landmark groups covering it,
// fb[i] is the first nonzero block column of block row i and depth the
// largest i - fb[i].  Returns the factorisation path (SPGO_FACTOR_DENSE,
// _ENVELOPE or _ENVELOPE_SPLIT), or -1 if the requested one cannot be used.
int profile_blocks(int n_poses, const HostLayout& L, std::vector<int>& fb) {
    const int P = 6 * n_poses, Pp = reduced_size(n_poses), BBk = chol_band_block();
    const std::vector<int>& fpose = L.fpose;
    const int nbk = Pp / BBk;
    fb.assign(nbk, 0);
    int depth = 0;
    for (int i = 0; i < nbk; ++i) {
        int f = i;
        for (int r = i * BBk; r < (i + 1) * BBk && r < P; ++r) f = std::min(f, (6 * fpose[r / 6]) / BBk);
        fb[i] = f;
        depth = std::max(depth, i - f);
    }
    return depth;
}

// The incremental kernels stage, per pose row, the groups whose anchors can
// reach it (k_form_system): at most 512.
static bool incremental_reach_ok(const spgo_solver* h, const HostLayout& L, std::string& why) {
    for (int j = 0; j < h->N; ++j) {
        const int amin = std::max(0, j - L.kmax + 1);
        if (L.anchor_grp[j + 1] - L.anchor_grp[amin] > 512) {
            why = "too many landmark groups reach pose " + std::to_string(j) + "; raise chunk_landmarks or group_chunks";
            return false;
        }
    }
    return true;
}

// The layout in the coordinates the options ask for; AUTO takes incremental
// coordinates when the measurements fit them.
int build_any_layout(spgo_solver* h, HostLayout& out) {
    const int c = h->opt.coordinates;
    if (c != SPGO_COORD_FRAMES) {
        int rc = build_layout(h, out);
        std::string why;
        if (rc == SPGO_OK) {
            if (incremental_reach_ok(h, out, why)) return SPGO_OK;
            if (c == SPGO_COORD_INCREMENTAL) return h->fail(SPGO_E_LAYOUT, why);
        } else if (rc != SPGO_E_LAYOUT || c == SPGO_COORD_INCREMENTAL) {
            return rc;
        }
    }
    return build_layout_frames(h, out);
}

static int factor_path(const spgo_solver* h, const HostLayout& L, std::vector<int>& fb, int& depth) {
    const int Pp = reduced_size(h->N), nbk = Pp / chol_band_block();
    depth = profile_blocks(h->N, L, fb);
    const int lim = 160 * 1024 - 1024;
    const bool fits = depth <= chol_band_max_blocks() && chol_band_lds(Pp, depth) <= (size_t)lim;
    const int nS = std::max(depth, 1);
    const bool split_fits = fits && nbk >= nS + 2 && chol_split_lds(Pp, nS, 2) <= (size_t)lim;
    switch (h->opt.factorization) {
    case SPGO_FACTOR_DENSE: return SPGO_FACTOR_DENSE;
    case SPGO_FACTOR_ENVELOPE: return fits ? SPGO_FACTOR_ENVELOPE : -1;
    case SPGO_FACTOR_ENVELOPE_SPLIT: return split_fits ? SPGO_FACTOR_ENVELOPE_SPLIT : -1;
    case SPGO_FACTOR_CR: return depth <= cr_max_depth() ? SPGO_FACTOR_CR : -1;
    default:   // AUTO, and the Cholesky behind CG
        // cyclic reduction once the profile is long: at 8 super-rows (384 rows,
        // depth 3) an LM iteration takes 0.98 of the split envelope's (its
        // back-substitution also runs the pose update), at 24 (1,536 rows,
        // depth 4) the solve takes 0.56 of it
        if (depth >= 1 && depth <= cr_max_depth() && nbk >= 8 * depth * 16 / chol_band_block())
            return SPGO_FACTOR_CR;
        if (split_fits && nbk >= 2 * depth + 6) return SPGO_FACTOR_ENVELOPE_SPLIT;
        return fits ? SPGO_FACTOR_ENVELOPE : SPGO_FACTOR_DENSE;
    }
}

template <typename T>
int alloc_state(spgo_solver* h, StateBuf<T>* b, const HostLayout& hl, int n_act) {
    for (int i = 0; i < 2; ++i) {
        b[i].X = h->mem.alloc<double>((size_t)h->N * 12);
        b[i].alpha = h->mem.alloc<double>((size_t)h->L);
        b[i].G = h->mem.alloc<double>((size_t)h->F * 12);
        b[i].GinvL = h->mem.alloc<double>((size_t)h->F * 12);
        b[i].GinvR = h->mem.alloc<double>((size_t)h->F * 12);
        b[i].v = h->mem.alloc<T>((size_t)hl.v_total);
        b[i].Ct = h->mem.alloc<T>((size_t)n_act);
        b[i].g = h->mem.alloc<T>((size_t)n_act);
        b[i].D = h->mem.alloc<T>((size_t)hl.D_total);
        b[i].d = h->mem.alloc<T>((size_t)hl.d_total);
        b[i].cost = h->mem.alloc<double>(hl.grp_anchor.size());
        b[i].Rz = hl.frame_mode ? h->mem.alloc<double>((size_t)hl.Rz_total) : nullptr;
        b[i].Dw = hl.frame_mode ? h->mem.alloc<double>((size_t)hl.D_total) : nullptr;
        if (!b[i].X || !b[i].alpha || !b[i].G || !b[i].GinvL || !b[i].GinvR || !b[i].v || !b[i].Ct || !b[i].g ||
            !b[i].D || !b[i].d || !b[i].cost || (hl.frame_mode && (!b[i].Rz || !b[i].Dw)))
            return h->fail(SPGO_E_DEVICE, "device allocation failed");
    }
    return SPGO_OK;
}

template <typename V>
int upload(spgo_solver* h, const std::vector<V>& v, const V** dst) {
    V* p = h->mem.alloc<V>(v.size());
    if (!p) return h->fail(SPGO_E_DEVICE, "device allocation failed");
    if (!v.empty()) {
        hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
        if (e != hipSuccess) return check_hip(h, e, "hipMemcpy");
    }
    *dst = p;
    return SPGO_OK;
}

int ensure_device(spgo_solver* h) {
    if (h->dev_ready) return SPGO_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return h->fail(SPGO_E_DEVICE, "no HIP device available");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, 0);
    if (e != hipSuccess) return check_hip(h, e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return h->fail(SPGO_E_DEVICE, std::string("device 0 is ") + prop.gcnArchName + ", this library is built for gfx950");
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return check_hip(h, e, "hipStreamCreate");
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return h->fail(SPGO_E_DEVICE, "hipEventCreate failed");
    h->dev_ready = true;
    return SPGO_OK;
}

int setup_layout(spgo_solver* h) {
    destroy_graph(h);
    h->mem.release();
    h->rec = nullptr;
    h->rec_cap = 0;
    HostLayout hl;
    int rc = build_any_layout(h, hl);
    if (rc) return rc;
    h->hl = std::move(hl);
    HostLayout& L = h->hl;
    Layout& lay = h->lay;
    lay = Layout{};
    lay.N = h->N;
    lay.F = h->F;
    lay.L = h->L;
    lay.n_act = (int)L.lm_id.size();
    lay.n_groups = (int)L.grp_anchor.size();
    lay.n_chunks = (int)L.chunk.size();
    lay.n_obs = (int)L.ob_t.size();
    lay.kmax = L.kmax;
    lay.multi_chunk = 0;
    for (const GroupDesc& d : L.grp_desc) lay.multi_chunk |= d.chunk_end - d.chunk_begin > 1;
    lay.form_rows = 1;   // one problem: one row per workgroup fills the device
    lay.frame_mode = L.frame_mode;
    lay.soff_stride = L.frame_mode ? L.kt + 2 : KMAX + 2;
    if (L.frame_mode &&
        ((rc = upload(h, L.grp_frames, &lay.grp_frames)) || (rc = upload(h, L.zp_row, &lay.zp_row)) ||
         (rc = upload(h, L.zp_col, &lay.zp_col)) || (rc = upload(h, L.zp_cbeg, &lay.zp_cbeg)) ||
         (rc = upload(h, L.zp_con, &lay.zp_con))))
        return rc;
    if ((rc = upload(h, L.ob_t, &lay.ob_t)) || (rc = upload(h, L.ob_meta, &lay.ob_meta)) ||
        (rc = upload(h, L.ob_uv, &lay.ob_uv)) || (rc = upload(h, L.lm_id, &lay.lm_id)) ||
        (rc = upload(h, L.lm_ray, &lay.lm_ray)) || (rc = upload(h, L.lm_pos, &lay.lm_pos)) ||
        (rc = upload(h, L.chunk, &lay.chunk)) || (rc = upload(h, L.chunk_soff, &lay.chunk_soff)) || (rc = upload(h, launch_order(L.grp_desc), &lay.grp_desc)) || (rc = upload(h, L.fs_begin, &lay.fs_begin)) || (rc = upload(h, L.fs_list, &lay.fs_list)))
        return rc;
    h->prec = h->opt.precision;
    if (h->prec == SPGO_PREC_MIXED)
        rc = alloc_state<float>(h, h->bf, L, lay.n_act);
    else
        rc = alloc_state<double>(h, h->bd, L, lay.n_act);
    if (rc) return rc;
    const int P = 6 * h->N;
    h->Pp = reduced_size(h->N);
    {
        std::vector<int> fb;
        int bw = 0;
        h->path = factor_path(h, L, fb, bw);
        if (h->path < 0)
            return h->fail(SPGO_E_LAYOUT, "the reduced-system profile does not fit the requested factorisation");
        h->band = h->path != SPGO_FACTOR_DENSE;
        h->pcg = h->opt.factorization == SPGO_FACTOR_PCG;
        h->lps.release();
        if (h->pcg) {
            // CG over the pose blocks: spans of K poses couple blocks up to K - 1 apart
            const int need = std::max(1, L.band_need);
            const int W = h->net ? lpcg_band_blocks(h->net) : std::min(need, LP_WMAX);
            std::vector<int> far_beg, far_col;
            if (need > W && !L.frame_mode)
                return h->fail(SPGO_E_LAYOUT, "landmark spans need a band of " + std::to_string(need) +
                                                  " blocks; the preconditioner network covers " + std::to_string(W));
            if (need > W) {
                // frame coordinates: the blocks farther than W from the diagonal
                // (loop revisits) join the CG product, not the preconditioner
                std::vector<std::vector<int>> rows(h->N);
                for (int r = 0; r < h->N; ++r)
                    for (int k = L.zp_row[r]; k < L.zp_row[r + 1]; ++k) {
                        const int c = L.zp_col[k];
                        if (r - c > W) {
                            rows[r].push_back(c);
                            rows[c].push_back(r);
                        }
                    }
                far_beg.assign(h->N + 1, 0);
                for (int r = 0; r < h->N; ++r) {
                    std::sort(rows[r].begin(), rows[r].end());
                    far_beg[r + 1] = far_beg[r] + (int)rows[r].size();
                    far_col.insert(far_col.end(), rows[r].begin(), rows[r].end());
                }
            }
            if (h->lps.setup(h->N, W, far_beg, far_col) != 0)
                return h->fail(SPGO_E_DEVICE, "device allocation failed (CG buffers)");
            if (!h->pcg_flag_host &&
                hipHostMalloc((void**)&h->pcg_flag_host, sizeof(int), hipHostMallocDefault) != hipSuccess) {
                h->pcg_flag_host = nullptr;
                return h->fail(SPGO_E_DEVICE, "pinned host allocation failed");
            }
        }
        if (h->band && !(h->Lpan = h->mem.alloc<double>(chol_panel_doubles(h->Pp))))
            return h->fail(SPGO_E_DEVICE, "device allocation failed");
        if ((rc = upload(h, fb, &h->first))) return rc;
        {
            // the dense path's block profile: zero blocks of the factor are skipped
            std::vector<int> f32, aoff, act;
            dense_profile(h->Pp, fb, f32, aoff, act, h->dense_nact);
            if ((rc = upload(h, f32, &h->dense_first)) || (rc = upload(h, aoff, &h->dense_act_off)) ||
                (rc = upload(h, act, &h->dense_act)))
                return rc;
            // the pipelined solves' forward solution, and their tickets and ready flags
            h->solve_y = h->mem.alloc<double>(h->Pp);
            h->solve_sync = h->mem.alloc<int>(2 + 2 * (h->Pp / NB));
            if (!h->solve_y || !h->solve_sync) return h->fail(SPGO_E_DEVICE, "device allocation failed");
        }
        h->band_depth = bw;
        if (h->path == SPGO_FACTOR_ENVELOPE && prepare_chol_band(bw, chol_band_lds(h->Pp, bw)) != 0)
            return h->fail(SPGO_E_DEVICE, "cannot reserve LDS for the envelope Cholesky");
        if (h->path == SPGO_FACTOR_ENVELOPE_SPLIT) {
            // separator of D block rows in the middle; the bottom part's profile in
            // reversed order: local row r is block row nb - 1 - r, its first column
            // nb - 1 - last(nb - 1 - r), last(i) the last block row j with fb[j] <= i
            const int nbk = h->Pp / chol_band_block(), nS = std::max(bw, 1);
            h->s0 = (nbk - nS + 1) / 2;
            h->s1 = h->s0 + nS;
            std::vector<int> last(nbk);
            for (int i = 0; i < nbk; ++i) {
                int j = i;
                while (j + 1 < nbk && fb[j + 1] <= i) ++j;
                last[i] = j;
            }
            std::vector<int> fB(nbk - h->s0);
            for (int r = 0; r < nbk - h->s0; ++r) fB[r] = nbk - 1 - last[nbk - 1 - r];
            if ((rc = upload(h, fB, &h->firstB))) return rc;
            const int EBk = chol_band_block();
            h->Ssep = h->mem.alloc<double>((size_t)2 * nS * nS * EBk * EBk + (size_t)2 * EBk * nS);
            h->Smrg = h->mem.alloc<double>((size_t)(EBk * nS) * (EBk * nS));
            h->winvS = h->mem.alloc<double>((size_t)nS * EBk * EBk);
            h->LpS = h->mem.alloc<double>(chol_panel_doubles(EBk * nS));
            if (!h->Ssep || !h->Smrg || !h->winvS || !h->LpS) return h->fail(SPGO_E_DEVICE, "device allocation failed");
            if (prepare_chol_split(h->Pp, bw) != 0)
                return h->fail(SPGO_E_DEVICE, "cannot reserve LDS for the split envelope Cholesky");
        }
        if (h->path == SPGO_FACTOR_CR) {
            const size_t n = cr_work_doubles(h->Pp, std::max(bw, 1));
            if (!(h->crw = h->mem.alloc<double>(n))) return h->fail(SPGO_E_DEVICE, "device allocation failed");
            launch_fill(h->crw, 0.0, (long long)n, h->stream);   // the reduction's counters start at zero
        }
    }
    h->H = h->mem.alloc<double>((size_t)h->Pp * h->Pp);
    h->gvec = h->mem.alloc<double>(h->Pp);
    h->delta = h->mem.alloc<double>(h->Pp);
    h->Linv = h->mem.alloc<double>((size_t)(h->Pp / NB) * NB * NB);
    h->ctl = h->mem.alloc<Control>(1);
    if (!h->H || !h->gvec || !h->delta || !h->Linv || !h->ctl)
        return h->fail(SPGO_E_DEVICE, "device allocation failed");
    {
        // a problem table of one entry: the batch kernels' view of this solver
        ProbDesc pd{};
        pd.N = h->N;
        pd.g_begin = 0;
        pd.g_end = lay.n_groups;
        pd.Pp = h->Pp;
        pd.D = h->band_depth;
        std::vector<ProbDesc> probs(1, pd);
        h->pd1 = pd;
        std::vector<int> pose_prob(h->N, 0);
        if ((rc = upload(h, probs, &lay.probs)) || (rc = upload(h, pose_prob, &lay.pose_prob))) return rc;
        lay.n_probs = 1;
    }
    launch_fill(h->H, 0.0, (long long)h->Pp * h->Pp, h->stream);
    launch_set_diag(h->H, h->Pp, P, 1.0, h->stream);
    launch_fill(h->gvec, 0.0, h->Pp, h->stream);
    launch_fill(h->delta, 0.0, h->Pp, h->stream);
    Control c{};
    c.cur = 0;
    c.chol_ok = 1;
    c.pcg_fail = 1;
    c.lambda = 0.0;
    rc = check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    h->layout_dirty = false;
    h->lin_valid = false;
    h->host_newer = true;   // state buffers are new
    h->lambda_reset = true;
    return SPGO_OK;
}

Camera make_camera(const spgo_solver* h) {
    Camera c;
    for (int k = 0; k < 2; ++k) {
        c.fx[k] = h->intr[4 * k + 0];
        c.fy[k] = h->intr[4 * k + 1];
        c.cx[k] = h->intr[4 * k + 2];
        c.cy[k] = h->intr[4 * k + 3];
    }
    for (int i = 0; i < 12; ++i) c.Trl[i] = h->Trl[i];
    project_rotation(c.Trl);
    return c;
}

AsmArgs make_asm(const spgo_solver* h, int init) {
    AsmArgs a;
    a.lay = h->lay;
    a.cam = make_camera(h);
    a.rb.kind = h->opt.robust;
    a.rb.near_boost = h->opt.near_boost;
    a.rb.param = h->opt.robust_param;
    a.eps = h->opt.eps;
    a.alpha_min = h->opt.alpha_min;
    a.alpha_max = h->opt.alpha_max;
    a.ctl = h->ctl;
    a.init = init;
    a.delta = h->delta;
    return a;
}

// the band factorisation and both solves: envelope (one or two kernels) or cyclic reduction.
// Cyclic reduction without CG also runs the pose update ps in its last
// back-substitution workgroup; returns whether it did.
static bool launch_chol(spgo_solver* h, const PoseStep* ps, hipStream_t s) {
    if (h->path == SPGO_FACTOR_CR)
        return launch_cr(h->H, h->gvec, h->delta, h->Pp, std::max(h->band_depth, 1), h->crw, h->ctl,
                         h->pcg ? 1 : 0, h->pcg ? nullptr : ps, s);
    if (h->path == SPGO_FACTOR_ENVELOPE_SPLIT)
        launch_chol_split(h->H, h->Pp, h->band_depth, h->s0, h->s1, h->first, h->firstB, h->gvec, h->delta, h->Linv,
                          h->Lpan, h->Ssep, h->Smrg, h->winvS, h->LpS, h->ctl, s);
    else
        launch_chol_band(h->lay.probs, 1, h->band_depth, chol_band_lds(h->Pp, h->band_depth), h->H, h->first, h->gvec,
                         h->delta, h->Linv, h->Lpan, h->ctl, s);
    return false;
}

// the candidate's pose update (step size s) as the cyclic reduction runs it
template <typename T>
static PoseStep pose_step(spgo_solver* h, StateBuf<T>* b) {
    PoseStep ps{};
    ps.pd = h->pd1;
    ps.cam = make_camera(h);
    ps.b = pose_bufs(b[0], b[1]);
    ps.step = (double)h->step_size;
    ps.frame_mode = h->lay.frame_mode;
    return ps;
}

template <typename T>
int launch_iteration(spgo_solver* h, StateBuf<T>* b) {
    hipStream_t s = h->stream;
    launch_form_system<T>(h->lay, b, h->ctl, h->H, h->gvec, s);
    bool factor = true;
    if (h->pcg) {
        // CG first; the host reads its convergence flag and launches the
        // Cholesky only when CG did not converge
        int rc = check_hip(h, (hipError_t)lp_solve_launch(h->lps, h->H, h->Pp, h->gvec, h->delta, h->net_params,
                                                          h->opt.pcg_tol, h->opt.pcg_max_iter, h->ctl, s),
                           "cooperative CG launch");
        if (rc) return rc;
        if ((rc = check_hip(h, hipMemcpyAsync(h->pcg_flag_host, &h->ctl->pcg_fail, sizeof(int),
                                              hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) ||
            (rc = check_hip(h, hipStreamSynchronize(s), "CG")))
            return rc;
        factor = *h->pcg_flag_host != 0;
    }
    bool posed = false;
    if (factor) {
        if (h->band) {
            const PoseStep ps = pose_step(h, b);
            posed = launch_chol(h, &ps, s);
        } else {
            launch_cholesky(h->H, h->Linv, h->Pp, h->dense_act_off, h->dense_act, h->dense_nact.data(), h->ctl, s);
            launch_chol_solve(h->H, h->Linv, h->gvec, h->delta, h->Pp, h->dense_first, h->dense_act_off,
                              h->dense_act, h->solve_y, h->solve_sync, h->ctl, s);
        }
    }
    if (!posed) launch_pose_update<T>(h->lay, make_camera(h), b, h->ctl, h->delta, (double)h->step_size, 0, s);
    launch_assemble<T>(make_asm(h, 0), b, s);
    launch_accept(b[0].cost, b[1].cost, h->lay, h->ctl, h->rec, h->opt.mode == SPGO_MODE_REFERENCE,
                  h->opt.lam_up, h->opt.lam_down, h->opt.lam_min, h->opt.lam_max, 0, s);
    return SPGO_OK;
}

template <typename T>
void launch_linearise(spgo_solver* h, StateBuf<T>* b) {
    hipStream_t s = h->stream;
    launch_pose_update<T>(h->lay, make_camera(h), b, h->ctl, h->delta, 0.0, 1, s);
    launch_assemble<T>(make_asm(h, 1), b, s);
    launch_accept(b[0].cost, b[1].cost, h->lay, h->ctl, h->rec, 0, 0, 0, 0, 0, 1, s);
}

// Graph slot k holds `iters` consecutive iterations (captured again when the
// count differs); iters == 0 keeps what the slot holds for a later call.
static int capture_iterations(spgo_solver* h, int k, int iters, bool mixed) {
    if (iters == 0 || (iters == h->graph_iters[k] && h->gexec[k])) return SPGO_OK;
    destroy_graph_slot(h, k);
    int rc = check_hip(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) return rc;
    for (int it = 0; it < iters && !rc; ++it)
        rc = mixed ? launch_iteration<float>(h, h->bf) : launch_iteration<double>(h, h->bd);
    const int rc2 = check_hip(h, hipStreamEndCapture(h->stream, &h->graph[k]), "hipStreamEndCapture");
    if (rc) return rc;
    if (rc2) return rc2;
    rc = check_hip(h, hipGraphInstantiate(&h->gexec[k], h->graph[k], nullptr, nullptr, 0), "hipGraphInstantiate");
    if (rc) return rc;
    h->graph_iters[k] = iters;
    return SPGO_OK;
}

// Copy the device state into the host mirror if the device is newer.
int pull_state(spgo_solver* h) {
    if (!h->dev_newer) return SPGO_OK;
    Control c;
    int rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    const double* X = h->prec == SPGO_PREC_MIXED ? h->bf[c.cur].X : h->bd[c.cur].X;
    const double* al = h->prec == SPGO_PREC_MIXED ? h->bf[c.cur].alpha : h->bd[c.cur].alpha;
    rc = check_hip(h, hipMemcpy(h->X.data(), X, h->X.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    rc = check_hip(h, hipMemcpy(h->alpha.data(), al, h->alpha.size() * sizeof(double), hipMemcpyDeviceToHost),
                   "hipMemcpy");
    if (rc) return rc;
    h->dev_newer = false;
    return SPGO_OK;
}

// The host mirror is about to be modified.
int begin_state_write(spgo_solver* h) {
    int rc = pull_state(h);
    if (rc) return rc;
    h->host_newer = true;
    h->lin_valid = false;
    h->lambda_reset = true;
    return SPGO_OK;
}

int push_state(spgo_solver* h) {
    Control c;
    int rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    for (int k = 0; k < h->N; ++k) project_rotation(&h->X[12 * k]);
    for (int i = 0; i < 2; ++i) {
        double* X = h->prec == SPGO_PREC_MIXED ? h->bf[i].X : h->bd[i].X;
        double* al = h->prec == SPGO_PREC_MIXED ? h->bf[i].alpha : h->bd[i].alpha;
        rc = check_hip(h, hipMemcpy(X, h->X.data(), h->X.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
        if (rc) return rc;
        rc = check_hip(h, hipMemcpy(al, h->alpha.data(), h->alpha.size() * sizeof(double), hipMemcpyHostToDevice),
                       "hipMemcpy");
        if (rc) return rc;
    }
    h->host_newer = false;
    h->lin_valid = false;
    return SPGO_OK;
}

int run_step(spgo_solver* h, int iterations) {
    int rc = ensure_device(h);
    if (rc) return rc;
    if (h->layout_dirty || h->prec != h->opt.precision) {
        rc = pull_state(h);
        if (rc) return rc;
        rc = setup_layout(h);
        if (rc) return rc;
    }
    if (h->host_newer) {
        rc = push_state(h);
        if (rc) return rc;
    }
    if (iterations > h->rec_cap) {
        destroy_graph(h);
        const int cap = std::max(iterations, 64);
        h->rec = h->mem.alloc<IterRecord>(cap);
        if (!h->rec) return h->fail(SPGO_E_DEVICE, "device allocation failed");
        h->rec_cap = cap;
    }
    // pinned copies of the control block and the records, read back once the
    // step's work is done (one stream synchronisation for all of it)
    if (!h->ctl_host && hipHostMalloc((void**)&h->ctl_host, sizeof(Control), hipHostMallocDefault) != hipSuccess) {
        h->ctl_host = nullptr;
        return h->fail(SPGO_E_DEVICE, "pinned host allocation failed");
    }
    if (h->rec_host_cap < h->rec_cap) {
        if (h->rec_host) (void)hipHostFree(h->rec_host);
        h->rec_host_cap = 0;
        if (hipHostMalloc((void**)&h->rec_host, h->rec_cap * sizeof(IterRecord), hipHostMallocDefault) != hipSuccess) {
            h->rec_host = nullptr;
            return h->fail(SPGO_E_DEVICE, "pinned host allocation failed");
        }
        h->rec_host_cap = h->rec_cap;
    }
    // on the device: restart the records and, with the linearisation kept, the
    // initial cost (a fresh linearisation sets both itself)
    launch_begin_step(h->ctl, h->lambda_reset ? 1 : 0, h->opt.mode == SPGO_MODE_REFERENCE ? 0.0 : h->opt.lambda0,
                      h->lin_valid ? 1 : 0, h->stream);
    h->lambda_reset = false;
    const bool mixed = h->prec == SPGO_PREC_MIXED;
    h->net_params = nullptr;
    if (h->pcg && h->net && !(h->net_params = lpcg_device_params(h->net)))
        return h->fail(SPGO_E_DEVICE, "cannot read the preconditioner network's parameters");
    if (!h->lin_valid) {
        if (mixed)
            launch_linearise<float>(h, h->bf);
        else
            launch_linearise<double>(h, h->bd);
        h->lin_valid = true;
    }
    rc = check_hip(h, hipGetLastError(), "kernel launch");
    if (rc) return rc;
    // CG decides on the host whether the Cholesky runs: no graph for it.  One
    // graph holds GRAPH_UNROLL iterations and is launched repeatedly; a second
    // one holds the remainder of the call.
    const bool graph = h->opt.use_graph && !h->pcg;
    const int unroll = std::min(iterations, GRAPH_UNROLL);
    const int n_full = graph && iterations > 0 ? iterations / unroll : 0;
    const int n_rem = graph && iterations > 0 ? iterations % unroll : 0;
    if (graph && iterations > 0) {
        if (h->graph_dirty) destroy_graph(h);
        if ((rc = capture_iterations(h, 0, unroll, mixed)) || (rc = capture_iterations(h, 1, n_rem, mixed)))
            return rc;
        h->graph_dirty = false;
    }
    rc = check_hip(h, hipEventRecord(h->ev0, h->stream), "hipEventRecord");
    if (rc) return rc;
    if (graph) {
        for (int it = 0; it < n_full && !rc; ++it)
            rc = check_hip(h, hipGraphLaunch(h->gexec[0], h->stream), "hipGraphLaunch");
        if (!rc && n_rem > 0) rc = check_hip(h, hipGraphLaunch(h->gexec[1], h->stream), "hipGraphLaunch");
        if (rc) return rc;
    } else {
        for (int it = 0; it < iterations; ++it) {
            rc = mixed ? launch_iteration<float>(h, h->bf) : launch_iteration<double>(h, h->bd);
            if (rc) return rc;
        }
    }
    rc = check_hip(h, hipEventRecord(h->ev1, h->stream), "hipEventRecord");
    if (rc) return rc;
    // every iteration writes one record: `iterations` of them, at most
    rc = check_hip(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(Control), hipMemcpyDeviceToHost, h->stream),
                   "hipMemcpyAsync");
    if (!rc && iterations > 0)
        rc = check_hip(h, hipMemcpyAsync(h->rec_host, h->rec, iterations * sizeof(IterRecord), hipMemcpyDeviceToHost,
                                         h->stream), "hipMemcpyAsync");
    if (rc) return rc;
    rc = check_hip(h, hipStreamSynchronize(h->stream), "hipStreamSynchronize");
    if (rc) return rc;
    rc = check_hip(h, hipGetLastError(), "kernel execution");
    if (rc) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, h->ev0, h->ev1);
    h->last_ms = ms;
    const Control c = *h->ctl_host;
    if (c.iter < 0 || c.iter > std::max(iterations, 0))
        return h->fail(SPGO_E_DEVICE, "the device wrote an unexpected number of iteration records");
    h->last_rec.assign(h->rec_host, h->rec_host + c.iter);
    h->last_cost0 = c.cost0;
    h->dev_newer = true;
    return SPGO_OK;
}

bool valid_opts(const spgo_options* o, std::string& why) {
    if (o->mode != SPGO_MODE_REFERENCE && o->mode != SPGO_MODE_LM) return why = "mode", false;
    if (o->robust < 0 || o->robust > 2) return why = "robust", false;
    if (o->precision != SPGO_PREC_FP64 && o->precision != SPGO_PREC_MIXED) return why = "precision", false;
    if (!(o->eps >= 0.0)) return why = "eps", false;
    if (o->robust != SPGO_ROBUST_NONE && !(o->robust_param > 0.0)) return why = "robust_param", false;
    if (!(o->alpha_min > 0.0) || !(o->alpha_max >= o->alpha_min)) return why = "alpha bounds", false;
    if (o->chunk_landmarks < 0 || o->chunk_landmarks > CHUNK_LM) return why = "chunk_landmarks", false;
    if (o->group_chunks < 0) return why = "group_chunks", false;
    if (o->factorization < 0 || o->factorization > 5) return why = "factorization", false;
    if (o->coordinates < 0 || o->coordinates > 2) return why = "coordinates", false;
    if (o->pcg_max_iter < 1) return why = "pcg_max_iter", false;
    if (!(o->pcg_tol > 0.0)) return why = "pcg_tol", false;
    return true;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int spgo_default_options(spgo_options* o) {
    if (!o) return SPGO_E_ARG;
    std::memset(o, 0, sizeof(*o));
    o->mode = SPGO_MODE_LM;
    o->robust = SPGO_ROBUST_CAUCHY;
    o->near_boost = 1;
    o->precision = SPGO_PREC_FP64;
    o->robust_param = 9.0;
    o->eps = 1e-4;
    o->lambda0 = 1e-4;
    o->lam_up = 10.0;
    o->lam_down = 0.1;
    o->lam_min = 1e-10;
    o->lam_max = 1e10;
    o->alpha_min = 1.0 / 30.0;
    o->alpha_max = 1.0 / 0.2;
    o->use_graph = 1;
    o->chunk_landmarks = CHUNK_LM;
    o->group_chunks = 0;   // chosen from the device's resident workgroups
    o->pcg_max_iter = 200;
    o->pcg_tol = 1e-8;
    return SPGO_OK;
}

int spgo_create(int32_t n_poses, int32_t m_per_frame, const spgo_options* opts, spgo_solver** out) {
    if (!out) return SPGO_E_ARG;
    *out = nullptr;
    if (n_poses < 1 || m_per_frame < 1) {
        g_create_error = "N and M must be positive";
        return SPGO_E_ARG;
    }
    if ((long long)n_poses * m_per_frame > (1ll << 30)) {
        g_create_error = "N*M too large";
        return SPGO_E_ARG;
    }
    std::unique_ptr<spgo_solver> h(new spgo_solver());
    h->N = n_poses;
    h->M = m_per_frame;
    h->L = n_poses * m_per_frame;
    h->F = n_poses + 1;
    h->max_meas = (long long)n_poses * (n_poses + 1) * m_per_frame;
    if (opts) {
        std::string why;
        if (!valid_opts(opts, why)) {
            g_create_error = "invalid option: " + why;
            return SPGO_E_ARG;
        }
        h->opt = *opts;
    } else {
        spgo_default_options(&h->opt);
    }
    // identity incremental poses, inverse depths 1.0 (include/Solver.hpp:36-44)
    h->X.assign((size_t)h->N * 12, 0.0);
    for (int k = 0; k < h->N; ++k) h->X[12 * k + 0] = h->X[12 * k + 5] = h->X[12 * k + 10] = 1.0;
    h->alpha.assign(h->L, 1.0);
    *out = h.release();
    return SPGO_OK;
}

int spgo_destroy(spgo_solver* h) {
    if (!h) return SPGO_E_ARG;
    // the solver's queued work finishes before its memory goes
    if (h->dev_ready) (void)hipStreamSynchronize(h->stream);
    destroy_graph(h);
    h->mem.release();
    h->lps.release();
    if (h->pcg_flag_host) (void)hipHostFree(h->pcg_flag_host);
    if (h->ctl_host) (void)hipHostFree(h->ctl_host);
    if (h->rec_host) (void)hipHostFree(h->rec_host);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPGO_OK;
}

int spgo_set_options(spgo_solver* h, const spgo_options* o) {
    if (!h || !o) return SPGO_E_ARG;
    std::string why;
    if (!valid_opts(o, why)) return h->fail(SPGO_E_ARG, "invalid option: " + why);
    const bool relayout = o->precision != h->opt.precision || o->chunk_landmarks != h->opt.chunk_landmarks ||
                          o->group_chunks != h->opt.group_chunks || o->factorization != h->opt.factorization ||
                          o->coordinates != h->opt.coordinates;
    const bool relin = relayout || o->robust != h->opt.robust || o->robust_param != h->opt.robust_param ||
                       o->near_boost != h->opt.near_boost || o->eps != h->opt.eps;
    if (o->mode != h->opt.mode || o->lambda0 != h->opt.lambda0) h->lambda_reset = true;
    if (relayout) {
        int rc = pull_state(h);
        if (rc) return rc;
        h->layout_dirty = true;
    }
    if (relin) h->lin_valid = false;
    h->opt = *o;
    destroy_graph(h);
    return SPGO_OK;
}

int spgo_get_options(spgo_solver* h, spgo_options* out) {
    if (!h || !out) return SPGO_E_ARG;
    *out = h->opt;
    return SPGO_OK;
}

int spgo_load_calibration(spgo_solver* h, const float intrinsics[8], const float T_r_to_l[16]) {
    if (!h || !intrinsics || !T_r_to_l) return SPGO_E_ARG;
    std::memcpy(h->intr, intrinsics, sizeof(h->intr));
    std::memcpy(h->Trl, T_r_to_l, sizeof(h->Trl));
    h->lin_valid = false;
    destroy_graph(h);
    return SPGO_OK;
}

int spgo_get_calibration(spgo_solver* h, float intrinsics[8], float T_r_to_l[16]) {
    if (!h || !intrinsics || !T_r_to_l) return SPGO_E_ARG;
    std::memcpy(intrinsics, h->intr, sizeof(h->intr));
    std::memcpy(T_r_to_l, h->Trl, sizeof(h->Trl));
    return SPGO_OK;
}

static int write_one(spgo_solver* h, int a, int t, int l, const float* left, const float* right) {
    if (a < 0 || a >= h->F || t < 0 || t >= h->F || l < 0 || l >= h->L)
        return h->fail(SPGO_E_ARG, "frame or landmark id out of range");
    if ((long long)h->ma.size() >= h->max_meas) return h->fail(SPGO_E_CAPACITY, "measurement capacity N(N+1)M exceeded");
    const long long key = (long long)t * h->L + l;
    auto it = h->cell_of.find(key);
    int cell;
    if (it == h->cell_of.end()) {
        cell = (int)(h->cells.size() / 6);
        h->cells.resize(h->cells.size() + 6);
        h->cell_of.emplace(key, cell);
    } else {
        cell = it->second;
    }
    float* c = &h->cells[6 * cell];
    for (int i = 0; i < 3; ++i) {
        c[i] = left[i];
        c[3 + i] = right[i];
    }
    h->ma.push_back(a);
    h->mt.push_back(t);
    h->ml.push_back(l);
    return SPGO_OK;
}

int spgo_write_observation(spgo_solver* h, int32_t anchor, int32_t target, int32_t feat, const float left[3],
                           const float right[3]) {
    if (!h || !left || !right) return SPGO_E_ARG;
    int rc = write_one(h, anchor, target, feat, left, right);
    if (rc) return rc;
    h->layout_dirty = true;
    h->lin_valid = false;
    return SPGO_OK;
}

int spgo_write_observations_bulk(spgo_solver* h, int32_t count, const int32_t* anchor, const int32_t* target,
                                 const int32_t* feat, const float* uvLR) {
    if (!h || count < 0 || (count > 0 && (!anchor || !target || !feat || !uvLR))) return SPGO_E_ARG;
    // every entry is checked before any is written, so a rejected call leaves
    // the measurement table as it was
    for (int i = 0; i < count; ++i)
        if (anchor[i] < 0 || anchor[i] >= h->F || target[i] < 0 || target[i] >= h->F || feat[i] < 0 ||
            feat[i] >= h->L)
            return h->fail(SPGO_E_ARG, "frame or landmark id out of range at entry " + std::to_string(i) +
                                           "; nothing was written");
    if ((long long)h->ma.size() + count > h->max_meas)
        return h->fail(SPGO_E_CAPACITY, "measurement capacity N(N+1)M exceeded; nothing was written");
    h->ma.reserve(h->ma.size() + count);
    h->mt.reserve(h->mt.size() + count);
    h->ml.reserve(h->ml.size() + count);
    h->cell_of.reserve(h->cell_of.size() + count);
    for (int i = 0; i < count; ++i) {
        const float left[3] = {uvLR[4 * i], uvLR[4 * i + 1], 1.f};
        const float right[3] = {uvLR[4 * i + 2], uvLR[4 * i + 3], 1.f};
        int rc = write_one(h, anchor[i], target[i], feat[i], left, right);
        if (rc) return rc;
    }
    h->layout_dirty = true;
    h->lin_valid = false;
    return SPGO_OK;
}

int spgo_get_observation(spgo_solver* h, int32_t frame, int32_t feat, float left[3], float right[3]) {
    if (!h || !left || !right) return SPGO_E_ARG;
    if (frame < 0 || frame >= h->F || feat < 0 || feat >= h->L) return h->fail(SPGO_E_ARG, "id out of range");
    auto it = h->cell_of.find((long long)frame * h->L + feat);
    for (int i = 0; i < 3; ++i) {
        left[i] = it == h->cell_of.end() ? 0.f : h->cells[6 * it->second + i];
        right[i] = it == h->cell_of.end() ? 0.f : h->cells[6 * it->second + 3 + i];
    }
    return SPGO_OK;
}

int spgo_step(spgo_solver* h, int32_t iterations) {
    if (!h || iterations < 0) return SPGO_E_ARG;
    return run_step(h, iterations);
}

int spgo_get_incremental_pose(spgo_solver* h, int32_t k, float T[16]) {
    if (!h || !T) return SPGO_E_ARG;
    if (k < 0 || k >= h->N) return h->fail(SPGO_E_ARG, "pose index out of range");
    int rc = pull_state(h);
    if (rc) return rc;
    for (int i = 0; i < 12; ++i) T[i] = (float)h->X[12 * k + i];
    T[12] = T[13] = T[14] = 0.f;
    T[15] = 1.f;
    return SPGO_OK;
}

int spgo_write_incremental_pose(spgo_solver* h, int32_t k, const float T[16]) {
    if (!h || !T) return SPGO_E_ARG;
    if (k < 0 || k >= h->N) return h->fail(SPGO_E_ARG, "pose index out of range");
    int rc = begin_state_write(h);
    if (rc) return rc;
    mat4_to_3x4(T, &h->X[12 * k]);   // rows 0-2 only (Solver.cpp:69-84)
    return SPGO_OK;
}

int spgo_load_inverse_depths(spgo_solver* h, const float* alpha) {
    if (!h || !alpha) return SPGO_E_ARG;
    int rc = begin_state_write(h);
    if (rc) return rc;
    for (int i = 0; i < h->L; ++i) h->alpha[i] = (double)alpha[i];
    return SPGO_OK;
}

int spgo_get_inverse_depths(spgo_solver* h, float* alpha) {
    if (!h || !alpha) return SPGO_E_ARG;
    int rc = pull_state(h);
    if (rc) return rc;
    for (int i = 0; i < h->L; ++i) alpha[i] = (float)h->alpha[i];
    return SPGO_OK;
}

int spgo_init_inverse_depths_from_disparity(spgo_solver* h, int32_t* n_set) {
    if (!h) return SPGO_E_ARG;
    int rc = begin_state_write(h);
    if (rc) return rc;
    const double fl[4] = {h->intr[0], h->intr[1], h->intr[2], h->intr[3]};
    const double fr[4] = {h->intr[4], h->intr[5], h->intr[6], h->intr[7]};
    const float* T = h->Trl;
    std::vector<char> done(h->L, 0);
    int cnt = 0;
    for (size_t m = 0; m < h->ma.size(); ++m) {
        const int l = h->ml[m];
        if (done[l]) continue;
        done[l] = 1;   // the anchor is the same for all of a landmark's measurements
        auto it = h->cell_of.find((long long)h->ma[m] * h->L + l);
        if (it == h->cell_of.end()) continue;
        const float* c = &h->cells[6 * it->second];
        // rays b = K^-1 [u, v, w] of the homogeneous pixels in both cameras
        const double bl[3] = {(c[0] - fl[2] * c[2]) / fl[0], (c[1] - fl[3] * c[2]) / fl[1], (double)c[2]};
        const double br[3] = {(c[3] - fr[2] * c[5]) / fr[0], (c[4] - fr[3] * c[5]) / fr[1], (double)c[5]};
        double v[3], t[3];
        for (int r = 0; r < 3; ++r) {
            v[r] = T[4 * r] * br[0] + T[4 * r + 1] * br[1] + T[4 * r + 2] * br[2];
            t[r] = T[4 * r + 3];
        }
        const double nb[3] = {v[1] * bl[2] - v[2] * bl[1], v[2] * bl[0] - v[0] * bl[2], v[0] * bl[1] - v[1] * bl[0]};
        const double nt[3] = {v[1] * t[2] - v[2] * t[1], v[2] * t[0] - v[0] * t[2], v[0] * t[1] - v[1] * t[0]};
        const double num = nb[0] * nt[0] + nb[1] * nt[1] + nb[2] * nt[2];
        const double den = nb[0] * nb[0] + nb[1] * nb[1] + nb[2] * nb[2];
        if (!(num > 0.0) || !(den > 0.0)) continue;
        h->alpha[l] = den / num;   // 1 / depth
        ++cnt;
    }
    if (n_set) *n_set = cnt;
    return SPGO_OK;
}

int spgo_set_step_size(spgo_solver* h, float s) {
    if (!h) return SPGO_E_ARG;
    if (s != h->step_size) destroy_graph(h);
    h->step_size = s;
    return SPGO_OK;
}

int spgo_get_state_f64(spgo_solver* h, double* poses, double* alpha) {
    if (!h) return SPGO_E_ARG;
    int rc = pull_state(h);
    if (rc) return rc;
    if (poses) {
        for (int k = 0; k < h->N; ++k) {
            for (int i = 0; i < 12; ++i) poses[16 * k + i] = h->X[12 * k + i];
            poses[16 * k + 12] = poses[16 * k + 13] = poses[16 * k + 14] = 0.0;
            poses[16 * k + 15] = 1.0;
        }
    }
    if (alpha) std::memcpy(alpha, h->alpha.data(), h->alpha.size() * sizeof(double));
    return SPGO_OK;
}

int spgo_set_state_f64(spgo_solver* h, const double* poses, const double* alpha) {
    if (!h) return SPGO_E_ARG;
    int rc = begin_state_write(h);
    if (rc) return rc;
    if (poses)
        for (int k = 0; k < h->N; ++k)
            for (int i = 0; i < 12; ++i) h->X[12 * k + i] = poses[16 * k + i];
    if (alpha) std::memcpy(h->alpha.data(), alpha, h->alpha.size() * sizeof(double));
    return SPGO_OK;
}

int spgo_get_stats(spgo_solver* h, spgo_iter_stats* out, int32_t max_records, int32_t* n_out, double* initial_cost) {
    if (!h) return SPGO_E_ARG;
    const int n = (int)h->last_rec.size();
    if (n_out) *n_out = n;
    if (initial_cost) *initial_cost = h->last_cost0;
    if (out) {
        for (int i = 0; i < n && i < max_records; ++i) {
            out[i].cost_candidate = h->last_rec[i].cost_candidate;
            out[i].cost = h->last_rec[i].cost;
            out[i].lambda = h->last_rec[i].lambda;
            out[i].accepted = h->last_rec[i].accepted;
            out[i].chol_ok = h->last_rec[i].chol_ok;
            out[i].pcg_iters = h->last_rec[i].pcg_iters;
            out[i].pcg_fallback = h->pcg ? h->last_rec[i].pcg_fail : 0;
        }
    }
    return SPGO_OK;
}

int spgo_get_lambda(spgo_solver* h, double* lambda) {
    if (!h || !lambda) return SPGO_E_ARG;
    if (!h->ctl || h->lambda_reset) {
        *lambda = h->opt.mode == SPGO_MODE_REFERENCE ? 0.0 : h->opt.lambda0;
        return SPGO_OK;
    }
    Control c;
    int rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    *lambda = c.lambda;
    return SPGO_OK;
}

int spgo_set_lambda(spgo_solver* h, double lambda) {
    if (!h || !(lambda >= 0.0)) return SPGO_E_ARG;
    if (!h->ctl) return h->fail(SPGO_E_STATE, "set_lambda needs a solver that has stepped once");
    Control c;
    int rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    c.lambda = lambda;
    h->lambda_reset = false;
    return check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
}

int spgo_linearize(spgo_solver* h, double lambda, double* H, double* g, double* cost) {
    if (!h || !(lambda >= 0.0)) return SPGO_E_ARG;
    int rc = run_step(h, 0);   // layout, state upload and linearisation of the current state
    if (rc) return rc;
    Control c;
    rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    const double lam_keep = c.lambda;
    c.lambda = lambda;
    rc = check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    if (h->prec == SPGO_PREC_MIXED)
        launch_form_system<float>(h->lay, h->bf, h->ctl, h->H, h->gvec, h->stream);
    else
        launch_form_system<double>(h->lay, h->bd, h->ctl, h->H, h->gvec, h->stream);
    rc = check_hip(h, hipStreamSynchronize(h->stream), "form system");
    if (rc) return rc;
    const int P = 6 * h->N;
    std::vector<double> Hp((size_t)h->Pp * h->Pp);
    rc = check_hip(h, hipMemcpy(Hp.data(), h->H, Hp.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    std::vector<double> gp(P);
    rc = check_hip(h, hipMemcpy(gp.data(), h->gvec, P * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    if (h->lay.frame_mode) {
        // the frame-coordinate system mapped to the incremental one at the
        // current global poses
        if ((rc = pull_state(h))) return rc;
        std::vector<double> G((size_t)h->F * 12), Hi((size_t)P * P), gi(P);
        global_chain(h->N, h->X.data(), G.data());
        frames_to_incremental(h->N, Hp.data(), h->Pp, gp.data(), G.data(), Hi.data(), gi.data());
        if (H) std::memcpy(H, Hi.data(), Hi.size() * sizeof(double));
        if (g) std::memcpy(g, gi.data(), P * sizeof(double));
    } else {
        if (H)
            for (int i = 0; i < P; ++i)
                for (int j = 0; j <= i; ++j) H[(size_t)i * P + j] = H[(size_t)j * P + i] = Hp[(size_t)i * h->Pp + j];
        if (g) std::memcpy(g, gp.data(), P * sizeof(double));
    }
    rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    if (cost) *cost = c.cost;
    c.lambda = lam_keep;
    return check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
}

int spgo_reduced_band(spgo_solver* h, double lambda, int32_t band_blocks, double* Hband, double* g, double* cost) {
    if (!h || !(lambda >= 0.0) || band_blocks < 0 || !Hband || !g) return SPGO_E_ARG;
    int rc = run_step(h, 0);
    if (rc) return rc;
    Control c;
    rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    const double lam_keep = c.lambda;
    c.lambda = lambda;
    rc = check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    if (h->prec == SPGO_PREC_MIXED)
        launch_form_system<float>(h->lay, h->bf, h->ctl, h->H, h->gvec, h->stream);
    else
        launch_form_system<double>(h->lay, h->bd, h->ctl, h->H, h->gvec, h->stream);
    const size_t nb = (size_t)h->N * (band_blocks + 1) * 36;
    double* Hb = nullptr;
    rc = check_hip(h, hipMalloc(&Hb, nb * sizeof(double)), "hipMalloc");
    if (rc) return rc;
    lp_launch_band_from_dense(h->N, band_blocks, h->H, h->Pp, Hb, h->stream);
    rc = check_hip(h, hipStreamSynchronize(h->stream), "reduced band");
    if (!rc) rc = check_hip(h, hipMemcpy(Hband, Hb, nb * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    (void)hipFree(Hb);
    if (!rc) rc = check_hip(h, hipMemcpy(g, h->gvec, (size_t)6 * h->N * sizeof(double), hipMemcpyDeviceToHost),
                            "hipMemcpy");
    if (!rc) rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    if (cost) *cost = c.cost;
    c.lambda = lam_keep;
    return check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
}

int spgo_profile_phases(spgo_solver* h, int32_t iterations, double ms[SPGO_N_PHASES]) {
    if (!h || !ms || iterations < 1) return SPGO_E_ARG;
    int rc = run_step(h, 0);
    if (rc) return rc;
    if (iterations > h->rec_cap) {
        destroy_graph(h);
        h->rec = h->mem.alloc<IterRecord>(iterations);
        if (!h->rec) return h->fail(SPGO_E_DEVICE, "device allocation failed");
        h->rec_cap = iterations;
    }
    hipEvent_t ev[SPGO_N_PHASES + 1];
    for (int i = 0; i <= SPGO_N_PHASES; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) return h->fail(SPGO_E_DEVICE, "hipEventCreate failed");
    double acc[SPGO_N_PHASES] = {0};
    hipStream_t s = h->stream;
    const bool mixed = h->prec == SPGO_PREC_MIXED;
    for (int it = 0; it < iterations && rc == SPGO_OK; ++it) {
        (void)hipEventRecord(ev[0], s);
        if (mixed) launch_form_system<float>(h->lay, h->bf, h->ctl, h->H, h->gvec, s);
        else launch_form_system<double>(h->lay, h->bd, h->ctl, h->H, h->gvec, s);
        (void)hipEventRecord(ev[1], s);
        bool factor = true;
        if (h->pcg) {
            rc = check_hip(h, (hipError_t)lp_solve_launch(h->lps, h->H, h->Pp, h->gvec, h->delta, h->net_params,
                                                          h->opt.pcg_tol, h->opt.pcg_max_iter, h->ctl, s),
                           "cooperative CG launch");
            if (!rc) rc = check_hip(h, hipMemcpyAsync(h->pcg_flag_host, &h->ctl->pcg_fail, sizeof(int),
                                                      hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
            if (!rc) rc = check_hip(h, hipStreamSynchronize(s), "CG");
            if (rc) break;
            factor = *h->pcg_flag_host != 0;
        }
        bool posed = false;
        if (!factor) {
            (void)hipEventRecord(ev[2], s);
        } else if (h->band) {
            // with cyclic reduction the pose update runs inside it, in this phase
            if (mixed) {
                const PoseStep ps = pose_step(h, h->bf);
                posed = launch_chol(h, &ps, s);
            } else {
                const PoseStep ps = pose_step(h, h->bd);
                posed = launch_chol(h, &ps, s);
            }
            (void)hipEventRecord(ev[2], s);
        } else {
            launch_cholesky(h->H, h->Linv, h->Pp, h->dense_act_off, h->dense_act, h->dense_nact.data(), h->ctl, s);
            (void)hipEventRecord(ev[2], s);
            launch_chol_solve(h->H, h->Linv, h->gvec, h->delta, h->Pp, h->dense_first, h->dense_act_off,
                              h->dense_act, h->solve_y, h->solve_sync, h->ctl, s);
        }
        (void)hipEventRecord(ev[3], s);
        if (!posed && mixed)
            launch_pose_update<float>(h->lay, make_camera(h), h->bf, h->ctl, h->delta, h->step_size, 0, s);
        else if (!posed)
            launch_pose_update<double>(h->lay, make_camera(h), h->bd, h->ctl, h->delta, h->step_size, 0, s);
        (void)hipEventRecord(ev[4], s);
        if (mixed) launch_assemble<float>(make_asm(h, 0), h->bf, s);
        else launch_assemble<double>(make_asm(h, 0), h->bd, s);
        (void)hipEventRecord(ev[5], s);
        if (mixed)
            launch_accept(h->bf[0].cost, h->bf[1].cost, h->lay, h->ctl, h->rec,
                          h->opt.mode == SPGO_MODE_REFERENCE, h->opt.lam_up, h->opt.lam_down, h->opt.lam_min,
                          h->opt.lam_max, 0, s);
        else
            launch_accept(h->bd[0].cost, h->bd[1].cost, h->lay, h->ctl, h->rec,
                          h->opt.mode == SPGO_MODE_REFERENCE, h->opt.lam_up, h->opt.lam_down, h->opt.lam_min,
                          h->opt.lam_max, 0, s);
        (void)hipEventRecord(ev[6], s);
        rc = check_hip(h, hipEventSynchronize(ev[6]), "phase timing");
        for (int p = 0; p < SPGO_N_PHASES && rc == SPGO_OK; ++p) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ev[p], ev[p + 1]);
            acc[p] += t;
        }
    }
    for (int i = 0; i <= SPGO_N_PHASES; ++i) (void)hipEventDestroy(ev[i]);
    if (rc) return rc;
    for (int p = 0; p < SPGO_N_PHASES; ++p) ms[p] = acc[p] / iterations;
    h->dev_newer = true;
    return SPGO_OK;
}

int spgo_set_preconditioner(spgo_solver* h, spgo_lpcg* net) {
    if (!h) return SPGO_E_ARG;
    h->net = net;
    h->layout_dirty = true;   // the CG band follows the network's
    destroy_graph(h);
    return SPGO_OK;
}

int spgo_problem_info(spgo_solver* h, spgo_layout_info* out) {
    if (!h || !out) return SPGO_E_ARG;
    HostLayout hl;
    int rc = build_any_layout(h, hl);
    if (rc) return rc;
    out->n_poses = h->N;
    out->n_slots = h->L;
    out->n_measurements = (int)h->ma.size();
    out->n_landmarks = (int)hl.lm_id.size();
    out->n_groups = (int)hl.grp_anchor.size();
    out->n_chunks = (int)hl.chunk.size();
    out->max_span = 0;
    for (size_t m = 0; m < h->ma.size(); ++m) out->max_span = std::max(out->max_span, h->mt[m] - h->ma[m]);
    out->coordinates = hl.frame_mode ? SPGO_COORD_FRAMES : SPGO_COORD_INCREMENTAL;
    out->n_anchors_used = hl.anchors_used;
    std::vector<int> fb;
    int depth = 0;
    out->factorization = factor_path(h, hl, fb, depth);
    if (h->opt.factorization == SPGO_FACTOR_PCG && out->factorization >= 0) out->factorization = SPGO_FACTOR_PCG;
    out->profile_depth = depth;
    out->reduced_size = reduced_size(h->N);
    return SPGO_OK;
}

int spgo_last_step_ms(spgo_solver* h, double* ms) {
    if (!h || !ms) return SPGO_E_ARG;
    *ms = h->last_ms;
    return SPGO_OK;
}

const char* spgo_last_error(spgo_solver* h) {
    if (!h) return g_create_error.c_str();
    return h->err.c_str();
}

int spgo_device_synchronize(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SPGO_E_DEVICE;
    return hipDeviceSynchronize() == hipSuccess ? SPGO_OK : SPGO_E_DEVICE;
}

}  // extern "C"
