// Host side of the frame-coordinate reduced system: the device layout of the
// measurements (landmark groups with their anchor and target frames, the plan
// of the system's blocks) and the map back to the reference's incremental-pose
// system for spgo_linearize.
//
// A measurement (a, t, l) with t > a couples the frames a and t only
// (k_assemble_frames.hip), so the layout needs no bound on t - a
// (Solver_cuda.cu:248-309 accepts any target after the anchor).  A group is
// the landmarks of one anchor frame whose distinct target frames together
// number at most kt (16, 32 or 64, the kernel shape), within one LDS chunk.
// Per group the device keeps, in double and world axes, the target-frame block
// B (6m x 6m, lower blocks), and in Rz the anchor's blocks R(s) = sum_s' B(s, s')
// (row sums of B by target frame) after their total, then the gradient total
// and the gradient b (6m).
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "spgo_host.h"
#include "spgo_kernels.h"

namespace spgo {

int build_layout_frames(spgo_solver* h, HostLayout& out) {
    const int O = (int)h->ma.size();
    const int N = h->N;
    HostLayout hl;
    hl.frame_mode = 1;
    std::vector<int> order(O);
    for (int i = 0; i < O; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (h->ml[x] != h->ml[y]) return h->ml[x] < h->ml[y];
        return h->mt[x] < h->mt[y];
    });
    struct Lm {
        int id, anchor, begin, count, m, last;
    };
    std::vector<Lm> lms;
    int mmax = 0;
    for (int i = 0; i < O;) {
        const int l = h->ml[order[i]];
        const int a = h->ma[order[i]];
        int j = i, m = 0, prev = -1, last = a;
        while (j < O && h->ml[order[j]] == l) {
            if (h->ma[order[j]] != a)
                return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) +
                                                  " has measurements with different anchor frames");
            const int t = h->mt[order[j]];
            if (t > a && t != prev) {
                ++m;
                prev = t;
                last = t;
            }
            ++j;
        }
        if (m > KTF)
            return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) + " is observed in " + std::to_string(m) +
                                              " frames after its anchor; this build supports at most " +
                                              std::to_string(KTF));
        if (j - i > CHUNK_OBS)
            return h->fail(SPGO_E_LAYOUT, "landmark " + std::to_string(l) + " has more than " +
                                              std::to_string(CHUNK_OBS) + " measurements");
        lms.push_back({l, a, i, j - i, m, last});
        mmax = std::max(mmax, m);
        i = j;
    }
    const int kt = mmax <= 16 ? 16 : (mmax <= 32 ? 32 : KTF);
    int cl = frames_landmarks_per_group(kt, h->opt.precision != SPGO_PREC_MIXED);
    if (h->opt.chunk_landmarks > 0) cl = std::min(cl, h->opt.chunk_landmarks);
    const int SO = kt + 2;
    hl.kt = kt;
    hl.cl = cl;
    hl.kmax = kt;
    // landmarks of one anchor with similar frames together: by last target frame
    std::stable_sort(lms.begin(), lms.end(), [](const Lm& x, const Lm& y) {
        if (x.anchor != y.anchor) return x.anchor < y.anchor;
        if (x.last != y.last) return x.last < y.last;
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
    std::vector<int> per_anchor(h->F, 0);
    auto targets_of = [&](const Lm& L, std::vector<int>& t) {
        t.clear();
        for (int r = 0; r < L.count; ++r) {
            const int tt = h->mt[order[L.begin + r]];
            if (tt > L.anchor && (t.empty() || t.back() != tt)) t.push_back(tt);
        }
    };
    std::vector<int> uni, tl, merged;
    int li = 0;
    while (li < n_act) {
        const int a = lms[li].anchor;
        const int lm0 = li;
        int nobs = 0;
        uni.clear();
        while (li < n_act && lms[li].anchor == a && li - lm0 < cl && nobs + lms[li].count <= CHUNK_OBS) {
            targets_of(lms[li], tl);
            merged.clear();
            std::set_union(uni.begin(), uni.end(), tl.begin(), tl.end(), std::back_inserter(merged));
            if ((int)merged.size() > kt) break;
            uni.swap(merged);
            nobs += lms[li].count;
            ++li;
        }
        const int g = (int)hl.grp_anchor.size();
        const int m = (int)uni.size();
        const int nl = li - lm0;
        const int ob0 = (int)hl.ob_t.size();
        auto slot_of = [&](int t) {
            if (t <= a) return 0;
            return 1 + (int)(std::lower_bound(uni.begin(), uni.end(), t) - uni.begin());
        };
        struct Ob {
            int slot, j, m, pos;
        };
        std::vector<Ob> obs;
        obs.reserve(nobs);
        int pos = 0;
        for (int j = 0; j < nl; ++j) {
            const Lm& L = lms[lm0 + j];
            hl.lm_id[lm0 + j] = L.id;
            hl.lm_pos[lm0 + j] = make_int2(pos, L.count);
            for (int r = 0; r < L.count; ++r) {
                const int mm = order[L.begin + r];
                obs.push_back({slot_of(h->mt[mm]), j, mm, pos});
                ++pos;
            }
            // ray from the anchor cell's left pixel (Solver_cuda.cu:147); zero if never written
            auto it = h->cell_of.find((long long)a * h->L + L.id);
            hl.lm_ray[lm0 + j] = it != h->cell_of.end()
                                     ? make_float2(h->cells[6 * it->second], h->cells[6 * it->second + 1])
                                     : make_float2(0.f, 0.f);
        }
        std::stable_sort(obs.begin(), obs.end(), [](const Ob& x, const Ob& y) {
            if (x.slot != y.slot) return x.slot < y.slot;
            return x.j < y.j;
        });
        const size_t base = hl.chunk_soff.size();
        hl.chunk_soff.resize(base + SO, nobs);
        for (int sl = 0, q = 0; sl < SO; ++sl) {
            while (q < nobs && obs[q].slot < sl) ++q;
            hl.chunk_soff[base + sl] = q;
        }
        for (const Ob& o : obs) {
            hl.ob_t.push_back(h->mt[o.m]);
            hl.ob_meta.push_back((uint32_t)o.pos | ((uint32_t)o.j << 16));
            const int cell = h->cell_of.at((long long)h->mt[o.m] * h->L + h->ml[o.m]);
            const float* c = &h->cells[6 * cell];
            hl.ob_uv.push_back(make_float4(c[0], c[1], c[3], c[4]));
        }
        const int ci = (int)hl.chunk.size();
        hl.chunk.push_back(make_int4(lm0, nl, ob0, nobs));
        hl.grp_anchor.push_back(a);
        hl.grp_lm_begin.push_back(lm0);
        hl.grp_chunks.push_back(make_int2(ci, ci + 1));
        hl.grp_K.push_back(m);
        hl.grp_v_off.push_back(hl.v_total);
        hl.grp_D_off.push_back(hl.D_total);
        hl.grp_d_off.push_back((int)hl.d_total);
        hl.grp_r_off.push_back(hl.Rz_total);
        hl.v_total += (long long)nl * 6 * m;
        hl.D_total += (long long)(6 * m) * (6 * m);
        hl.d_total += 6 * m;
        hl.Rz_total += 36LL * (m + 1) + 6 + 6LL * m;
        const size_t fb = hl.grp_frames.size();
        hl.grp_frames.resize(fb + KTF + 1, -1);
        hl.grp_frames[fb] = a;
        for (int s = 0; s < m; ++s) hl.grp_frames[fb + 1 + s] = uni[s];
        per_anchor[a] += 1;
        (void)g;
    }
    for (int a = 0; a < h->F; ++a) {
        hl.anchor_grp[a + 1] = hl.anchor_grp[a] + per_anchor[a];
        if (per_anchor[a]) hl.anchors_used++;
    }
    const int G = (int)hl.grp_anchor.size();
    hl.grp_desc.resize(G);
    for (int g = 0; g < G; ++g) {
        GroupDesc& d = hl.grp_desc[g];
        d = GroupDesc{};
        d.anchor = d.a_loc = hl.grp_anchor[g];
        d.K = hl.grp_K[g];
        d.chunk_begin = hl.grp_chunks[g].x;
        d.chunk_end = hl.grp_chunks[g].y;
        d.lm_begin = hl.grp_lm_begin[g];
        d.d_off = hl.grp_d_off[g];
        d.v_off = hl.grp_v_off[g];
        d.D_off = hl.grp_D_off[g];
        d.r_off = hl.grp_r_off[g];
        d.chunk0 = hl.chunk[d.chunk_begin];
        for (int sl = 0; sl < KMAX + 2; ++sl)
            d.soff0[sl] = hl.chunk_soff[(size_t)d.chunk_begin * SO + std::min(sl, SO - 1)];
    }
    // plan of the reduced system: block (r, c), c <= r < N, sums the entries
    // (s_r, s_c) of every group holding frames r = f_{s_r} and c = f_{s_c}
    struct Ent {
        int r, c, g, sr, sc;
    };
    std::vector<Ent> ent;
    for (int g = 0; g < G; ++g) {
        const int* fr = &hl.grp_frames[(size_t)g * (KTF + 1)];
        const int m = hl.grp_K[g];
        for (int sr = 0; sr <= m; ++sr) {
            if (fr[sr] >= N) continue;
            for (int sc = 0; sc <= sr; ++sc) ent.push_back({fr[sr], fr[sc], g, sr, sc});
        }
    }
    std::stable_sort(ent.begin(), ent.end(), [](const Ent& x, const Ent& y) {
        if (x.r != y.r) return x.r < y.r;
        if (x.c != y.c) return x.c < y.c;
        return x.g < y.g;
    });
    hl.zp_row.assign(N + 1, 0);
    hl.zp_con.reserve(ent.size());
    hl.fpose.resize(N);
    for (int r = 0; r < N; ++r) hl.fpose[r] = std::max(r - 1, 0);   // the damping couples r - 1
    size_t q = 0;
    for (int r = 0; r < N; ++r) {
        hl.zp_row[r] = (int)hl.zp_col.size();
        while (q < ent.size() && ent[q].r == r) {
            const int c = ent[q].c;
            hl.zp_col.push_back(c);
            hl.zp_cbeg.push_back((int)hl.zp_con.size());
            hl.fpose[r] = std::min(hl.fpose[r], c);
            for (; q < ent.size() && ent[q].r == r && ent[q].c == c; ++q) {
                const Ent& e = ent[q];
                const int m = hl.grp_K[e.g], K6 = 6 * m;
                const long long ro = hl.grp_r_off[e.g];
                ZCon z;
                if (e.sc > 0) {          // two target frames: a block of B
                    z.off = hl.grp_D_off[e.g] + (long long)(6 * (e.sr - 1)) * K6 + 6 * (e.sc - 1);
                    z.K6 = K6;
                    z.doff = e.sr == e.sc ? (int)(ro + 36LL * (m + 1) + 6 + 6 * (e.sr - 1)) : -1;
                } else if (e.sr > 0) {   // target frame against the anchor: -R(s)
                    z.off = ro + 36LL * e.sr;
                    z.K6 = -1;
                    z.doff = -1;
                } else {                 // the anchor itself
                    z.off = ro;
                    z.K6 = -2;
                    z.doff = (int)(ro + 36LL * (m + 1));
                }
                hl.zp_con.push_back(z);
            }
        }
    }
    hl.zp_row[N] = (int)hl.zp_col.size();
    hl.zp_cbeg.push_back((int)hl.zp_con.size());
    for (int r = 0; r < N; ++r) hl.band_need = std::max(hl.band_need, r - hl.fpose[r]);
    // the incremental kernels' covering lists are not used in these coordinates
    hl.fs_begin.assign(N + 1, 0);
    out = std::move(hl);
    return SPGO_OK;
}

// H = U^T Z U, g = U^T gz with z = U delta, z_k = sum_{i >= k} E_i delta_i,
// E_i = Ad(G_i): H_ij = E_i^T (sum_{k <= i, l <= j} Z_kl) E_j and
// g_i = E_i^T sum_{k <= i} gz_k.
void frames_to_incremental(int N, const double* Z, int ld, const double* gz, const double* G, double* H, double* g) {
    const int P = 6 * N;
    std::vector<double> M((size_t)P * P);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j <= i; ++j) M[(size_t)i * P + j] = M[(size_t)j * P + i] = Z[(size_t)i * ld + j];
    // prefix sums over column blocks, then over row blocks
    for (int i = 0; i < P; ++i)
        for (int j = 6; j < P; ++j) M[(size_t)i * P + j] += M[(size_t)i * P + j - 6];
    for (int i = 6; i < P; ++i)
        for (int j = 0; j < P; ++j) M[(size_t)i * P + j] += M[(size_t)(i - 6) * P + j];
    std::vector<double> E((size_t)N * 36);
    for (int k = 0; k < N; ++k) {
        const double* T = G + 12 * k;
        double* A = &E[(size_t)k * 36];
        // Ad = [[R, 0], [[t]x R, R]]
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const double R = T[4 * r + c];
                A[6 * r + c] = R;
                A[6 * r + c + 3] = 0.0;
                A[6 * (r + 3) + c + 3] = R;
            }
        const double t0 = T[3], t1 = T[7], t2 = T[11];
        for (int c = 0; c < 3; ++c) {
            const double R0 = T[c], R1 = T[4 + c], R2 = T[8 + c];
            A[6 * 3 + c] = t1 * R2 - t2 * R1;
            A[6 * 4 + c] = t2 * R0 - t0 * R2;
            A[6 * 5 + c] = t0 * R1 - t1 * R0;
        }
    }
    for (int bi = 0; bi < N; ++bi)
        for (int bj = 0; bj < N; ++bj) {
            const double* Ei = &E[(size_t)bi * 36];
            const double* Ej = &E[(size_t)bj * 36];
            double t[36];
            for (int p = 0; p < 6; ++p)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int x = 0; x < 6; ++x) s += M[(size_t)(6 * bi + x) * P + 6 * bj + c] * Ei[6 * x + p];
                    t[6 * p + c] = s;
                }
            for (int p = 0; p < 6; ++p)
                for (int c = 0; c < 6; ++c) {
                    double s = 0.0;
                    for (int x = 0; x < 6; ++x) s += t[6 * p + x] * Ej[6 * x + c];
                    H[(size_t)(6 * bi + p) * P + 6 * bj + c] = s;
                }
        }
    std::vector<double> cum(6, 0.0);
    for (int bi = 0; bi < N; ++bi) {
        for (int x = 0; x < 6; ++x) cum[x] += gz[6 * bi + x];
        const double* Ei = &E[(size_t)bi * 36];
        for (int p = 0; p < 6; ++p) {
            double s = 0.0;
            for (int x = 0; x < 6; ++x) s += Ei[6 * x + p] * cum[x];
            g[6 * bi + p] = s;
        }
    }
}

}  // namespace spgo
