// Host runtime of the learned-preconditioner PCG behind the spgo_lpcg_* C ABI.
//
// No reference counterpart (the reference solves the reduced system with a
// dense Cholesky, linear_solvers.cu:195-218).  A handle owns the network
// parameters with their Adam state and a loaded set of block-banded systems
// with everything the forward and backward passes keep; all arithmetic runs
// in the k_lp_* kernels on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spgo.h"
#include "spgo_host.h"
#include "spgo_lpcg.h"

using namespace spgo;

namespace {
thread_local std::string g_lpcg_create_error;

// splitmix64 stream and Box-Muller normals for the parameter initialisation
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};
}  // namespace

struct spgo_lpcg {
    int W = 0;
    std::string err;
    std::vector<float> host_params;
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    DevMem pmem;                 // parameters, gradient, Adam moments
    float *P = nullptr, *gP = nullptr, *mA = nullptr, *vA = nullptr;
    bool params_on_host_newer = true;
    int adam_step = 0;

    DevMem smem;                 // the loaded systems and the pass buffers
    bool loaded = false;
    LpSet set{};
    std::vector<int> sys_begin_h;
    double *Hb = nullptr, *g = nullptr, *C = nullptr, *N1 = nullptr, *sfeat = nullptr, *wsum = nullptr;
    double *Gd = nullptr, *Go = nullptr, *tnode = nullptr, *ldnode = nullptr, *Tsys = nullptr, *lsys = nullptr;
    int *node_sys = nullptr, *sys_begin = nullptr, *bad = nullptr;
    float *X = nullptr, *Z[LP_NLAYER + 1] = {}, *Hh[LP_NLAYER + 1] = {}, *M[LP_NLAYER] = {};
    float *OUT = nullptr, *dOUT = nullptr, *dH = nullptr, *dZ = nullptr, *dM = nullptr;
    float* part = nullptr;       // partial weight gradients (lp_wgrad_blocks(T) x (64 64 + 64))
    double *px = nullptr, *pr = nullptr, *pz = nullptr, *pp = nullptr, *pq = nullptr, *ptmp = nullptr;
    int* piters = nullptr;
    double* prel = nullptr;
    double last_ms[2] = {0.0, 0.0};   // preconditioner build, PCG

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace {

int lcheck(spgo_lpcg* h, hipError_t e, const char* what) {
    if (e == hipSuccess) return SPGO_OK;
    return h->fail(SPGO_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure_device(spgo_lpcg* h) {
    if (h->dev_ready) return SPGO_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return h->fail(SPGO_E_DEVICE, "no HIP device available");
    hipDeviceProp_t prop;
    int rc = lcheck(h, hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    if (rc) return rc;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return h->fail(SPGO_E_DEVICE, std::string("device 0 is ") + prop.gcnArchName +
                                          ", this library is built for gfx950");
    if ((rc = lcheck(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate"))) return rc;
    for (auto& e : h->ev)
        if ((rc = lcheck(h, hipEventCreate(&e), "hipEventCreate"))) return rc;
    h->P = h->pmem.alloc<float>(LP_NPARAM);
    h->gP = h->pmem.alloc<float>(LP_NPARAM);
    h->mA = h->pmem.alloc<float>(LP_NPARAM);
    h->vA = h->pmem.alloc<float>(LP_NPARAM);
    if (!h->P || !h->gP || !h->mA || !h->vA) return h->fail(SPGO_E_DEVICE, "device allocation failed");
    std::vector<float> z(LP_NPARAM, 0.f);
    if ((rc = lcheck(h, hipMemcpy(h->mA, z.data(), z.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")) ||
        (rc = lcheck(h, hipMemcpy(h->vA, z.data(), z.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")))
        return rc;
    h->dev_ready = true;
    return SPGO_OK;
}

int push_params(spgo_lpcg* h) {
    if (!h->params_on_host_newer) return SPGO_OK;
    int rc = lcheck(h, hipMemcpy(h->P, h->host_params.data(), LP_NPARAM * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    h->params_on_host_newer = false;
    return SPGO_OK;
}

const float* layer_ws(const float* P, int l) { return P + LP_OFF_L0 + l * LP_LAYER_SIZE; }
const float* layer_wm(const float* P, int l) { return layer_ws(P, l) + LP_HID * LP_HID; }
const float* layer_b(const float* P, int l) { return layer_wm(P, l) + LP_HID * LP_HID; }

// network forward over the loaded set: OUT = head(h_3, x)
void forward(spgo_lpcg* h) {
    LpNet b;
    b.X = h->X;
    for (int l = 0; l <= LP_NLAYER; ++l) {
        b.Z[l] = h->Z[l];
        b.Hh[l] = h->Hh[l];
        if (l < LP_NLAYER) b.M[l] = h->M[l];
    }
    b.OUT = h->OUT;
    b.sfeat = h->sfeat;
    b.wsum = h->wsum;
    lp_forward(h->set, h->P, b, h->stream);
}

// loss per system into lsys
void loss_terms(spgo_lpcg* h) {
    lp_launch_kap(h->set, h->OUT, h->N1, h->tnode, h->ldnode, h->Tsys, h->lsys, h->stream);
}

void backward(spgo_lpcg* h) {
    const int T = h->set.T;
    hipStream_t st = h->stream;
    float* G = h->gP;
    lp_launch_kap_dout(h->set, h->OUT, h->N1, h->Tsys, h->dOUT, st);
    lp_launch_wgrad(T, LP_NOUT, LP_HID, h->dOUT, h->Hh[LP_NLAYER], G + LP_OFF_WO, G + LP_OFF_BO, h->part, st);
    lp_launch_wgrad(T, LP_NOUT, LP_NIN, h->dOUT, h->X, G + LP_OFF_WK, nullptr, h->part, st);
    lp_launch_linearT(T, LP_NOUT, LP_HID, h->dOUT, h->P + LP_OFF_WO, h->dH, 0, st);
    for (int l = LP_NLAYER - 1; l >= 0; --l) {
        float* gl = G + LP_OFF_L0 + l * LP_LAYER_SIZE;
        lp_launch_reluback(T * LP_HID, h->dH, h->Z[l + 1], h->dZ, st);
        lp_launch_wgrad(T, LP_HID, LP_HID, h->dZ, h->Hh[l], gl, gl + 2 * LP_HID * LP_HID, h->part, st);
        lp_launch_wgrad(T, LP_HID, LP_HID, h->dZ, h->M[l], gl + LP_HID * LP_HID, nullptr, h->part, st);
        lp_launch_linearT(T, LP_HID, LP_HID, h->dZ, layer_wm(h->P, l), h->dM, 0, st);
        lp_launch_linearT(T, LP_HID, LP_HID, h->dZ, layer_ws(h->P, l), h->dH, 0, st);
        lp_launch_agg(h->set, h->sfeat, h->wsum, h->dM, h->dH, 1, st);
    }
    lp_launch_reluback(T * LP_HID, h->dH, h->Z[0], h->dZ, st);
    lp_launch_wgrad(T, LP_HID, LP_NIN, h->dZ, h->X, G + LP_OFF_WIN, G + LP_OFF_BIN, h->part, st);
}

int mean_loss(spgo_lpcg* h, double* loss) {
    const int S = h->set.S;
    std::vector<double> ls(S);
    int rc = lcheck(h, hipMemcpy(ls.data(), h->lsys, S * 8, hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    double s = 0.0;
    for (int i = 0; i < S; ++i) s += ls[i];
    *loss = s / S;
    return SPGO_OK;
}

}  // namespace

namespace spgo {

void lp_forward(const LpSet& s, const float* P, const LpNet& b, hipStream_t st) {
    const int T = s.T;
    lp_launch_linear(T, LP_HID, LP_NIN, b.X, P + LP_OFF_WIN, 0, nullptr, nullptr, P + LP_OFF_BIN, b.Z[0], b.Hh[0], st);
    for (int l = 0; l < LP_NLAYER; ++l) {
        lp_launch_agg(s, b.sfeat, b.wsum, b.Hh[l], b.M[l], 0, st);
        lp_launch_linear(T, LP_HID, LP_HID, b.Hh[l], layer_ws(P, l), LP_HID, b.M[l], layer_wm(P, l), layer_b(P, l),
                         b.Z[l + 1], b.Hh[l + 1], st);
    }
    lp_launch_linear(T, LP_NOUT, LP_HID, b.Hh[LP_NLAYER], P + LP_OFF_WO, LP_NIN, b.X, P + LP_OFF_WK, P + LP_OFF_BO,
                     b.OUT, nullptr, st);
}

int LpSolve::setup(int n_poses, int band_blocks, const std::vector<int>& far_beg_h, const std::vector<int>& far_col_h) {
    release();
    N = n_poses;
    W = band_blocks;
    auto take = [&](size_t bytes) -> void* {
        void* ptr = nullptr;
        if (hipMalloc(&ptr, bytes ? bytes : 8) != hipSuccess) return nullptr;
        ptrs.push_back(ptr);
        return ptr;
    };
    const size_t T = (size_t)N;
    Hb = (double*)take(T * (W + 1) * 36 * 8);
    C = (double*)take(T * 36 * 8);
    N1 = (double*)take(T * 36 * 8);
    sfeat = (double*)take(T * (W + 1) * 8);
    wsum = (double*)take(T * 8);
    Gd = (double*)take(T * 36 * 8);
    Go = (double*)take(T * 36 * 8);
    r = (double*)take(T * 6 * 8);
    z = (double*)take(T * 6 * 8);
    p = (double*)take(T * 6 * 8);
    q = (double*)take(T * 6 * 8);
    tmp = (double*)take(T * 6 * 8);
    bad = (int*)take(4);
    node_sys = (int*)take(T * 4);
    sys_begin = (int*)take(8);
    net.X = (float*)take(T * LP_NIN * 4);
    for (int l = 0; l <= LP_NLAYER; ++l) {
        net.Z[l] = (float*)take(T * LP_HID * 4);
        net.Hh[l] = (float*)take(T * LP_HID * 4);
        if (l < LP_NLAYER) net.M[l] = (float*)take(T * LP_HID * 4);
    }
    net.OUT = (float*)take(T * LP_NOUT * 4);
    net.sfeat = sfeat;
    net.wsum = wsum;
    // the cooperative PCG: one work item per unknown, at most all resident workgroups
    const int maxb = lp_coop_max_blocks();
    coop_blocks = std::min(maxb, (int)((6 * T + 255) / 256));
    part = (double*)take((size_t)4 * std::max(coop_blocks, 1) * 8);
    if (coop_blocks < 1) return (int)hipErrorLaunchFailure;
    if (ptrs.size() != 17 + 2 * (LP_NLAYER + 1) + LP_NLAYER + 1) return (int)hipErrorOutOfMemory;   // one failed
    std::vector<int> ns(T, 0);
    const int sb[2] = {0, N};
    hipError_t e = hipMemcpy(node_sys, ns.data(), T * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(sys_begin, sb, 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return (int)e;
    nfar = far_col_h.empty() ? 0 : (int)far_col_h.size();
    if (nfar > 0) {
        std::vector<int> rows(nfar);
        for (int i = 0; i < N; ++i)
            for (int k = far_beg_h[i]; k < far_beg_h[i + 1]; ++k) rows[k] = i;
        far_beg = (int*)take((size_t)(N + 1) * 4);
        far_col = (int*)take((size_t)nfar * 4);
        far_row = (int*)take((size_t)nfar * 4);
        far_val = (double*)take((size_t)nfar * 36 * 8);
        if (!far_beg || !far_col || !far_row || !far_val) return (int)hipErrorOutOfMemory;
        e = hipMemcpy(far_beg, far_beg_h.data(), (size_t)(N + 1) * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(far_col, far_col_h.data(), (size_t)nfar * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(far_row, rows.data(), (size_t)nfar * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) return (int)e;
    }
    set = LpSet{N, 1, W, Hb, nullptr, node_sys, sys_begin};
    return 0;
}

void LpSolve::release() {
    for (void* q2 : ptrs)
        if (q2) (void)hipFree(q2);
    ptrs.clear();
}

int lp_solve_launch(const LpSolve& b, const double* H, int Pp, const double* g, double* x, const float* P,
                    double tol, int maxit, Control* ctl, hipStream_t st) {
    LpSet s = b.set;
    s.g = g;
    lp_launch_band_from_dense(b.N, b.W, H, Pp, b.Hb, st);
    lp_launch_zero_flag(b.bad, st);
    lp_launch_features(s, b.C, b.sfeat, b.net.X, b.N1, b.wsum, b.bad, st);
    if (P) {
        lp_forward(s, P, b.net, st);
        lp_launch_buildG(s, b.net.OUT, b.C, b.Gd, b.Go, st);
    }
    if (b.nfar > 0) lp_launch_far_from_dense(b.nfar, b.far_row, b.far_col, H, Pp, b.far_val, st);
    return lp_launch_pcg_coop(b.N, b.W, b.Hb, P ? b.Gd : b.C, P ? b.Go : nullptr, g, x, b.r, b.z, b.p, b.q, b.tmp,
                              b.part, b.coop_blocks, tol, maxit, ctl, b.bad, b.nfar > 0 ? b.far_beg : nullptr,
                              b.far_col, b.far_val, st);
}

const float* lpcg_device_params(spgo_lpcg* h) {
    if (!h || ensure_device(h) || push_params(h) || hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
    return h->P;
}

int lpcg_band_blocks(const spgo_lpcg* h) { return h ? h->W : 0; }

}  // namespace spgo

extern "C" {

int spgo_lpcg_create(int32_t band_blocks, uint64_t seed, spgo_lpcg** out) {
    if (!out) return SPGO_E_ARG;
    *out = nullptr;
    if (band_blocks < 1 || band_blocks > LP_WMAX) {
        g_lpcg_create_error = "spgo_lpcg_create: band_blocks must be in [1, " + std::to_string(LP_WMAX) + "]";
        return SPGO_E_ARG;
    }
    spgo_lpcg* h = new spgo_lpcg();
    h->W = band_blocks;
    h->host_params.assign(LP_NPARAM, 0.f);
    Rng r{seed};
    const double s_in = std::sqrt(2.0 / LP_NIN), s_h = std::sqrt(1.0 / LP_HID);
    for (int k = 0; k < LP_HID * LP_NIN; ++k) h->host_params[LP_OFF_WIN + k] = (float)(s_in * r.normal());
    for (int l = 0; l < LP_NLAYER; ++l)
        for (int k = 0; k < 2 * LP_HID * LP_HID; ++k)
            h->host_params[LP_OFF_L0 + l * LP_LAYER_SIZE + k] = (float)(s_h * r.normal());
    *out = h;
    return SPGO_OK;
}

int spgo_lpcg_destroy(spgo_lpcg* h) {
    if (!h) return SPGO_OK;
    if (h->dev_ready) (void)hipStreamSynchronize(h->stream);
    h->smem.release();
    h->pmem.release();
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPGO_OK;
}

int spgo_lpcg_n_params(spgo_lpcg* h, int32_t* n) {
    if (!h || !n) return SPGO_E_ARG;
    *n = LP_NPARAM;
    return SPGO_OK;
}

int spgo_lpcg_get_params(spgo_lpcg* h, float* out) {
    if (!h || !out) return SPGO_E_ARG;
    if (!h->params_on_host_newer && h->dev_ready) {
        int rc = lcheck(h, hipMemcpy(h->host_params.data(), h->P, LP_NPARAM * 4, hipMemcpyDeviceToHost), "hipMemcpy");
        if (rc) return rc;
    }
    std::memcpy(out, h->host_params.data(), LP_NPARAM * 4);
    return SPGO_OK;
}

int spgo_lpcg_set_params(spgo_lpcg* h, const float* in) {
    if (!h || !in) return SPGO_E_ARG;
    std::memcpy(h->host_params.data(), in, LP_NPARAM * 4);
    h->params_on_host_newer = true;
    h->adam_step = 0;
    if (h->dev_ready) {
        std::vector<float> z(LP_NPARAM, 0.f);
        int rc;
        if ((rc = lcheck(h, hipMemcpy(h->mA, z.data(), z.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")) ||
            (rc = lcheck(h, hipMemcpy(h->vA, z.data(), z.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")))
            return rc;
    }
    return SPGO_OK;
}

int spgo_lpcg_load_systems(spgo_lpcg* h, int32_t n_systems, const int32_t* n_nodes, const double* Hband,
                           const double* g) {
    if (!h || n_systems <= 0 || !n_nodes || !Hband || !g) return SPGO_E_ARG;
    int rc = ensure_device(h);
    if (rc) return rc;
    std::vector<int> sb(n_systems + 1, 0), ns;
    for (int s = 0; s < n_systems; ++s) {
        if (n_nodes[s] <= 0) return h->fail(SPGO_E_ARG, "system " + std::to_string(s) + " has no nodes");
        sb[s + 1] = sb[s] + n_nodes[s];
        for (int i = 0; i < n_nodes[s]; ++i) ns.push_back(s);
    }
    const int T = sb[n_systems], W = h->W;
    (void)hipStreamSynchronize(h->stream);
    h->smem.release();
    h->loaded = false;
    DevMem& m = h->smem;
    h->Hb = m.alloc<double>((size_t)T * (W + 1) * 36);
    h->g = m.alloc<double>((size_t)T * 6);
    h->C = m.alloc<double>((size_t)T * 36);
    h->N1 = m.alloc<double>((size_t)T * 36);
    h->sfeat = m.alloc<double>((size_t)T * (W + 1));
    h->wsum = m.alloc<double>(T);
    h->Gd = m.alloc<double>((size_t)T * 36);
    h->Go = m.alloc<double>((size_t)T * 36);
    h->tnode = m.alloc<double>(T);
    h->ldnode = m.alloc<double>(T);
    h->Tsys = m.alloc<double>(n_systems);
    h->lsys = m.alloc<double>(n_systems);
    h->node_sys = m.alloc<int>(T);
    h->sys_begin = m.alloc<int>(n_systems + 1);
    h->bad = m.alloc<int>(1);
    h->X = m.alloc<float>((size_t)T * LP_NIN);
    for (int l = 0; l <= LP_NLAYER; ++l) {
        h->Z[l] = m.alloc<float>((size_t)T * LP_HID);
        h->Hh[l] = m.alloc<float>((size_t)T * LP_HID);
        if (l < LP_NLAYER) h->M[l] = m.alloc<float>((size_t)T * LP_HID);
    }
    h->OUT = m.alloc<float>((size_t)T * LP_NOUT);
    h->dOUT = m.alloc<float>((size_t)T * LP_NOUT);
    h->dH = m.alloc<float>((size_t)T * LP_HID);
    h->dZ = m.alloc<float>((size_t)T * LP_HID);
    h->dM = m.alloc<float>((size_t)T * LP_HID);
    h->part = m.alloc<float>((size_t)lp_wgrad_blocks(T) * (LP_HID * LP_HID + LP_HID));
    h->px = m.alloc<double>((size_t)T * 6);
    h->pr = m.alloc<double>((size_t)T * 6);
    h->pz = m.alloc<double>((size_t)T * 6);
    h->pp = m.alloc<double>((size_t)T * 6);
    h->pq = m.alloc<double>((size_t)T * 6);
    h->ptmp = m.alloc<double>((size_t)T * 6);
    h->piters = m.alloc<int>(n_systems);
    h->prel = m.alloc<double>(n_systems);
    if (!h->prel || !h->piters || !h->ptmp || !h->pq || !h->pp || !h->pz || !h->pr || !h->px || !h->part || !h->dM || !h->dZ ||
        !h->dH || !h->dOUT || !h->OUT || !h->X || !h->bad || !h->sys_begin || !h->node_sys || !h->lsys || !h->Tsys ||
        !h->ldnode || !h->tnode || !h->Go || !h->Gd || !h->wsum || !h->sfeat || !h->N1 || !h->C || !h->g || !h->Hb)
        return h->fail(SPGO_E_DEVICE, "device allocation failed");
    for (int l = 0; l <= LP_NLAYER; ++l)
        if (!h->Z[l] || !h->Hh[l] || (l < LP_NLAYER && !h->M[l])) return h->fail(SPGO_E_DEVICE, "device allocation failed");
    const int zero = 0;
    if ((rc = lcheck(h, hipMemcpy(h->Hb, Hband, (size_t)T * (W + 1) * 36 * 8, hipMemcpyHostToDevice), "hipMemcpy")) ||
        (rc = lcheck(h, hipMemcpy(h->g, g, (size_t)T * 6 * 8, hipMemcpyHostToDevice), "hipMemcpy")) ||
        (rc = lcheck(h, hipMemcpy(h->node_sys, ns.data(), (size_t)T * 4, hipMemcpyHostToDevice), "hipMemcpy")) ||
        (rc = lcheck(h, hipMemcpy(h->sys_begin, sb.data(), sb.size() * 4, hipMemcpyHostToDevice), "hipMemcpy")) ||
        (rc = lcheck(h, hipMemcpy(h->bad, &zero, 4, hipMemcpyHostToDevice), "hipMemcpy")))
        return rc;
    h->set = LpSet{T, n_systems, W, h->Hb, h->g, h->node_sys, h->sys_begin};
    h->sys_begin_h = sb;
    lp_launch_features(h->set, h->C, h->sfeat, h->X, h->N1, h->wsum, h->bad, h->stream);
    int bad = 0;
    if ((rc = lcheck(h, hipMemcpyAsync(&bad, h->bad, 4, hipMemcpyDeviceToHost, h->stream), "hipMemcpyAsync")) ||
        (rc = lcheck(h, hipStreamSynchronize(h->stream), "features")))
        return rc;
    if (bad) return h->fail(SPGO_E_ARG, "a diagonal block is not positive definite");
    h->loaded = true;
    return SPGO_OK;
}

int spgo_lpcg_loss_grad(spgo_lpcg* h, double* loss, float* grad) {
    if (!h || !h->loaded) return h ? h->fail(SPGO_E_STATE, "no systems loaded") : SPGO_E_ARG;
    int rc = push_params(h);
    if (rc) return rc;
    forward(h);
    loss_terms(h);
    backward(h);
    if ((rc = lcheck(h, hipStreamSynchronize(h->stream), "loss and gradient"))) return rc;
    if (loss && (rc = mean_loss(h, loss))) return rc;
    if (grad && (rc = lcheck(h, hipMemcpy(grad, h->gP, LP_NPARAM * 4, hipMemcpyDeviceToHost), "hipMemcpy"))) return rc;
    return SPGO_OK;
}

int spgo_lpcg_train(spgo_lpcg* h, int32_t steps, float lr, float* losses) {
    if (!h || steps < 0 || !(lr > 0.f)) return SPGO_E_ARG;
    if (!h->loaded) return h->fail(SPGO_E_STATE, "no systems loaded");
    int rc = push_params(h);
    if (rc) return rc;
    for (int s = 0; s < steps; ++s) {
        forward(h);
        loss_terms(h);
        backward(h);
        ++h->adam_step;
        lp_launch_adam(LP_NPARAM, h->P, h->gP, h->mA, h->vA, lr, 0.9f, 0.999f, 1e-8f, h->adam_step, h->stream);
        if (losses) {
            if ((rc = lcheck(h, hipStreamSynchronize(h->stream), "training step"))) return rc;
            double l = 0.0;
            if ((rc = mean_loss(h, &l))) return rc;
            losses[s] = (float)l;
        }
    }
    if ((rc = lcheck(h, hipStreamSynchronize(h->stream), "training"))) return rc;
    return SPGO_OK;
}

int spgo_lpcg_solve(spgo_lpcg* h, int32_t learned, double tol, int32_t max_iter, int32_t* iters, double* relres,
                    double* x) {
    if (!h || !(tol > 0.0) || max_iter <= 0) return SPGO_E_ARG;
    if (!h->loaded) return h->fail(SPGO_E_STATE, "no systems loaded");
    int rc = push_params(h);
    if (rc) return rc;
    hipStream_t st = h->stream;
    (void)hipEventRecord(h->ev[0], st);
    if (learned) {
        forward(h);
        lp_launch_buildG(h->set, h->OUT, h->C, h->Gd, h->Go, st);
    }
    (void)hipEventRecord(h->ev[1], st);
    if (learned)
        lp_launch_pcg(h->set, h->Gd, h->Go, h->px, h->pr, h->pz, h->pp, h->pq, h->ptmp, tol, max_iter, h->piters,
                      h->prel, st);
    else
        lp_launch_pcg(h->set, h->C, nullptr, h->px, h->pr, h->pz, h->pp, h->pq, h->ptmp, tol, max_iter, h->piters,
                      h->prel, st);
    (void)hipEventRecord(h->ev[2], st);
    if ((rc = lcheck(h, hipStreamSynchronize(st), "PCG"))) return rc;
    float a = 0.f, b = 0.f;
    (void)hipEventElapsedTime(&a, h->ev[0], h->ev[1]);
    (void)hipEventElapsedTime(&b, h->ev[1], h->ev[2]);
    h->last_ms[0] = a;
    h->last_ms[1] = b;
    const int S = h->set.S;
    if (iters && (rc = lcheck(h, hipMemcpy(iters, h->piters, S * 4, hipMemcpyDeviceToHost), "hipMemcpy"))) return rc;
    if (relres && (rc = lcheck(h, hipMemcpy(relres, h->prel, S * 8, hipMemcpyDeviceToHost), "hipMemcpy"))) return rc;
    if (x && (rc = lcheck(h, hipMemcpy(x, h->px, (size_t)h->set.T * 6 * 8, hipMemcpyDeviceToHost), "hipMemcpy")))
        return rc;
    return SPGO_OK;
}

int spgo_lpcg_last_ms(spgo_lpcg* h, double ms[2]) {
    if (!h || !ms) return SPGO_E_ARG;
    ms[0] = h->last_ms[0];
    ms[1] = h->last_ms[1];
    return SPGO_OK;
}

int spgo_lpcg_get_features(spgo_lpcg* h, float* X, double* C) {
    if (!h) return SPGO_E_ARG;
    if (!h->loaded) return h->fail(SPGO_E_STATE, "no systems loaded");
    int rc;
    if (X && (rc = lcheck(h, hipMemcpy(X, h->X, (size_t)h->set.T * LP_NIN * 4, hipMemcpyDeviceToHost), "hipMemcpy")))
        return rc;
    if (C && (rc = lcheck(h, hipMemcpy(C, h->C, (size_t)h->set.T * 36 * 8, hipMemcpyDeviceToHost), "hipMemcpy")))
        return rc;
    return SPGO_OK;
}

const char* spgo_lpcg_last_error(spgo_lpcg* h) {
    if (!h) return g_lpcg_create_error.c_str();
    return h->err.c_str();
}

}  // extern "C"
