// spgo_reduced_band_noalloc: spgo_reduced_band without any device allocation,
// for callers that cannot allocate (code under graph capture, serving loops
// with a fixed memory budget).  The band is staged in a file-scope device
// array instead of a buffer taken per call; the solver must already hold its
// device state (stepped or linearised once), so the linearisation allocates
// nothing either.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/spgo.h"
#include "spgo_host.h"
#include "spgo_kernels.h"
#include "spgo_lpcg.h"

namespace {

// staging for the block band: N (W + 1) 36 doubles (8 MiB per device, held
// until the process ends)
constexpr long long SCRATCH_DOUBLES = 1 << 20;
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
double* g_scratch_ptr[MAX_DEVICES] = {};
std::once_flag g_scratch_once[MAX_DEVICES];
hipError_t g_scratch_err[MAX_DEVICES] = {};

// device address of this device's copy of `scratch`, looked up once per device
// with that device current (the array's name is not a device address in host code)
double* scratch_on(int dev, hipError_t& err) {
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = g_scratch_err[dev] == hipSuccess ? (double*)p : nullptr;
    });
    err = g_scratch_err[dev];
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int spgo_reduced_band_noalloc(spgo_solver* h, double lambda, int32_t band_blocks, double* Hband, double* g,
                              double* cost) {
    using namespace spgo;
    if (!h || !(lambda >= 0.0) || band_blocks < 0 || !Hband || !g) return SPGO_E_ARG;
    const long long need = (long long)h->N * (band_blocks + 1) * 36;
    if (need > SCRATCH_DOUBLES)
        return h->fail(SPGO_E_ARG, "spgo_reduced_band_noalloc: N (band_blocks + 1) 36 = " + std::to_string(need) +
                                       " doubles exceeds the scratch capacity of " +
                                       std::to_string(SCRATCH_DOUBLES));
    // everything the linearisation uses must exist already
    if (!h->dev_ready || h->layout_dirty || !h->ctl || !h->ctl_host || h->prec != h->opt.precision)
        return h->fail(SPGO_E_STATE, "spgo_reduced_band_noalloc needs a solver whose device state exists "
                                     "(step or linearise it once after writing measurements or options)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
        return h->fail(SPGO_E_DEVICE, "spgo_reduced_band_noalloc: no current device");
    hipError_t e;
    double* buf = scratch_on(dev, e);
    if (!buf) return check_hip(h, e, "hipGetSymbolAddress(scratch)");
    int rc = run_step(h, 0);   // state upload and linearisation: no allocation on a set-up solver
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
    lp_launch_band_from_dense(h->N, band_blocks, h->H, h->Pp, buf, h->stream);
    rc = check_hip(h, hipStreamSynchronize(h->stream), "reduced band");
    if (!rc) rc = check_hip(h, hipMemcpy(Hband, buf, need * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = check_hip(h, hipMemcpy(g, h->gvec, (size_t)6 * h->N * sizeof(double), hipMemcpyDeviceToHost),
                            "hipMemcpy");
    if (!rc) rc = check_hip(h, hipMemcpy(&c, h->ctl, sizeof(Control), hipMemcpyDeviceToHost), "hipMemcpy");
    if (rc) return rc;
    if (cost) *cost = c.cost;
    c.lambda = lam_keep;
    return check_hip(h, hipMemcpy(h->ctl, &c, sizeof(Control), hipMemcpyHostToDevice), "hipMemcpy");
}

}  // extern "C"
