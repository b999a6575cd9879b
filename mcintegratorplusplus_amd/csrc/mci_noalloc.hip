// mci_noalloc.hip — the allocation-free entry point mci_estimate_device: the standalone
// estimators of mci_estimate on arrays the caller already holds on the device, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).  The only memory beyond the
// caller's input and outputs is the general MJBlocker's copy for its in-place blocking (more than
// 64 values per component); it lives in `scratch` below, a fixed array in every device's memory,
// whose address on each device is looked up once with hipGetSymbolAddress.  The array's name is
// used only for that lookup: on the host it is a stand-in, not a device address.
#include "mci_host.hpp"
#include "mci_kernels.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

namespace
{
// MCI_ESTIMATE_DEVICE_MAX doubles (16 MiB), held by every device the library runs on until the
// process ends
__device__ __align__(256) double scratch[MCI_ESTIMATE_DEVICE_MAX];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_ptr[MAX_DEVICES] = {};

// this device's address of `scratch` (looked up on the first call on the device)
mci_status scratch_ptr(double **out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) { return mcih::fail(MCI_E_DEVICE, "device %d beyond the %d this library tracks", dev, MAX_DEVICES); }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) { return mcih::fail(MCI_E_DEVICE, "hipGetSymbolAddress returned no address for the scratch array"); }
        g_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_ptr[dev];
    return MCI_OK;
}
} // namespace

extern "C" {

mci_status mci_estimate_device(int estim, int64_t n, int ndim, const double *x, double *average, double *error, void *stream)
{
    if (!x || !average || !error || ndim < 1) { return mcih::fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    if (estim < MCI_ESTIM_NOOP || estim > MCI_ESTIM_MJBLOCKER) { return mcih::fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator."); }
    if (n < 1) { return mcih::fail(MCI_E_INVALID_ARGUMENT, "n must be at least 1"); }
    TRY(mcih::check_estimator(estim, n));
    if (n > (int64_t)MCI_ESTIMATE_DEVICE_MAX / ndim) {
        return mcih::fail(MCI_E_INVALID_ARGUMENT, "mci_estimate_device takes at most %d values (n * ndim), got %lld * %d", (int)MCI_ESTIMATE_DEVICE_MAX,
                          (long long)n, ndim);
    }
    double *scr = nullptr;
    TRY(scratch_ptr(&scr));
    mcid::EstimJob j;
    j.data = x;
    j.scratch = scr;  // [n][ndim], written before it is read (no clearing needed)
    j.n = n;
    j.W = 1;
    j.nobs = ndim;
    j.off = 0;
    j.estim = estim;
    j.scale = 1.0;
    j.active = nullptr;
    j.avg = average;
    j.err = error;
    HIPCHK(mcid::launch_estimate(j, static_cast<hipStream_t>(stream)));
    return MCI_OK;
}

} // extern "C"
