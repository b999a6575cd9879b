// mci_host.hpp — host-side helpers shared by the C-ABI translation units: the thread-local
// error message, status macros, a grow-only device buffer and the deterministic sum order
// used for every cross-walker reduction.
#pragma once

#include "../../include/mci_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

namespace mcih
{
inline thread_local std::string g_err;

inline mci_status fail(mci_status code, const char *fmt, ...)
{
    char buf[600];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// device buffer that only grows
template <class T>
struct DBuf
{
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    ~DBuf() { if (p) { (void)hipFree(p); } }
    hipError_t ensure(size_t m)
    {
        if (m <= n && p) { return hipSuccess; }
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(m, 1) * sizeof(T));
        if (e == hipSuccess) { n = m; }
        return e;
    }
};

inline bool is_pow2(int64_t n) { return n > 0 && (((uint64_t)n & ((uint64_t)n - 1)) == 0); }

// the estimator's own size requirements (Estimators.cpp:39,85,127,197; MJBlocker.cpp:25-30),
// checked before any device work
inline mci_status check_estimator(int estim, int64_t n)
{
    switch (estim) {
    case MCI_ESTIM_UNCORRELATED:
        if (n < 2) { return fail(MCI_E_INVALID_ARGUMENT, "[UncorrelatedEstimator] n must be larger than 1"); }
        break;
    case MCI_ESTIM_FCBLOCKER:
        if (n < 50) { return fail(MCI_E_INVALID_ARGUMENT, "[FCBlockerEstimator] n must be >= 50"); }
        break;
    case MCI_ESTIM_MJBLOCKER:
        if (!is_pow2(n)) { return fail(MCI_E_INVALID_ARGUMENT, "[MJBlocker] ndata must be a power of two."); }
        break;
    case MCI_ESTIM_CORRELATED:
        if (!is_pow2(n) && n < 50) { return fail(MCI_E_INVALID_ARGUMENT, "[FCBlockerEstimator] n must be >= 50"); }
        break;
    }
    return MCI_OK;
}

// the deterministic walker sum used on the device (chunks of 256, pairwise tree, chunks in order)
inline double det_sum(const double *v, int64_t n)
{
    double total = 0.0;
    bool first = true;
    for (int64_t c0 = 0; c0 < n; c0 += 256) {
        double buf[256];
        for (int i = 0; i < 256; ++i) { buf[i] = (c0 + i < n) ? v[c0 + i] : 0.0; }
        for (int s = 128; s > 0; s >>= 1) {
            for (int i = 0; i < s; ++i) { buf[i] = buf[i] + buf[i + s]; }
        }
        if (first) { total = buf[0]; first = false; }
        else { total = total + buf[0]; }
    }
    return total;
}
} // namespace mcih

#define HIPCHK(expr)                                                                                     \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) { return mcih::fail(MCI_E_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); } \
    } while (0)
#define TRY(expr)                                                                                        \
    do {                                                                                                 \
        mci_status _s = (expr);                                                                          \
        if (_s != MCI_OK) { return _s; }                                                                 \
    } while (0)
