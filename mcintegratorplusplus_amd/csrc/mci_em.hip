// mci_em.hip — the diagonal Gaussian-mixture fit on the device.  Every sum over samples is taken per
// workgroup of EM_BLOCK samples in a fixed order and the workgroups' sums are added in block order,
// so the fit is deterministic.  The arithmetic follows the host fit (mci_gmm_fit_diag): the same
// expressions, squared distances summed over coordinates in order (these are bit-identical to the
// host's), sums over samples in another order (these agree to rounding).
#include "mci_em.hpp"

#include <float.h>

namespace mcie
{
namespace
{
__device__ __forceinline__ double dist2(const double *x, int d, int64_t r, int64_t c)
{
    double e = 0.;
    for (int j = 0; j < d; ++j) {
        const double t = x[r * d + j] - x[c * d + j];
        e += t * t;
    }
    return e;
}

// pairwise tree over the workgroup (EM_BLOCK threads); thread 0 returns the sum
__device__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = EM_BLOCK / 2; o > 0; o >>= 1) {
        if (t < o) { sh[t] = sh[t] + sh[t + o]; }
        __syncthreads();
    }
    const double s = sh[0];
    __syncthreads();
    return s;
}

__global__ __launch_bounds__(EM_BLOCK) void k_em_dist(const double *x, int64_t n, int d, int64_t pick, double *d2, int first)
{
    const int64_t r = (int64_t)blockIdx.x * EM_BLOCK + threadIdx.x;
    if (r >= n) { return; }
    const double e = dist2(x, d, r, pick);
    d2[r] = first ? e : fmin(d2[r], e);
}

__global__ __launch_bounds__(EM_BLOCK) void k_em_potential(const double *x, int64_t n, int d, const double *d2, const int64_t *cand, int nc,
                                                           double *part)
{
    __shared__ double sh[EM_BLOCK];
    const int64_t r = (int64_t)blockIdx.x * EM_BLOCK + threadIdx.x;
    for (int l = 0; l < nc; ++l) {
        const double v = r < n ? fmin(d2[r], dist2(x, d, r, cand[l])) : 0.;
        const double s = block_sum(v, sh);
        if (threadIdx.x == 0) { part[(int64_t)blockIdx.x * EM_NCAND + l] = s; }
    }
}

__global__ void k_em_lc(const double *pi, const double *var, int K, int d, double *lc)
{
    const int k = threadIdx.x;
    if (k >= K) { return; }
    double ld = 0.;
    for (int j = 0; j < d; ++j) { ld += log(2. * M_PI * var[k * d + j]); }
    lc[k] = log(pi[k]) - 0.5 * ld;
}

// responsibilities of one sample into rs [K] (normalised), the host fit's expressions
__device__ void resp(const double *xr, int d, int K, const double *mu, const double *var, const double *lc, double *rs)
{
    double mx = -HUGE_VAL;
    for (int k = 0; k < K; ++k) {
        double q = 0.;
        for (int j = 0; j < d; ++j) {
            const double t = xr[j] - mu[k * d + j];
            q += t * t / var[k * d + j];
        }
        rs[k] = lc[k] - 0.5 * q;
        mx = fmax(mx, rs[k]);
    }
    double z = 0.;
    for (int k = 0; k < K; ++k) {
        rs[k] = exp(rs[k] - mx);
        z += rs[k];
    }
    for (int k = 0; k < K; ++k) { rs[k] = rs[k] / z; }
}

// the E step's sums of one workgroup: nk [K], s1 [K][d] (r x), s2 [K][d] (r x x), each over the
// workgroup's samples in order
__global__ __launch_bounds__(EM_BLOCK) void k_em_estep(const double *x, int64_t n, int d, int K, const double *mu, const double *var,
                                                       const double *lc, double *part)
{
    __shared__ double rsh[EM_BLOCK * EM_KMAX];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * EM_BLOCK;
    const int m = (int)((n - r0) < EM_BLOCK ? (n - r0) : EM_BLOCK);
    if (t < m) { resp(x + (r0 + t) * d, d, K, mu, var, lc, rsh + t * EM_KMAX); }
    __syncthreads();
    const int ne = K * (1 + 2 * d);
    double *out = part + (int64_t)blockIdx.x * ne;
    for (int e = t; e < ne; e += EM_BLOCK) {
        double s = 0.;
        if (e < K) {
            for (int q = 0; q < m; ++q) { s += rsh[q * EM_KMAX + e]; }
        }
        else if (e < K + K * d) {
            const int k = (e - K) / d, j = (e - K) % d;
            for (int q = 0; q < m; ++q) { s += rsh[q * EM_KMAX + k] * x[(r0 + q) * d + j]; }
        }
        else {
            const int k = (e - K - K * d) / d, j = (e - K - K * d) % d;
            for (int q = 0; q < m; ++q) {
                const double xv = x[(r0 + q) * d + j];
                s += rsh[q * EM_KMAX + k] * xv * xv;
            }
        }
        out[e] = s;
    }
}

// the M step (one workgroup): the blocks' sums in block order, then the host fit's update
__global__ __launch_bounds__(EM_BLOCK) void k_em_mstep(const double *part, int64_t nb, int64_t n, int d, int K, double *pi, double *mu,
                                                       double *var, double *lc)
{
    __shared__ double nk[EM_KMAX];
    const int t = threadIdx.x;
    const int ne = K * (1 + 2 * d);
    if (t < K) {
        double s = 0.;
        for (int64_t b = 0; b < nb; ++b) { s += part[b * ne + t]; }
        nk[t] = s + 1e-9;
        pi[t] = fmax(nk[t] / (double)n, 1e-12);
    }
    __syncthreads();
    for (int e = t; e < K * d; e += EM_BLOCK) {
        const int k = e / d;
        double s1 = 0., s2 = 0.;
        for (int64_t b = 0; b < nb; ++b) {
            s1 += part[b * ne + K + e];
            s2 += part[b * ne + K + K * d + e];
        }
        const double m1 = s1 / nk[k];
        mu[e] = m1;
        var[e] = fmax(s2 / nk[k] - m1 * m1, 0.) + 1e-6;
    }
    __syncthreads();
    if (t < K) {
        double ld = 0.;
        for (int j = 0; j < d; ++j) { ld += log(2. * M_PI * var[t * d + j]); }
        lc[t] = log(pi[t]) - 0.5 * ld;
    }
}

__global__ __launch_bounds__(EM_BLOCK) void k_em_loglik(const double *x, int64_t n, int d, int K, const double *mu, const double *var,
                                                        const double *lc, double *part)
{
    __shared__ double sh[EM_BLOCK];
    const int64_t r = (int64_t)blockIdx.x * EM_BLOCK + threadIdx.x;
    double v = 0.;
    if (r < n) {
        double q[EM_KMAX];
        double mx = -HUGE_VAL;
        for (int k = 0; k < K; ++k) {
            double e = 0.;
            for (int j = 0; j < d; ++j) {
                const double t = x[r * d + j] - mu[k * d + j];
                e += t * t / var[k * d + j];
            }
            q[k] = lc[k] - 0.5 * e;
            mx = fmax(mx, q[k]);
        }
        double z = 0.;
        for (int k = 0; k < K; ++k) { z += exp(q[k] - mx); }
        v = mx + log(z);
    }
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = s; }
}
} // namespace

hipError_t launch_em_dist(const double *x, int64_t n, int d, int64_t pick, double *d2, bool first, hipStream_t s)
{
    hipLaunchKernelGGL(k_em_dist, dim3((unsigned)em_nblocks(n)), dim3(EM_BLOCK), 0, s, x, n, d, pick, d2, first ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_em_potential(const double *x, int64_t n, int d, const double *d2, const int64_t *cand, int nc, double *part,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_em_potential, dim3((unsigned)em_nblocks(n)), dim3(EM_BLOCK), 0, s, x, n, d, d2, cand, nc, part);
    return hipGetLastError();
}

hipError_t launch_em_lc(const double *pi, const double *var, int K, int d, double *lc, hipStream_t s)
{
    hipLaunchKernelGGL(k_em_lc, dim3(1), dim3(EM_KMAX), 0, s, pi, var, K, d, lc);
    return hipGetLastError();
}

hipError_t launch_em_iteration(const double *x, int64_t n, int d, int K, double *pi, double *mu, double *var, double *lc, double *part,
                               hipStream_t s)
{
    const int64_t nb = em_nblocks(n);
    hipLaunchKernelGGL(k_em_estep, dim3((unsigned)nb), dim3(EM_BLOCK), 0, s, x, n, d, K, mu, var, lc, part);
    hipLaunchKernelGGL(k_em_mstep, dim3(1), dim3(EM_BLOCK), 0, s, part, nb, n, d, K, pi, mu, var, lc);
    return hipGetLastError();
}

hipError_t launch_em_loglik(const double *x, int64_t n, int d, int K, const double *mu, const double *var, const double *lc, double *part,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_em_loglik, dim3((unsigned)em_nblocks(n)), dim3(EM_BLOCK), 0, s, x, n, d, K, mu, var, lc, part);
    return hipGetLastError();
}
} // namespace mcie
