// mci_em.hpp — the diagonal Gaussian-mixture fit on the device (greedy k-means++ seeding and EM in
// FP64), the start of the learned proposal's network.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcie
{
constexpr int EM_BLOCK = 128;   // samples per workgroup
constexpr int EM_KMAX = 16;     // components
constexpr int EM_DMAX = 128;    // coordinates
constexpr int EM_NCAND = 16;    // seeding candidates per seed

// d2[r] = |x_r - x_pick|^2 (first) or min(d2[r], that)
hipError_t launch_em_dist(const double *x, int64_t n, int d, int64_t pick, double *d2, bool first, hipStream_t s);
// part[b][l] = sum over block b of min(d2[r], |x_r - x_cand_l|^2), l < nc
hipError_t launch_em_potential(const double *x, int64_t n, int d, const double *d2, const int64_t *cand, int nc, double *part,
                               hipStream_t s);
// lc[k] = log pi_k - sum_j log(2 pi var_kj) / 2
hipError_t launch_em_lc(const double *pi, const double *var, int K, int d, double *lc, hipStream_t s);
// one EM iteration: responsibilities and the per-block sums nk [K], s1 [K][d], s2 [K][d] into part
// [nblocks][K (1 + 2 d)], then the update of pi, mu, var and lc from their sums in block order
hipError_t launch_em_iteration(const double *x, int64_t n, int d, int K, double *pi, double *mu, double *var, double *lc, double *part,
                               hipStream_t s);
// part[b] = sum over block b of log sum_k exp(lc_k - sum_j (x_rj - mu_kj)^2 / var_kj / 2)
hipError_t launch_em_loglik(const double *x, int64_t n, int d, int K, const double *mu, const double *var, const double *lc, double *part,
                            hipStream_t s);
inline int64_t em_nblocks(int64_t n) { return (n + EM_BLOCK - 1) / EM_BLOCK; }
} // namespace mcie
