// mci_lp.hpp — kernels of the learned-proposal sampler: a d-dimensional target that is a
// mixture of K full-covariance Gaussians, random-walk Metropolis (single-coordinate uniform
// moves, FP64 or FP32 chain) and the MH-corrected move drawn from the network's Gaussian
// mixture q(x' | x) (FP32).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcil
{
constexpr int WG = 256;   // walkers per workgroup
constexpr int KMAX = 8;   // target components (fewer are padded with inert ones)
constexpr int DMAX = 32;  // dimensions

template <class R>
struct Target
{
    int d;
    const R *mu;        // [KMAX][d]
    const R *P;         // [KMAX][d][d] precision matrices
    const R *U;         // [KMAX][d][d] upper Cholesky factors, P_k = U_k^T U_k
    const R *Um;        // [KMAX][d] U_k mu_k
    const R *v;         // [KMAX][d] P_k mu_k
    const R *c;         // [KMAX] log w_k - log det(2 pi Sigma_k)/2 (-inf for padding)
    const float *L;     // [KMAX][d][d] Cholesky factor of Sigma_k (exact start)
    const float *muf;   // [KMAX][d]
    const float *cw;    // [KMAX] cumulative weights
    int K;              // real components
    const float *Ut;    // [KMAX][d][32] U_k transposed, rows zero-padded to 32 (FP32; matrix-core operand)
};

// walker state (HBM, walker index fastest)
template <class R>
struct State
{
    int64_t W;
    R *x;               // [d][W]
    R *q;               // [KMAX][W] (x - mu_k)^T P_k (x - mu_k)
    R *m, *S;           // [W] log-sum-exp shift and sum: log p(x) = m + log S
    double *sq;         // [W] |x|^2
    R *sb, *cb;         // [2][W] block sums and Kahan compensations of |x|^2 and 1[x_0 > 0]
    int *jb, *bi;       // [W] steps in the current block, block index
    double *bmeans;     // [nblocks][2][W]
    int B;              // block length (0: no accumulation)
    int32_t *nacc;      // [W] accepted moves of the current phase
};

size_t rw_lds_bytes(int d, int esize);
template <class R>
hipError_t launch_lp_init(const Target<R> &T, const State<R> &S, uint32_t k0, uint32_t k1, hipStream_t s);
template <class R>
hipError_t launch_lp_rw(const Target<R> &T, const State<R> &S, R step, uint64_t t0, int64_t nsteps, uint32_t k0, uint32_t k1,
                        hipStream_t s);
// the network for the fused learned move: BF16 weights stored [out][in]: W1t [H][dp] with the input
// padded to dp = max(16, d) columns, W2t [H][H], W3t [Op][H] with Op = 2 d KMAX + 32 rows: component
// k at rows 2 d k + 2 j (mean of coordinate j) and 2 d k + 2 j + 1 (log-scale), zero rows for
// components k >= K, the K logits from row 2 d KMAX; FP32 biases in the same row order
struct NetDev
{
    int d, dp, H, K, O, Op;
    const uint16_t *W1t, *W2t, *W3t;
    const float *b1, *b2, *b3;
};
// the fused kernel exists for d in {8, 16, 32}, H in {64, 128}, 1 <= K <= KMAX
bool learned_supported(int d, int H, int K);
// nl consecutive learned moves with the conditional proposal q(x' | x) (the network at x), steps
// t0 .. t0 + nl - 1, for every walker (FP32 chain)
hipError_t launch_lp_learned(const NetDev &N, const Target<float> &T, const State<float> &S, uint64_t t0, int nl, uint32_t k0, uint32_t k1,
                             int32_t *nacc_ind, hipStream_t s);
// the independence proposal's mixture (device arrays): q(x) = sum_k exp(lc_k - |(x - mu_k) isig_k|^2 / 2)
struct IndepMix
{
    const float *mu;   // [KMAX][d] means
    const float *isig; // [KMAX][d] 1 / sigma
    const float *sig;  // [KMAX][d] sigma (the move's scale: x' = mu + sigma z)
    const float *lc;   // [KMAX] log pi_k + sum_j log(1 / sigma_kj) - d log(2 pi) / 2 (-inf past K)
    const float *cw;   // [KMAX] cumulative weights (1 from component K - 1 on)
    int K;
};
// nl consecutive independence moves, steps t0 .. t0 + nl - 1, for every walker (FP32 chain; d in
// {8, 16, 32})
hipError_t launch_lp_indep(const IndepMix &M, const Target<float> &T, const State<float> &S, uint64_t t0, int nl, uint32_t k0, uint32_t k1,
                           int32_t *nacc_ind, hipStream_t s);
// the target's state (q, m, S, |x|^2) at the positions S.x, empty block sums (FP32 chain)
hipError_t launch_lp_state(const Target<float> &T, const State<float> &S, hipStream_t s);
// the same kernel's network at positions xin [d][S.W], the mixture to out [K (1 + 2 d)][S.W]
// (per component d means and d log-scales, transformed; then K logits)
hipError_t launch_lp_network(const NetDev &N, const Target<float> &T, const State<float> &S, const float *xin, float *out, hipStream_t s);
// x [d][W] -> dst [d][W] FP32 (snapshots)
template <class R>
hipError_t launch_lp_snapshot(const State<R> &S, int d, float *dst, hipStream_t s);
// training pairs: column p of ctx/tgt [d][B] from snapshots [nsnap][d][W]; with probability
// same_frac the target is the same walker at another snapshot, else another walker (24-bit index);
// with fixed_ctx (device [d]) every input is that position instead
hipError_t launch_lp_pairs(const float *snap, int nsnap, int d, int64_t W, int B, uint64_t iter, uint32_t k0, uint32_t k1,
                           float same_frac, const float *fixed_ctx, float *ctx, float *tgt, hipStream_t s);
} // namespace mcil
