// mci_mlp.hpp — the learned proposal's network: an MLP d -> H -> H -> K (1 + 2 d) whose output
// parameterises a K-component diagonal Gaussian mixture, with BF16 MFMA forward and backward
// GEMMs (FP32 accumulation), the mixture log-likelihood head and Adam.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcim
{
typedef uint16_t bf16_t; // bfloat16 bits

// Activations are feature-major ([feature][sample]) so that one thread per sample reads and
// writes coalesced.  Output features of one sample, component-major: for each component k, d
// mean rows then d log-scale pre-activation rows (rows 2 d k .. 2 d k + 2 d - 1), then K logits.
constexpr float IN_SCALE = 0.25f;  // network input = x * IN_SCALE
constexpr float MU_SCALE = 4.0f;   // mean = MU_SCALE * output
constexpr float LS_MAX = 3.0f;     // log-scale = output clamped to [-LS_MAX, LS_MAX]

enum { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_SILU = 2 };

// C[M x N] = op(A)[M x K] op(B)[K x N] with A, B in BF16 and C in FP32.
//   TA: A is stored K x M (lda = M), else M x K (lda = K); TB: B is stored N x K (ldb = K),
//   else K x N (ldb = N).  EPI_BIAS adds bias[M] (one value per output row: activations are
//   feature-major); EPI_BIAS_SILU also writes H = bf16(silu(C)).
hipError_t gemm_bf16(bool ta, bool tb, int epi, int M, int N, int K, const bf16_t *A, const bf16_t *B, const float *bias,
                     float *C, int ldc, bf16_t *H, int ldh, hipStream_t s);
// the same without epilogue, the reduction split over up to 32 workgroups per output tile (for
// long reductions over few tiles: weight gradients over a minibatch); partial products in ws
// (ws_elems floats, S M N used), summed in split order (deterministic)
hipError_t gemm_bf16_splitk(bool ta, bool tb, int M, int N, int K, const bf16_t *A, const bf16_t *B, float *C, int ldc, float *ws,
                            size_t ws_elems, hipStream_t s);

// x [n][d] FP32 -> bf16(x * IN_SCALE) [n][d]
hipError_t launch_to_bf16_scaled(const float *x, bf16_t *xb, int64_t n, float scale, hipStream_t s);
hipError_t launch_to_bf16(const float *x, bf16_t *xb, int64_t n, hipStream_t s);

// feature-major [rows][ld] outputs and [d][ld] targets: per sample r < n, loss_r = -log q(y_r | out_r)
// and dOut_r = inv_n * d(loss_r)/d(out_r) (FP32 and BF16; rows beyond K (1 + 2 d) get 0)
hipError_t launch_mixture_head(const float *out, int ld, const float *y, int n, int d, int K, int rows, float *loss, float *dout,
                               bf16_t *doutb, float inv_n, hipStream_t s);

// dZ = dH * silu'(Z): FP32 in place of dH and BF16 copy
hipError_t launch_silu_back(const float *dh, const float *z, bf16_t *dzb, int64_t n, hipStream_t s);
// row sums of a [rows][cols] FP32 matrix (fixed order) -> out[rows]
hipError_t launch_rowsum(const float *a, int rows, int cols, int lda, float *out, hipStream_t s);
// Adam on FP32 parameters with BF16 copies (copy may be null)
hipError_t launch_adam(float *p, float *m, float *v, const float *g, bf16_t *copy, int64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t s);
} // namespace mcim
