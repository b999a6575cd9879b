// mci_mlp.hip — BF16 MFMA GEMM (FP32 accumulation) with fused bias/SiLU epilogue, the
// Gaussian-mixture likelihood head, SiLU backward, row sums and Adam.
//
// GEMM tiling: a 256-thread workgroup computes a 64 x 64 tile of C as four 32 x 32 wave tiles
// with v_mfma_f32_32x32x16_bf16.  Per 32-wide slice of the reduction dimension the A and B
// tiles are staged in LDS as [row][k] and [col][k] (k contiguous), so each lane's MFMA operand
// (8 consecutive k of one row or column) is one 16-byte LDS read.  Operand layout of the
// instruction (checked on the device): lane l supplies row/column l % 32 and k = 8 (l / 32) .. +7;
// accumulator element i of lane l is C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
// Loads are bounds-checked (zero fill) and stores guarded, so any M, N, K are valid.
#include "mci_mlp.hpp"

namespace mcim
{
namespace
{
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TILE = 64, KC = 32, KP = KC + 8; // LDS row stride: 40 bf16 = 80 bytes

__device__ __forceinline__ bf16_t f2bf(float f)
{
    // round to nearest even (finite inputs)
    const uint32_t u = __float_as_uint(f);
    return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

// split z = blockIdx.z takes the reduction range [z kchunk, min(K, (z + 1) kchunk)) and writes
// C + z zstride (kchunk a multiple of KC; one split: kchunk >= K, zstride 0)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16_t *__restrict__ A, const bf16_t *__restrict__ B,
                                              const float *__restrict__ bias, float *__restrict__ C, int ldc, bf16_t *__restrict__ H,
                                              int ldh, int kchunk, int64_t zstride)
{
    __shared__ __attribute__((aligned(16))) bf16_t As[TILE][KP];
    __shared__ __attribute__((aligned(16))) bf16_t Bs[TILE][KP];
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
    const int kb = (int)blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    C += (int64_t)blockIdx.z * zstride;
    f32x16 acc = {};
    for (int k0 = kb; k0 < ke; k0 += KC) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid * 8 + q; // 0 .. 2047
            int mm, kk;
            if (TA) { kk = e / TILE; mm = e % TILE; }
            else { mm = e / KC; kk = e % KC; }
            const int gm = m0 + mm, gk = k0 + kk;
            bf16_t v = 0;
            if (gm < M && gk < ke) { v = TA ? A[(int64_t)gk * M + gm] : A[(int64_t)gm * K + gk]; }
            As[mm][kk] = v;
            int nn;
            if (TB) { nn = e / KC; kk = e % KC; }
            else { kk = e / TILE; nn = e % TILE; }
            const int gn = n0 + nn, gk2 = k0 + kk;
            bf16_t u = 0;
            if (gn < N && gk2 < ke) { u = TB ? B[(int64_t)gn * K + gk2] : B[(int64_t)gk2 * N + gn]; }
            Bs[nn][kk] = u;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KC; ks += 16) {
            const int kb = ks + 8 * (lane >> 5);
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + (lane & 31)][kb]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + (lane & 31)][kb]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int col = n0 + wn * 32 + (lane & 31);
    if (col >= N) { return; }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = m0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        if (row < M) {
            const float v = acc[i] + ((EPI != EPI_NONE) ? bias[row] : 0.f);
            C[(int64_t)row * ldc + col] = v;
            if (EPI == EPI_BIAS_SILU) { H[(int64_t)row * ldh + col] = f2bf(silu(v)); }
        }
    }
}

// C[r][c] = sum over the S splits of ws[z][r][c], in split order
__global__ void k_sum_splits(const float *ws, int S, int M, int N, float *C, int ldc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)M * N) { return; }
    float v = 0.f;
    for (int z = 0; z < S; ++z) { v += ws[(int64_t)z * M * N + e]; }
    C[(e / N) * ldc + e % N] = v;
}

__global__ void k_to_bf16(const float *x, bf16_t *xb, int64_t n, float scale)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { xb[i] = f2bf(x[i] * scale); }
}

// eight threads per sample (thread c of the group takes component c), feature-major layout
// (element c of sample r at out[c * ld + r]): -log of the mixture density of y under the sample's
// parameters and the gradient of (inv_n * that) with respect to the sample's raw outputs.  Output
// rows of one sample: for each component k, d mean rows then d log-scale rows (rows 2 d k ..
// 2 d k + 2 d - 1); then K logits.
__device__ __forceinline__ float group8_max(float v)
{
    v = fmaxf(v, __shfl_xor(v, 1));
    v = fmaxf(v, __shfl_xor(v, 2));
    return fmaxf(v, __shfl_xor(v, 4));
}

__device__ __forceinline__ float group8_sum(float v)
{
    v += __shfl_xor(v, 1);
    v += __shfl_xor(v, 2);
    return v + __shfl_xor(v, 4);
}

__global__ __launch_bounds__(256) void k_mixture_head(const float *out, int ld, const float *y, int n, int d, int K, int rows, float *loss,
                                                      float *dout, bf16_t *doutb, float inv_n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = t >> 3, k = t & 7;
    const bool ok = r < n;                // whole groups of eight agree
    const bool on = ok && k < K;
    const int rr = ok ? r : n - 1;
    const float *o = out + rr;
    const float *yr = y + rr;
    float *g = dout + rr;
    bf16_t *gb = doutb + rr;
    constexpr float HALF_LOG_2PI = 0.91893853320467274f;
    const int L0 = 2 * d * K; // first logit row
    const float a = on ? o[(int64_t)(L0 + k) * ld] : -__builtin_huge_valf();
    const float amax = group8_max(a);
    const float lsa = amax + __logf(group8_sum(on ? __expf(a - amax) : 0.f));
    float c = -__builtin_huge_valf();
    if (on) {
        c = a - lsa;
        for (int j = 0; j < d; ++j) {
            const float mu = MU_SCALE * o[(int64_t)(2 * d * k + j) * ld];
            const float ls = fminf(fmaxf(o[(int64_t)(2 * d * k + d + j) * ld], -LS_MAX), LS_MAX);
            const float z = (yr[(int64_t)j * ld] - mu) * __expf(-ls);
            c += -ls - HALF_LOG_2PI - 0.5f * z * z;
        }
    }
    const float cmax = group8_max(c);
    const float L = cmax + __logf(group8_sum(on ? __expf(c - cmax) : 0.f));
    if (ok && k == 0) { loss[r] = -L; }
    if (on) {
        const float gam = __expf(c - L);
        const float ga = inv_n * (__expf(a - lsa) - gam);
        g[(int64_t)(L0 + k) * ld] = ga;
        gb[(int64_t)(L0 + k) * ld] = f2bf(ga);
        for (int j = 0; j < d; ++j) {
            const int64_t im = (int64_t)(2 * d * k + j) * ld, il = (int64_t)(2 * d * k + d + j) * ld;
            const float mu = MU_SCALE * o[im];
            const float ol = o[il];
            const float ls = fminf(fmaxf(ol, -LS_MAX), LS_MAX);
            const float isg = __expf(-ls);
            const float z = (yr[(int64_t)j * ld] - mu) * isg;
            const float gm = inv_n * (-gam * z * isg) * MU_SCALE;
            const float gl = (fabsf(ol) < LS_MAX) ? inv_n * gam * (1.0f - z * z) : 0.0f;
            g[im] = gm;
            g[il] = gl;
            gb[im] = f2bf(gm);
            gb[il] = f2bf(gl);
        }
    }
    if (ok) {
        for (int cc = K * (1 + 2 * d) + k; cc < rows; cc += 8) {
            g[(int64_t)cc * ld] = 0.f;
            gb[(int64_t)cc * ld] = 0;
        }
    }
}

__global__ void k_silu_back(const float *dh, const float *z, bf16_t *dzb, float *dz, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    const float zz = z[i];
    const float s = 1.0f / (1.0f + __expf(-zz));
    const float v = dh[i] * (s * (1.0f + zz * (1.0f - s)));
    dz[i] = v;
    dzb[i] = f2bf(v);
}

// out[f] = sum of row f of a [rows][cols] (one wave per row, fixed order: lane partial sums in
// column order, then a tree over lanes)
__global__ void k_rowsum(const float *a, int rows, int cols, int lda, float *out)
{
    const int f = blockIdx.x;
    const int lane = threadIdx.x;
    if (f >= rows) { return; }
    float s = 0.f;
    for (int c = lane; c < cols; c += 64) { s += a[(int64_t)f * lda + c]; }
    for (int o = 32; o > 0; o >>= 1) { s += __shfl_down(s, o); }
    if (lane == 0) { out[f] = s; }
}

__global__ void k_adam(float *p, float *m, float *v, const float *g, bf16_t *copy, int64_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    if (copy) { copy[i] = f2bf(pi); }
}

dim3 grid1(int64_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

template <bool TA, bool TB>
hipError_t gemm_epi(int epi, dim3 g, hipStream_t s, int M, int N, int K, const bf16_t *A, const bf16_t *B, const float *bias, float *C,
                    int ldc, bf16_t *H, int ldh)
{
    const int kc = (int)g.z > 1 ? ((K + (int)g.z - 1) / (int)g.z + KC - 1) / KC * KC : K;
    const int64_t zs = (int)g.z > 1 ? (int64_t)M * N : 0;
    switch (epi) {
    case EPI_NONE: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_NONE>), g, dim3(256), 0, s, M, N, K, A, B, bias, C, ldc, H, ldh, kc, zs); break;
    case EPI_BIAS: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS>), g, dim3(256), 0, s, M, N, K, A, B, bias, C, ldc, H, ldh, kc, zs); break;
    default: hipLaunchKernelGGL((k_gemm<TA, TB, EPI_BIAS_SILU>), g, dim3(256), 0, s, M, N, K, A, B, bias, C, ldc, H, ldh, kc, zs); break;
    }
    return hipGetLastError();
}
} // namespace

hipError_t gemm_bf16(bool ta, bool tb, int epi, int M, int N, int K, const bf16_t *A, const bf16_t *B, const float *bias, float *C,
                     int ldc, bf16_t *H, int ldh, hipStream_t s)
{
    const dim3 g((unsigned)((N + TILE - 1) / TILE), (unsigned)((M + TILE - 1) / TILE));
    if (ta && tb) { return gemm_epi<true, true>(epi, g, s, M, N, K, A, B, bias, C, ldc, H, ldh); }
    if (ta) { return gemm_epi<true, false>(epi, g, s, M, N, K, A, B, bias, C, ldc, H, ldh); }
    if (tb) { return gemm_epi<false, true>(epi, g, s, M, N, K, A, B, bias, C, ldc, H, ldh); }
    return gemm_epi<false, false>(epi, g, s, M, N, K, A, B, bias, C, ldc, H, ldh);
}

hipError_t gemm_bf16_splitk(bool ta, bool tb, int M, int N, int K, const bf16_t *A, const bf16_t *B, float *C, int ldc, float *ws,
                            size_t ws_elems, hipStream_t s)
{
    const int nt = (N + TILE - 1) / TILE, mt = (M + TILE - 1) / TILE;
    int S = 1;
    while (S < 32 && nt * mt * S < 512 && K / (2 * S) >= 4 * KC && (size_t)(2 * S) * M * N <= ws_elems) { S *= 2; }
    if (S == 1) { return gemm_bf16(ta, tb, EPI_NONE, M, N, K, A, B, nullptr, C, ldc, nullptr, 0, s); }
    const dim3 g((unsigned)nt, (unsigned)mt, (unsigned)S);
    hipError_t e;
    if (ta && tb) { e = gemm_epi<true, true>(EPI_NONE, g, s, M, N, K, A, B, nullptr, ws, N, nullptr, 0); }
    else if (ta) { e = gemm_epi<true, false>(EPI_NONE, g, s, M, N, K, A, B, nullptr, ws, N, nullptr, 0); }
    else if (tb) { e = gemm_epi<false, true>(EPI_NONE, g, s, M, N, K, A, B, nullptr, ws, N, nullptr, 0); }
    else { e = gemm_epi<false, false>(EPI_NONE, g, s, M, N, K, A, B, nullptr, ws, N, nullptr, 0); }
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL(k_sum_splits, grid1((int64_t)M * N, 256), dim3(256), 0, s, ws, S, M, N, C, ldc);
    return hipGetLastError();
}

hipError_t launch_to_bf16_scaled(const float *x, bf16_t *xb, int64_t n, float scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf16, grid1(n, 256), dim3(256), 0, s, x, xb, n, scale);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float *x, bf16_t *xb, int64_t n, hipStream_t s) { return launch_to_bf16_scaled(x, xb, n, 1.0f, s); }

hipError_t launch_mixture_head(const float *out, int ld, const float *y, int n, int d, int K, int rows, float *loss, float *dout,
                               bf16_t *doutb, float inv_n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mixture_head, grid1((int64_t)n * 8, 256), dim3(256), 0, s, out, ld, y, n, d, K, rows, loss, dout, doutb, inv_n);
    return hipGetLastError();
}

hipError_t launch_silu_back(const float *dh, const float *z, bf16_t *dzb, int64_t n, hipStream_t s)
{
    // dz (FP32) overwrites dh in place
    hipLaunchKernelGGL(k_silu_back, grid1(n, 256), dim3(256), 0, s, dh, z, dzb, const_cast<float *>(dh), n);
    return hipGetLastError();
}

hipError_t launch_rowsum(const float *a, int rows, int cols, int lda, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned)rows), dim3(64), 0, s, a, rows, cols, lda, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *p, float *m, float *v, const float *g, bf16_t *copy, int64_t n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, hipStream_t s)
{
    hipLaunchKernelGGL(k_adam, grid1(n, 256), dim3(256), 0, s, p, m, v, g, copy, n, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}
} // namespace mcim
