// mci_lp.hip — the learned-proposal sampler's kernels.
//
// Target: p(x) = sum_k w_k N(x; mu_k, Sigma_k) in d dimensions with full covariances, held as
// precision matrices P_k and c_k = log w_k - log det(2 pi Sigma_k)/2, so that
// log p(x) = LSE_k (c_k - q_k/2) with q_k = (x - mu_k)^T P_k (x - mu_k).
//
// Random-walk Metropolis (k_lp_rw): single-coordinate uniform moves (SRRDVecMove, veclen 1).  A
// move x_j -> x_j + delta changes q_k by 2 delta (P_k (x - mu_k))_j + delta^2 (P_k)_jj, i.e. one row
// of each P_k against x: 32 K fused multiply-adds at d = 32, all operands in LDS (the P_k rows
// padded to d + 1 so that lanes at different j read different banks).  The test u <= p'/p is
// evaluated as u S < exp(m' - m) S' with log p = m + log S (no logarithm).  Every block end
// recomputes q_k from x.  The kernel runs in FP64 (the plain chain) or FP32.
//
// Learned move (FP32, k_lp_learned): the network maps the current x to a K-component diagonal
// Gaussian mixture q(. | x); propose x' ~ q(. | x), evaluate the network at x' and accept with
// min(1, p(x') q(x | x') / (p(x) q(x' | x))), so the target stays exact whatever the network
// computes.  Both forward passes (at x and at x') run inside the move's one launch.
//
// Random words (Philox4x32-10, key = seed, counter (c0, walker, step low, step high)):
//   c0 = 0            random-walk step: [0] coordinate (Lemire), [1] proposal, [2..3] u
//   c0 = 0x80000000+k Lemire rejection words
//   c0 = 0x10000000   learned move: [0] component, [1] u;  0x10000001 + n: polar-method normals
//   c0 = 0x40000000+n exact start draw (step 0): call 0 the component, then normals
#include "mci_lp.hpp"
#include "mci_f32math.hpp"
#include "mci_mlp.hpp"

#include <float.h>

namespace mcil
{
namespace
{
using mcif::bexp;
using mcif::blog;
using mcif::bsincos2pi;
using mcif::xor3;
using mcif::sym24;

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t o[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    o[0] = c0;
    o[1] = c1;
    o[2] = c2;
    o[3] = c3;
}

__device__ __forceinline__ float rexp(float x) { return __expf(x); }
__device__ __forceinline__ double rexp(double x) { return exp(x); }

template <class R>
__device__ __forceinline__ void kahan(R &s, R &c, R y)
{
    const R yy = y - c;
    const R tt = s + yy;
    c = (tt - s) - yy;
    s = tt;
}

__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Marsaglia's polar method: the next standard normal of this lane's stream (c0base + call)
struct Normals
{
    uint32_t o[4];
    int avail;    // unused attempts (pairs) left in o
    float spare;  // second normal of the last accepted pair
    bool has_spare;
    uint32_t call;
};

__device__ __forceinline__ float next_normal(Normals &N, uint32_t c0base, uint32_t id, uint32_t tl, uint32_t th, uint32_t k0, uint32_t k1)
{
    if (N.has_spare) {
        N.has_spare = false;
        return N.spare;
    }
    for (;;) {
        if (N.avail == 0) {
            philox(c0base + N.call, id, tl, th, k0, k1, N.o);
            ++N.call;
            N.avail = 2;
        }
        const bool second = N.avail == 1;
        --N.avail;
        const float v1 = sym24(second ? N.o[2] : N.o[0]);
        const float v2 = sym24(second ? N.o[3] : N.o[1]);
        const float s = fmaf(v1, v1, v2 * v2);
        if (s < 1.0f) {
            const float f = sqrtf(-2.0f * blog(s) / s);
            N.spare = v2 * f;
            N.has_spare = true;
            return v1 * f;
        }
    }
}

// four standard normals from one Philox block by the Box-Muller transform (no rejection, so no
// divergence): u1 in (0, 1], u2 in [0, 1) from the top 24 bits of a word; r = sqrt(-2 ln u1),
// (r cos 2 pi u2, r sin 2 pi u2), with this library's own log and sine/cosine (mci_f32math.hpp)
// and the correctly rounded square root, so that the CPU reproduces every normal bit for bit; each
// is within a few ulp of the exact transform of its two words.  The learned moves draw their
// normals with it.
__device__ __forceinline__ void normal4(uint32_t c0, uint32_t id, uint32_t tl, uint32_t th, uint32_t k0, uint32_t k1, float z[4])
{
    uint32_t o[4];
    philox(c0, id, tl, th, k0, k1, o);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const float u1 = (float)((o[2 * p] >> 8) + 1u) * 0x1p-24f;
        const float u2 = (float)(o[2 * p + 1] >> 8) * 0x1p-24f;
        const float r = sqrtf(-2.0f * blog(u1));
        float s, c;
        bsincos2pi(u2, s, c);
        z[2 * p] = r * c;
        z[2 * p + 1] = r * s;
    }
}

// log-density state (q_k, m, S) of a position held in registers: q_k = |U_k x - U_k mu_k|^2 with
// U_k the upper Cholesky factor of P_k (d (d + 1)/2 fused multiply-adds per component); U and
// U mu are read at the same address by every lane
template <class R>
__device__ void target_eval_reg(const Target<R> &T, const R *U, const R xr[DMAX], R q[KMAX], R &m, R &S)
{
    const int d = T.d;
    m = -__builtin_huge_val();
    R e[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const R *Uk = U + (int64_t)k * d * d;
        const R *bk = T.Um + k * d;
        R qk = 0;
#pragma unroll
        for (int i = 0; i < DMAX; ++i) {
            if (i < d) {
                R r = -bk[i];
#pragma unroll
                for (int j = i; j < DMAX; ++j) {
                    if (j < d) { r = fma(Uk[i * d + j], xr[j], r); }
                }
                qk = fma(r, r, qk);
            }
        }
        q[k] = qk;
        e[k] = T.c[k] - (R)0.5 * qk;
        m = e[k] > m ? e[k] : m;
    }
    S = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { S = S + rexp(e[k] - m); }
}

// the same for column w of x [d][ldx] in global memory
template <class R>
__device__ void target_eval(const Target<R> &T, const R *x, int64_t ldx, R q[KMAX], R &m, R &S)
{
    R xr[DMAX];
#pragma unroll
    for (int j = 0; j < DMAX; ++j) { xr[j] = (j < T.d) ? x[j * ldx] : (R)0; }
    target_eval_reg(T, T.U, xr, q, m, S);
}

template <class R>
__device__ __forceinline__ void accumulate(const State<R> &St, int64_t w, bool live, double sq, R x0, R &s0, R &c0, R &s1, R &c1, int &jb,
                                           int &bi)
{
    kahan(s0, c0, (R)sq);
    kahan(s1, c1, x0 > 0 ? (R)1 : (R)0);
    if (++jb == St.B) {
        if (live) {
            St.bmeans[((int64_t)bi * 2 + 0) * St.W + w] = ((double)s0 - (double)c0) / (double)St.B;
            St.bmeans[((int64_t)bi * 2 + 1) * St.W + w] = ((double)s1 - (double)c1) / (double)St.B;
        }
        s0 = c0 = s1 = c1 = 0;
        jb = 0;
        ++bi;
    }
}

// exact start: component by cumulative weight, z ~ N(0, I) by the polar method, x = mu + L z
template <class R>
__global__ __launch_bounds__(WG) void k_lp_init(Target<R> T, State<R> St, uint32_t k0, uint32_t k1)
{
    extern __shared__ float zs[]; // [d][WG]
    const int tid = threadIdx.x, d = T.d;
    const int64_t w = (int64_t)blockIdx.x * WG + tid;
    if (w >= St.W) { return; }
    uint32_t o[4];
    philox(0x40000000u, (uint32_t)w, 0u, 0u, k0, k1, o);
    const float uc = (float)(o[0] >> 8) * 0x1p-24f;
    int kc = 0;
    for (int k = 0; k < T.K - 1; ++k) { kc += uc >= T.cw[k] ? 1 : 0; }
    Normals N;
    N.avail = 0;
    N.has_spare = false;
    N.call = 1;
    for (int j = 0; j < d; ++j) { zs[j * WG + tid] = next_normal(N, 0x40000000u, (uint32_t)w, 0u, 0u, k0, k1); }
    const float *Lk = T.L + (int64_t)kc * d * d;
    double sq = 0.;
    for (int i = 0; i < d; ++i) {
        float xi = T.muf[kc * d + i];
        for (int j = 0; j <= i; ++j) { xi = fmaf(Lk[i * d + j], zs[j * WG + tid], xi); }
        St.x[(int64_t)i * St.W + w] = (R)xi;
        sq = fma((double)xi, (double)xi, sq);
    }
    R q[KMAX], m, S;
    target_eval(T, St.x + w, St.W, q, m, S);
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { St.q[(int64_t)k * St.W + w] = q[k]; }
    St.m[w] = m;
    St.S[w] = S;
    St.sq[w] = sq;
    for (int o2 = 0; o2 < 2; ++o2) {
        St.sb[o2 * St.W + w] = 0;
        St.cb[o2 * St.W + w] = 0;
    }
    St.jb[w] = 0;
    St.bi[w] = 0;
    St.nacc[w] = 0;
}

// the target's state at given positions (no draw): q_k, m, S and |x|^2 of every walker, empty
// block sums and counters
template <class R>
__global__ __launch_bounds__(WG) void k_lp_state(Target<R> T, State<R> St)
{
    const int64_t w = (int64_t)blockIdx.x * WG + threadIdx.x;
    if (w >= St.W) { return; }
    R q[KMAX], m, S;
    target_eval(T, St.x + w, St.W, q, m, S);
    double sq = 0.;
    for (int i = 0; i < T.d; ++i) { const double v = (double)St.x[(int64_t)i * St.W + w]; sq = fma(v, v, sq); }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { St.q[(int64_t)k * St.W + w] = q[k]; }
    St.m[w] = m;
    St.S[w] = S;
    St.sq[w] = sq;
    for (int o2 = 0; o2 < 2; ++o2) {
        St.sb[o2 * St.W + w] = 0;
        St.cb[o2 * St.W + w] = 0;
    }
    St.jb[w] = 0;
    St.bi[w] = 0;
    St.nacc[w] = 0;
}

template <class R>
__device__ __forceinline__ void rw_draw(uint32_t o1, uint32_t o2, uint32_t o3, R step, R &delta, R &ua);
template <>
__device__ __forceinline__ void rw_draw<double>(uint32_t o1, uint32_t o2, uint32_t o3, double step, double &delta, double &ua)
{
    delta = step * (((double)o1 + 0.5) * 0x1p-31 - 1.0);
    ua = ((double)(o2 >> 5) * 67108864.0 + (double)(o3 >> 6)) * 0x1p-53;
}
template <>
__device__ __forceinline__ void rw_draw<float>(uint32_t o1, uint32_t o2, uint32_t, float step, float &delta, float &ua)
{
    delta = step * sym24(o1);
    ua = (float)(o2 >> 8) * 0x1p-24f;
}

// per-lane chain state held in registers by the multi-step kernels
template <class R>
struct Lane
{
    R q[KMAX], c[KMAX];
    R m, S;
    double sq;
    R s0, c0, s1, c1;
    int jb, bi;
    int32_t nacc, nacc_ind;
};

template <class R>
__device__ void lane_load(const Target<R> &T, const State<R> &St, int64_t wl, Lane<R> &L)
{
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        L.q[k] = St.q[(int64_t)k * St.W + wl];
        L.c[k] = T.c[k];
    }
    L.m = St.m[wl];
    L.S = St.S[wl];
    L.sq = St.sq[wl];
    L.s0 = St.sb[wl];
    L.c0 = St.cb[wl];
    L.s1 = St.sb[St.W + wl];
    L.c1 = St.cb[St.W + wl];
    L.jb = St.jb[wl];
    L.bi = St.bi[wl];
    L.nacc = 0;
    L.nacc_ind = 0;
}

template <class R, int WGS>
__device__ void lane_store(const State<R> &St, int d, int64_t w, const R *xs, int tid, const Lane<R> &L)
{
    for (int i = 0; i < d; ++i) { St.x[(int64_t)i * St.W + w] = xs[i * WGS + tid]; }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { St.q[(int64_t)k * St.W + w] = L.q[k]; }
    St.m[w] = L.m;
    St.S[w] = L.S;
    St.sq[w] = L.sq;
    St.sb[w] = L.s0;
    St.cb[w] = L.c0;
    St.sb[St.W + w] = L.s1;
    St.cb[St.W + w] = L.c1;
    St.jb[w] = L.jb;
    St.bi[w] = L.bi;
    St.nacc[w] += L.nacc;
}

// LDS images of the target used by the multi-step kernels: P_k rows padded to d + 1 and P_k mu_k
template <class R, int WGS>
__device__ void load_target_lds(const Target<R> &T, R *Ps, R *vs)
{
    const int d = T.d, dp = d + 1;
    for (int e = threadIdx.x; e < KMAX * d * d; e += WGS) {
        const int k = e / (d * d), r = (e / d) % d, c = e % d;
        Ps[(k * d + r) * dp + c] = T.P[e];
    }
    for (int e = threadIdx.x; e < KMAX * d; e += WGS) { vs[e] = T.v[e]; }
}

// one random-walk step at time t (single-coordinate uniform move)
template <class R, int WGS>
__device__ __forceinline__ bool rw_step(const Target<R> &T, const R *Ps, const R *vs, R *xs, int tid, uint32_t wid, Lane<R> &L, R step,
                                        uint64_t t, uint32_t k0, uint32_t k1)
{
    const int d = T.d, dp = d + 1;
    const uint32_t thr = (uint32_t)(0u - (uint32_t)d) % (uint32_t)d;
    const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32);
    uint32_t o[4];
    philox(0u, wid, tl, th, k0, k1, o);
    uint64_t prod = (uint64_t)o[0] * (uint32_t)d;
    if (__builtin_expect(__any((uint32_t)prod < thr), 0)) {
        uint32_t low = (uint32_t)prod, kx = 0;
        while (low < thr) {
            const uint32_t kk = 0x80000000u + (kx++);
            uint32_t e[4];
            philox(kk >> 2, wid, tl, th, k0, k1, e);
            prod = (uint64_t)e[kk & 3u] * (uint32_t)d;
            low = (uint32_t)prod;
        }
    }
    const int j = (int)(prod >> 32);
    R delta, ua;
    rw_draw<R>(o[1], o[2], o[3], step, delta, ua);
    const R xo = xs[j * WGS + tid];
    const R xn = xo + delta;
    R g[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { g[k] = -vs[k * d + j]; }
    for (int i = 0; i < d; ++i) {
        const R xi = xs[i * WGS + tid];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) { g[k] = fma(Ps[(k * d + j) * dp + i], xi, g[k]); }
    }
    R qn[KMAX], en[KMAX];
    R mn = -__builtin_huge_val();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const R pjj = Ps[(k * d + j) * dp + j];
        qn[k] = L.q[k] + fma(delta + delta, g[k], delta * delta * pjj);
        en[k] = L.c[k] - (R)0.5 * qn[k];
        mn = en[k] > mn ? en[k] : mn;
    }
    R Sn = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { Sn = Sn + rexp(en[k] - mn); }
    const R E = rexp(mn - L.m);
    const bool acc = ua * L.S < E * Sn;
    if (acc) { xs[j * WGS + tid] = xn; }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { L.q[k] = acc ? qn[k] : L.q[k]; }
    L.m = acc ? mn : L.m;
    L.S = acc ? Sn : L.S;
    L.sq = acc ? L.sq + fma((double)xn, (double)xn, -((double)xo * (double)xo)) : L.sq;
    L.nacc += acc ? 1 : 0;
    return acc;
}

// observables after a step; at a block end the block means are written and q_k, m, S, |x|^2
// recomputed from x (bounding the drift of the incremental updates)
template <class R, int WGS>
__device__ __forceinline__ void after_step(const Target<R> &T, const State<R> &St, const R *Ps, const R *xs, int tid, int64_t w, bool live,
                                          Lane<R> &L)
{
    if (St.B <= 0) { return; }
    const int d = T.d, dp = d + 1;
    const int bprev = L.bi;
    accumulate(St, w, live, L.sq, xs[tid], L.s0, L.c0, L.s1, L.c1, L.jb, L.bi);
    if (L.bi == bprev) { return; }
    double q2 = 0.;
    for (int i = 0; i < d; ++i) {
        const double v = (double)xs[i * WGS + tid];
        q2 = fma(v, v, q2);
    }
    L.sq = q2;
    L.m = -__builtin_huge_val();
    R e2[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        const R *mu = T.mu + k * d;
        R qk = 0;
        for (int i = 0; i < d; ++i) {
            R row = 0;
            for (int jj = 0; jj < d; ++jj) { row = fma(Ps[(k * d + i) * dp + jj], xs[jj * WGS + tid] - mu[jj], row); }
            qk = fma(xs[i * WGS + tid] - mu[i], row, qk);
        }
        L.q[k] = qk;
        e2[k] = L.c[k] - (R)0.5 * qk;
        L.m = e2[k] > L.m ? e2[k] : L.m;
    }
    L.S = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { L.S = L.S + rexp(e2[k] - L.m); }
}

// nsteps random-walk steps t0 .. t0 + nsteps - 1 for 256 walkers per workgroup
template <class R>
__global__ __launch_bounds__(WG) void k_lp_rw(Target<R> T, State<R> St, R step, uint64_t t0, int64_t nsteps, uint32_t k0, uint32_t k1)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int d = T.d, tid = threadIdx.x;
    R *Ps = reinterpret_cast<R *>(lds_raw); // [KMAX][d][d + 1]
    R *vs = Ps + KMAX * d * (d + 1);         // [KMAX][d]
    R *xs = vs + KMAX * d;                   // [d][WG]
    const int64_t w = (int64_t)blockIdx.x * WG + tid;
    const bool live = w < St.W;
    const int64_t wl = live ? w : St.W - 1;
    load_target_lds<R, WG>(T, Ps, vs);
    for (int i = 0; i < d; ++i) { xs[i * WG + tid] = St.x[(int64_t)i * St.W + wl]; }
    __syncthreads();
    Lane<R> L;
    lane_load(T, St, wl, L);
    for (int64_t s = 0; s < nsteps; ++s) {
        (void)rw_step<R, WG>(T, Ps, vs, xs, tid, (uint32_t)wl, L, step, t0 + (uint64_t)s, k0, k1);
        after_step<R, WG>(T, St, Ps, xs, tid, w, live, L);
    }
    if (live) { lane_store<R, WG>(St, d, w, xs, tid, L); }
}

// ---- the learned move, fused: one launch per run of consecutive learned steps ----
// A workgroup takes TW walkers (32: four waves; 64: eight waves) for the whole run.  The network is
// evaluated at the walkers' positions when the run starts (random-walk moves in between need no
// refresh) and at every proposal; the mixture at the current position is carried in registers
// from step to step.  Network: v_mfma_f32_32x32x16_bf16 tiles, weights read from L2 as [out][in]
// rows, biases as the accumulators' start values, activations walker-major in LDS, one activation
// buffer per workgroup (a layer's output stays in accumulators until every wave has read its
// input).  The last layer stays in registers: wave w holds components 2 (w % 4), 2 (w % 4) + 1 for
// the walker half w / 4, its rows ordered as (mean_j, log-scale_j) pairs so that each lane holds
// whole pairs of one walker, each pair kept as (mean, 1 / sigma) in two halves; wave 0 (and
// wave 4) also computes the K logits.  A step:
//  1. the component from the logits at x, x' = mu + sigma z from the lanes holding that component,
//     log q(x' | x) folded from the registers; log p(x') (one target component per wave);
//  2. the network at x', log q(x | x') folded the same way;
//  3. the MH test; accepted walkers take x', its target state, logits and mixture.
// Mode 0 evaluates the network only and writes the mixture (means, log-scales, logits) for tests.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __fp16 h2_t __attribute__((ext_vector_type(2)));
#ifndef MCIL_NS
#define MCIL_NS 2 // last-layer row blocks per pass
#endif
#ifndef MCIL_PD
#define MCIL_PD 3 // last-layer weight fragments in flight (reduction steps ahead, from L2)
#endif
#ifndef MCIL_TW
#define MCIL_TW 64
#endif
constexpr int TW = MCIL_TW;    // walkers per workgroup (32 or 64)
constexpr int NCB = TW / 32;   // walker blocks of 32 (matrix-core columns)
constexpr int NWV = 4 * NCB;   // waves per workgroup: 4 component pairs x walker blocks
constexpr int TT = 64 * NWV;   // threads per workgroup
constexpr int NPART = TT / TW; // threads per walker in the per-walker loops (8)
static_assert(NPART == KMAX, "one thread per walker and target component");

template <int D, int H>
struct LCfg
{
    static constexpr int DP = D < 16 ? 16 : D;  // padded input width
    static constexpr int LDX = DP + 8;          // LDS row pitches (BF16 elements)
    static constexpr int LDH = H + 8;
    static constexpr int NRB = D / 8;           // 32-row blocks of a wave's two components
    static constexpr int LOFF = 2 * D * KMAX;   // first logit row
    static constexpr int NB = 2 * H + LOFF + 32; // biases
};

__device__ __forceinline__ float silu(float v) { return v * __builtin_amdgcn_rcpf(1.0f + __expf(-v)); }

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// acc[r] = bias + rows row0 + 32 r .. + 31 of Wt [.][ldw] (BF16, row = output unit, KD used) times walkers
// 32 cb .. + 31 of act [64][lda] (BF16, LDS), r < NR; bias in LDS, indexed by row
template <int KD, int NR, int PD = 1>
__device__ __forceinline__ void mfma_tile(const uint16_t *__restrict__ Wt, int ldw, const float *bias, int row0, const uint16_t *act, int lda,
                                          int cb, f32x16 acc[NR])
{
    const int lane = threadIdx.x & 63;
    constexpr int NQ = KD / 16, NB = PD + 1;
    const uint16_t *wr = Wt + (int64_t)(row0 + (lane & 31)) * ldw + 8 * (lane >> 5);
    const uint16_t *br = act + (32 * cb + (lane & 31)) * lda + 8 * (lane >> 5);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float4 b = *reinterpret_cast<const float4 *>(bias + row0 + 32 * r + 8 * g + 4 * (lane >> 5));
            acc[r][4 * g + 0] = b.x;
            acc[r][4 * g + 1] = b.y;
            acc[r][4 * g + 2] = b.z;
            acc[r][4 * g + 3] = b.w;
        }
    }
    // PD reduction steps of fragments in flight ahead of the products (a ring of PD + 1 sets;
    // the depth bounds the live registers)
    bf16x8 a[NB][NR], b[NB];
#pragma unroll
    for (int q = 0; q < PD && q < NQ; ++q) {
#pragma unroll
        for (int r = 0; r < NR; ++r) { a[q][r] = *reinterpret_cast<const bf16x8 *>(wr + (int64_t)32 * r * ldw + 16 * q); }
        b[q] = *reinterpret_cast<const bf16x8 *>(br + 16 * q);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int cu = q % NB, nx = (q + PD) % NB;
        if (q + PD < NQ) {
#pragma unroll
            for (int r = 0; r < NR; ++r) { a[nx][r] = *reinterpret_cast<const bf16x8 *>(wr + (int64_t)32 * r * ldw + 16 * (q + PD)); }
            b[nx] = *reinterpret_cast<const bf16x8 *>(br + 16 * (q + PD));
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) { acc[r] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[cu][r], b[cu], acc[r], 0, 0, 0); }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// hidden layer: Hb [64][LDH] = bf16(silu(W act + b)); the 2 H / 32 blocks of 32 rows x 32 walkers
// go to waves 0 .. 2 H / 32 - 1 (row block w % (H / 32), walker half w / (H / 32)).  The caller
// has synchronised so that act is complete; when act is Hb itself, the waves synchronise between
// the products and the write-back.
template <int KD, int H>
__device__ __forceinline__ void hidden_layer(const uint16_t *Wt, int ldw, const float *bias, const uint16_t *act, int lda, uint16_t *Hb,
                                             bool in_place)
{
    constexpr int NRH = H / 32, LDH = H + 8;
    const int w = wave_id(), lane = threadIdx.x & 63;
    const bool mine = w < NRH * NCB;
    const int rb = w % NRH, cb = w / NRH;
    f32x16 acc[1];
    if (mine) { mfma_tile<KD, 1>(Wt, ldw, bias, 32 * rb, act, lda, cb, acc); }
    if (in_place) { __syncthreads(); }
    if (mine) {
        const int col = 32 * cb + (lane & 31);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int r = 32 * rb + 8 * g + 4 * (lane >> 5);
            uint32_t p[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const __bf16 lo = (__bf16)silu(acc[0][4 * g + 2 * h]);
                const __bf16 hi = (__bf16)silu(acc[0][4 * g + 2 * h + 1]);
                p[h] = (uint32_t)__builtin_bit_cast(uint16_t, lo) | ((uint32_t)__builtin_bit_cast(uint16_t, hi) << 16);
            }
            *reinterpret_cast<uint2 *>(Hb + col * LDH + r) = make_uint2(p[0], p[1]);
        }
    }
}

// the network at the positions in Xb.  Wave w keeps the last layer of components 2 (w % 4),
// 2 (w % 4) + 1 for walkers 32 (w / 4) + lane % 32 in registers: pair (rb 4 + g) 2 + h of mx holds
// (mean, 1 / sigma) of one coordinate as two halves (rounded toward zero; the move samples and
// evaluates with exactly these values), lpi[cl] the sum of log(1 / sigma) over the lane's pairs of
// component 2 (w % 4) + cl.  The logits go to lg [KMAX][64]; mode 0 also writes the mixture to out
// [K (1 + 2 D)][n] (walker w0 + column): means, log-scales -log(1 / sigma), logits.
__device__ __forceinline__ constexpr int pair_index(int rb, int g, int h) { return (rb * 4 + g) * 2 + h; }

template <int D, int H>
__device__ __forceinline__ void network(const NetDev &N, const uint16_t *W1s, const uint16_t *W2s, const float *bs, const uint16_t *Xb,
                                        uint16_t *Hb, h2_t mx[D], float lpi[2], float *lg, float *out, int64_t n, int64_t w0)
{
    using C = LCfg<D, H>;
    const int w = wave_id(), lane = threadIdx.x & 63, K = N.K;
    const int wc = w & 3, cb = w >> 2;
    const float *b3 = bs + 2 * H;
    hidden_layer<C::DP, H>(W1s, C::LDX, bs, Xb, C::LDX, Hb, false);
    __syncthreads();
    hidden_layer<H, H>(W2s, C::LDH, bs + H, Hb, C::LDH, Hb, true);
    __syncthreads();
    const int col = 32 * cb + (lane & 31);
    if (wc == 0) {
        f32x16 al[1];
        mfma_tile<H, 1, MCIL_PD>(N.W3t, H, b3, C::LOFF, Hb, C::LDH, cb, al);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int k = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            if (k < K) {
                lg[k * TW + col] = al[0][i];
                if (out != nullptr && w0 + col < n) { out[(int64_t)(2 * D * K + k) * n + w0 + col] = al[0][i]; }
            }
        }
    }
    if (2 * wc < K) {
        constexpr int NS = C::NRB >= MCIL_NS ? MCIL_NS : 1; // row blocks per pass (bounds the live accumulators)
        float pr[2] = {1.f, 1.f};
#pragma unroll
        for (int rb0 = 0; rb0 < C::NRB; rb0 += NS) {
            f32x16 acc[NS];
            mfma_tile<H, NS, MCIL_PD>(N.W3t, H, b3, 4 * D * wc + 32 * rb0, Hb, C::LDH, cb, acc);
#pragma unroll
            for (int rr = 0; rr < NS; ++rr) {
                const int rb = rb0 + rr;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int cl = (32 * rb + 8 * g) / (2 * D);
                    const int r = 32 * rb + 8 * g + 4 * (lane >> 5); // wave-local row of the first pair
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const float mu = __builtin_amdgcn_fmed3f(mcim::MU_SCALE * acc[rr][4 * g + 2 * h], -6.0e4f, 6.0e4f);
                        const float ls = __builtin_amdgcn_fmed3f(acc[rr][4 * g + 2 * h + 1], -mcim::LS_MAX, mcim::LS_MAX);
                        const h2_t p = __builtin_amdgcn_cvt_pkrtz(mu, bexp(-ls));
                        mx[pair_index(rb, g, h)] = p;
                        pr[cl] *= (float)p[1];
                        const int k = 2 * wc + cl;
                        if (out != nullptr && k < K && w0 + col < n) {
                            const int j = ((r % (2 * D)) >> 1) + h;
                            out[(int64_t)(2 * D * k + j) * n + w0 + col] = (float)p[0];
                            out[(int64_t)(2 * D * k + D + j) * n + w0 + col] = -blog((float)p[1]);
                        }
                    }
                }
            }
        }
        lpi[0] = blog(pr[0]);
        lpi[1] = blog(pr[1]);
    }
}

// per component and walker: sum_j log(1 / sigma_j) - |(y - mu) / sigma|^2 / 2 over the wave's two
// components, y [D][64] in LDS, into cf [KMAX][64]
template <int D>
__device__ __forceinline__ void fold(const h2_t mx[D], const float lpi[2], const float *y, float *cf, int K)
{
    const int w = wave_id(), lane = threadIdx.x & 63;
    const int wc = w & 3, col = 32 * (w >> 2) + (lane & 31);
    if (2 * wc >= K) { return; }
    float s[2] = {0.f, 0.f};
#pragma unroll
    for (int rb = 0; rb < D / 8; ++rb) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int cl = (32 * rb + 8 * g) / (2 * D);
            const int j = ((32 * rb + 8 * g + 4 * (lane >> 5)) % (2 * D)) >> 1;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const h2_t p = mx[pair_index(rb, g, h)];
                const float z = (y[(j + h) * TW + col] - (float)p[0]) * (float)p[1];
                s[cl] = fmaf(z, z, s[cl]);
            }
        }
    }
#pragma unroll
    for (int cl = 0; cl < 2; ++cl) {
        float v = lpi[cl] - 0.5f * s[cl];
        v += __shfl_xor(v, 32);
        if (lane < 32 && 2 * wc + cl < K) { cf[(2 * wc + cl) * TW + col] = v; }
    }
}

// q_k = |U_k y - U_k mu_k|^2 with v_mfma_f32_32x32x2_f32 (FP32 products, FP32 sums) for two
// (component, walker block of 32) pairs: rows of U_k (zero-padded to 32) times y [D][TW] (LDS),
// the accumulators starting at -U_k mu_k; lanes 0 .. 31 return q of walker 32 cb + lane.  U comes
// transposed, Ut [k][j][32 rows], so that a reduction step's operand is two 128-byte lines; all
// of a pass's operands are loaded (global address space: the loads need not wait on LDS traffic)
// before the products.  SAMEK: k[0] == k[1] (one set of operands).
typedef float f32x16v __attribute__((ext_vector_type(16)));
typedef const float __attribute__((address_space(1))) *gfloat_p;
template <int D, bool SAMEK>
__device__ __forceinline__ void target_q(const float *Ut, const float *Um, const int k[2], const int cb[2], const float *y, int lane, float out[2])
{
    const int row = lane & 31, kk = lane >> 5;
    const gfloat_p Ug = (gfloat_p)Ut, Umg = (gfloat_p)Um;
    constexpr int NA = SAMEK ? 1 : 2;
    float a[NA][D / 2];
#pragma unroll
    for (int c = 0; c < NA; ++c) {
#pragma unroll
        for (int q = 0; q < D / 2; ++q) { a[c][q] = Ug[((int64_t)k[c] * D + 2 * q + kk) * 32 + row]; }
    }
    f32x16v acc[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = 8 * (i >> 2) + 4 * kk + (i & 3);
            acc[c][i] = r < D ? -Umg[k[c] * D + r] : 0.f;
        }
    }
#pragma unroll
    for (int q = 0; q < D / 2; ++q) {
        const int j = 2 * q + kk;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[SAMEK ? 0 : c][q], y[j * TW + 32 * cb[c] + row], acc[c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        float s = 0.f;
#pragma unroll
        for (int i = 0; i < 16; ++i) { s = fmaf(acc[c][i], acc[c][i], s); }
        s += __shfl_xor(s, 32);
        out[c] = s;
    }
}

// log q(y | the mixture whose logits are lg and component sums cf), for walker col
__device__ __forceinline__ float mixture_logq(const float *lg, const float *cf, int K, int D, int col)
{
    constexpr float HALF_LOG_2PI = 0.91893853320467274f;
    float amax = -__builtin_huge_valf(), cmax = -__builtin_huge_valf();
    float comp[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < K) {
            const float a = lg[k * TW + col];
            amax = fmaxf(amax, a);
            comp[k] = a + cf[k * TW + col];
            cmax = fmaxf(cmax, comp[k]);
        }
    }
    float sa = 0.f, sc = 0.f;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k < K) {
            sa += bexp(lg[k * TW + col] - amax);
            sc += bexp(comp[k] - cmax);
        }
    }
    return (cmax + blog(sc)) - (amax + blog(sa)) - (float)D * HALF_LOG_2PI;
}

template <int D, int H>
size_t learned_lds()
{
    using C = LCfg<D, H>;
    return sizeof(float) * ((size_t)3 * D * TW + (size_t)6 * KMAX * TW + C::NB) + sizeof(int) * 3 * TW +
           sizeof(uint16_t) * ((size_t)TW * C::LDX + (size_t)TW * C::LDH + (size_t)H * C::LDX + (size_t)H * C::LDH);
}

template <int MODE, int D, int H> // 0: the network at xin, written to out; 1: nl learned steps from step t0
__global__ __launch_bounds__(TT) void k_lp_learned(NetDev N, Target<float> T, State<float> St, const float *xin, float *out, uint64_t t0, int nl,
                                                   uint32_t k0, uint32_t k1, int32_t *nacc_ind)
{
    using C = LCfg<D, H>;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int K = N.K;
    const int tid = threadIdx.x, col = tid % TW, part = tid / TW, lane = tid & 63;
    const int wv = wave_id(), wc = wv & 3, cbw = wv >> 2;
    const int64_t W = St.W;
    const int64_t w0 = (int64_t)blockIdx.x * TW;
    const int64_t w = w0 + col;
    const bool live = w < W;
    const int64_t wl = live ? w : W - 1;
    float *xs = reinterpret_cast<float *>(lds_raw);      // [D][64] current x
    float *xps = xs + D * TW;                             // [D][64] the proposal
    float *lgx = xps + D * TW;                            // [KMAX][64] logits at x
    float *lg = lgx + KMAX * TW;                          // [KMAX][64] logits at x'
    float *cf = lg + KMAX * TW;                           // [KMAX][64] component sums of the fold at x'
    float *cf2 = cf + KMAX * TW;                          // [KMAX][64] the same of the fold at x
    float *zs = cf2 + KMAX * TW;                          // [D][64] normals of the next proposal
    float *qn = zs + D * TW;                              // [KMAX][64] target quadratic forms at x'
    float *qx = qn + KMAX * TW;                           // [KMAX][64] the same at x (accepted moves)
    float *bs = qx + KMAX * TW;                           // biases: b1 [H], b2 [H], b3 [LOFF + 32]
    int *kst = reinterpret_cast<int *>(bs + C::NB);       // [64] chosen component
    int *af = kst + TW;                                   // [64] accepted in this step
    int *anyacc = af + TW;                                // [64] accepted in this run
    uint16_t *Xb = reinterpret_cast<uint16_t *>(anyacc + TW); // [64][LDX]
    uint16_t *Hb = Xb + TW * C::LDX;                      // [64][LDH]
    uint16_t *W1s = Hb + TW * C::LDH;                     // [H][LDX] first-layer weights, read every step
    uint16_t *W2s = W1s + H * C::LDX;                     // [H][LDH] second-layer weights
    const float *xsrc = MODE == 0 ? xin : St.x;

    for (int e = tid; e < H * (C::DP / 8); e += TT) {
        const int h = e / (C::DP / 8), c = 8 * (e % (C::DP / 8));
        *reinterpret_cast<uint4 *>(W1s + h * C::LDX + c) = *reinterpret_cast<const uint4 *>(N.W1t + (int64_t)h * C::DP + c);
    }
    for (int e = tid; e < H * (H / 8); e += TT) {
        const int h = e / (H / 8), c = 8 * (e % (H / 8));
        *reinterpret_cast<uint4 *>(W2s + h * C::LDH + c) = *reinterpret_cast<const uint4 *>(N.W2t + (int64_t)h * H + c);
    }

    for (int i = tid; i < H; i += TT) {
        bs[i] = N.b1[i];
        bs[H + i] = N.b2[i];
    }
    for (int i = tid; i < C::LOFF + 32; i += TT) { bs[2 * H + i] = N.b3[i]; }
    for (int j = part; j < D; j += NPART) { xs[j * TW + col] = xsrc[(int64_t)j * W + wl]; }
    for (int j = part; j < C::DP; j += NPART) {
        const float v = j < D ? xsrc[(int64_t)j * W + wl] : 0.f;
        Xb[col * C::LDX + j] = f2bf(v * mcim::IN_SCALE);
    }
    __syncthreads();
    h2_t mx[D];
    float lpi[2] = {0.f, 0.f};
    network<D, H>(N, W1s, W2s, bs, Xb, Hb, mx, lpi, lgx, MODE == 0 ? out : nullptr, W, w0);
    if (MODE == 0) { return; }
    // the logits of walkers 32 .. 63 come from wave 4
    __syncthreads();
    // the chain state of walker col, held by wave 0
    float m = 0.f, S = 1.f, s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f;
    double sq = 0.;
    int jb = 0, bi = 0, nacc = 0;
    if (part == 0) {
        m = St.m[wl];
        S = St.S[wl];
        sq = St.sq[wl];
        if (St.B > 0) {
            s0 = St.sb[wl];
            c0 = St.cb[wl];
            s1 = St.sb[W + wl];
            c1 = St.cb[W + wl];
            jb = St.jb[wl];
            bi = St.bi[wl];
        }
    }
    // normals of step t go to zs, drawn by waves 1 .. 7 (generator g = wave - 1 takes coordinates
    // g, g + 7, ...); they are drawn for the next step while wave 0 runs the MH test
    auto draw_normals = [&](uint64_t t) {
        if (part > 0) {
            constexpr int NZ = (D + NPART - 2) / (NPART - 1); // normals per generator at most
            float zq[4 * ((NZ + 3) / 4)];
            const uint32_t base = 0x10000000u + ((uint32_t)part << 20);
#pragma unroll
            for (int c = 0; c < (NZ + 3) / 4; ++c) { normal4(base + 1u + (uint32_t)c, (uint32_t)wl, (uint32_t)t, (uint32_t)(t >> 32), k0, k1, zq + 4 * c); }
#pragma unroll
            for (int i = 0; i < NZ; ++i) {
                const int j = part - 1 + (NPART - 1) * i;
                if (j < D) { zs[j * TW + col] = zq[i]; }
            }
        }
    };
    // the component of step t from the logits at x (wave 0; o holds the step's words: [0] the
    // component, [1] the acceptance uniform)
    uint32_t o[4] = {0u, 0u, 0u, 0u};
    auto choose = [&](uint64_t t) {
        philox(0x10000000u, (uint32_t)wl, (uint32_t)t, (uint32_t)(t >> 32), k0, k1, o);
        float amax = -__builtin_huge_valf();
        for (int k = 0; k < K; ++k) { amax = fmaxf(amax, lgx[k * TW + col]); }
        float e[KMAX], sa = 0.f;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            e[k] = (k < K) ? bexp(lgx[k * TW + col] - amax) : 0.f;
            sa += e[k];
        }
        const float uc = (float)(o[0] >> 8) * 0x1p-24f * sa;
        int kk = K - 1;
        float cum = 0.f;
#pragma unroll
        for (int k = 0; k < KMAX - 1; ++k) {
            cum += e[k];
            if (k < K - 1 && kk == K - 1 && uc < cum) { kk = k; }
        }
        kst[col] = kk;
    };
    if (part == 0) { choose(t0); }
    else { draw_normals(t0); }
    __syncthreads();
    for (int st = 0; st < nl; ++st) {
        // the weights and target factors are read afresh each step: an empty asm makes their
        // addresses opaque, so the compiler does not hoist those loads out of the loop (they would
        // not fit the registers)
        NetDev Ns = N;
        const float *Ubase = T.Ut, *Umbase = T.Um;
        asm volatile("" : "+s"(Ns.W3t), "+s"(Ubase), "+s"(Umbase));
        const uint64_t t = t0 + (uint64_t)st;
        // x' = mu + sigma z from the lanes holding the chosen component
        if (2 * wc < K) {
            const int cc = 32 * cbw + (lane & 31);
            const int kc = kst[cc] - 2 * wc;
            // the lane's normals first (one wait), then x'
            float zv[D];
#pragma unroll
            for (int rb = 0; rb < D / 8; ++rb) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int j = ((32 * rb + 8 * g + 4 * (lane >> 5)) % (2 * D)) >> 1;
#pragma unroll
                    for (int h = 0; h < 2; ++h) { zv[pair_index(rb, g, h)] = zs[(j + h) * TW + cc]; }
                }
            }
#pragma unroll
            for (int rb = 0; rb < D / 8; ++rb) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int cl = (32 * rb + 8 * g) / (2 * D);
                    const int j = ((32 * rb + 8 * g + 4 * (lane >> 5)) % (2 * D)) >> 1;
                    if (kc == cl) {
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const h2_t p = mx[pair_index(rb, g, h)];
                            xps[(j + h) * TW + cc] = fmaf(zv[pair_index(rb, g, h)], __builtin_amdgcn_rcpf((float)p[1]), (float)p[0]);
                        }
                    }
                }
            }
        }
        __syncthreads();
        fold<D>(mx, lpi, xps, cf, K);
        for (int j = part; j < C::DP; j += NPART) { Xb[col * C::LDX + j] = f2bf(j < D ? xps[j * TW + col] * mcim::IN_SCALE : 0.f); }
        {
            // target component wv at x' on the matrix cores (FP32 products and sums):
            // U_k x' - U_k mu_k for both walker halves, then |.|^2 over the 32 rows
            static_assert(D == 32 || D == 16 || D == 8, "dimension");
            // wave w: components w for both walker blocks (64 walkers) or 2 w, 2 w + 1 (32)
            const int kq[2] = {NCB == 2 ? wv : 2 * wv, NCB == 2 ? wv : 2 * wv + 1};
            const int cq[2] = {0, NCB == 2 ? 1 : 0};
            float qv[2];
            target_q<D, NCB == 2>(Ubase, Umbase, kq, cq, xps, lane, qv);
            if (lane < 32) {
                qn[kq[0] * TW + 32 * cq[0] + lane] = qv[0];
                qn[kq[1] * TW + 32 * cq[1] + lane] = qv[1];
            }
        }
        __syncthreads();
        // the network at x' and log q(x | x')
        h2_t mx2[D];
        float lpi2[2] = {0.f, 0.f};
        network<D, H>(Ns, W1s, W2s, bs, Xb, Hb, mx2, lpi2, lg, nullptr, W, w0);
        fold<D>(mx2, lpi2, xs, cf2, K);
        __syncthreads();
        // wave 0: the MH test, the accepted walkers' new state, then the next step's component;
        // waves 1 .. 7 meanwhile draw the next step's normals
        if (part == 0) {
            const float logq_f = mixture_logq(lgx, cf, K, D, col);
            const float logq_r = mixture_logq(lg, cf2, K, D, col);
            float mn = -__builtin_huge_valf();
            float e[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                e[k] = T.c[k] - 0.5f * qn[k * TW + col];
                mn = fmaxf(mn, e[k]);
            }
            float Sn = 0.f;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) { Sn += bexp(e[k] - mn); }
            const float ua = ((float)(o[1] >> 8) + 0.5f) * 0x1p-24f;
            const bool acc = live && (blog(ua) < (mn + blog(Sn)) - (m + blog(S)) + (logq_r - logq_f));
            af[col] = acc ? 1 : 0;
            const float x0 = acc ? xps[col] : xs[col];
            if (acc) {
                double s2 = 0.;
#pragma unroll
                for (int j0 = 0; j0 < D; j0 += 8) {
                    float v[8]; // eight reads in flight, then the writes
#pragma unroll
                    for (int j = 0; j < 8; ++j) { v[j] = xps[(j0 + j) * TW + col]; }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        xs[(j0 + j) * TW + col] = v[j];
                        s2 = fma((double)v[j], (double)v[j], s2);
                    }
                }
#pragma unroll
                for (int k = 0; k < KMAX; ++k) { qx[k * TW + col] = qn[k * TW + col]; }
                for (int k = 0; k < K; ++k) { lgx[k * TW + col] = lg[k * TW + col]; }
                m = mn;
                S = Sn;
                sq = s2;
                ++nacc;
            }
            if (live && St.B > 0) { accumulate(St, w, true, sq, x0, s0, c0, s1, c1, jb, bi); }
            if (st + 1 < nl) { choose(t + 1); }
        }
        else if (st + 1 < nl) { draw_normals(t + 1); }
        __syncthreads();
        // accepted walkers: the mixture at x takes the proposal's
        {
            const bool a = af[32 * cbw + (lane & 31)] != 0;
#pragma unroll
            for (int i = 0; i < D; ++i) { mx[i] = a ? mx2[i] : mx[i]; }
            lpi[0] = a ? lpi2[0] : lpi[0];
            lpi[1] = a ? lpi2[1] : lpi[1];
        }
    }
    // the run's end: accepted walkers' position and quadratic forms, every walker's chain state
    if (part == 0) { anyacc[col] = nacc; }
    __syncthreads();
    if (live && anyacc[col] > 0) {
        for (int j = part; j < D; j += NPART) { St.x[(int64_t)j * W + w] = xs[j * TW + col]; }
        St.q[(int64_t)part * W + w] = qx[part * TW + col];
    }
    if (part == 0 && live) {
        St.m[w] = m;
        St.S[w] = S;
        St.sq[w] = sq;
        if (St.B > 0) {
            St.sb[w] = s0;
            St.cb[w] = c0;
            St.sb[W + w] = s1;
            St.cb[W + w] = c1;
            St.jb[w] = jb;
            St.bi[w] = bi;
        }
        nacc_ind[w] += nacc;
    }
}

// ---- the independence proposal: q(x') = sum_k pi_k N(x'; mu_k, diag sigma_k^2), one mixture for
// every walker ----
// A workgroup takes TW = 64 walkers for a run of nl consecutive learned steps with eight waves.  The
// mixture lives in LDS, read by a whole wave at one address (wave k folds component k) or with a
// padded pitch (the x' build, where lanes hold different components), so that no lane keeps it in
// registers and two workgroups share a compute unit (the phases of one overlap the other's).  A step:
//  1. every thread builds D / 8 coordinates of x' = mu_k + sigma_k z (k and z were drawn at the end
//     of the previous step);
//  2. wave k: the fold of x' under component k, lc_k - |(x' - mu_k) / sigma_k|^2 / 2, and the
//     target's q_k at x' on the matrix cores;
//  3. wave 0: log q(x') by log-sum-exp over the folds, the MH test against the log q(x) kept from the
//     step that last moved the walker, the accepted walkers' new state, the next component; waves
//     1 .. 7 meanwhile draw the next step's normals.
// The random words are those of the learned move (component and acceptance uniform from counter
// 0x10000000, normals from 0x10000000 + (generator << 20) + call).
template <int D>
__global__ __launch_bounds__(TT) __attribute__((amdgpu_waves_per_eu(4))) void k_lp_indep(IndepMix M, Target<float> T, State<float> St, uint64_t t0,
                                                                                        int nl, uint32_t k0, uint32_t k1, int32_t *nacc_ind)
{
    static_assert(TW == 64 && NPART == KMAX, "64 walkers, one wave per component");
    constexpr int DP = D + 1; // mixture pitch
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int K = M.K;
    const int tid = threadIdx.x, col = tid % TW, part = tid / TW, lane = tid & 63;
    const int wv = wave_id();
    const int64_t W = St.W;
    const int64_t w = (int64_t)blockIdx.x * TW + col;
    const bool live = w < W;
    const int64_t wl = live ? w : W - 1;
    float *xs = reinterpret_cast<float *>(lds_raw); // [D][64] current x
    float *xps = xs + D * TW;                        // [D][64] the proposal
    float *zs = xps + D * TW;                        // [D][64] normals of the next proposal
    float *cf = zs + D * TW;                         // [KMAX][64] folds (at x', at the start at x)
    float *qn = cf + KMAX * TW;                      // [KMAX][64] target quadratic forms at x'
    float *qx = qn + KMAX * TW;                      // [KMAX][64] the same at x (accepted moves)
    float *mus = qx + KMAX * TW;                     // [KMAX][DP]
    float *isg = mus + KMAX * DP;                    // [KMAX][DP] 1 / sigma
    float *sgs = isg + KMAX * DP;                    // [KMAX][DP] sigma
    float *lcs = sgs + KMAX * DP;                    // [KMAX]
    float *cws = lcs + KMAX;                         // [KMAX]
    int *kst = reinterpret_cast<int *>(cws + KMAX);  // [64] chosen component
    int *af = kst + TW;                              // [64] accepted in this step
    int *anyacc = af + TW;                           // [64] accepted in this run
    for (int e = tid; e < KMAX * D; e += TT) {
        const int k = e / D, j = e % D;
        mus[k * DP + j] = M.mu[e];
        isg[k * DP + j] = M.isig[e];
        sgs[k * DP + j] = M.sig[e];
    }
    if (tid < KMAX) {
        lcs[tid] = M.lc[tid];
        cws[tid] = M.cw[tid];
    }
    for (int j = part; j < D; j += NPART) { xs[j * TW + col] = St.x[(int64_t)j * W + wl]; }
    __syncthreads();
    // wave k: the fold of y [D][64] under component k into cf
    auto fold_k = [&](const float *y) {
        if (wv < K) {
            float s = 0.f;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const float z = (y[j * TW + lane] - mus[wv * DP + j]) * isg[wv * DP + j];
                s = fmaf(z, z, s);
            }
            cf[wv * TW + lane] = lcs[wv] - 0.5f * s;
        }
    };
    // log q at walker col from the folds in cf (wave 0)
    auto logq = [&]() {
        float v[KMAX], mx = -__builtin_huge_valf();
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            v[k] = k < K ? cf[k * TW + col] : -__builtin_huge_valf();
            mx = fmaxf(mx, v[k]);
        }
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) { s += k < K ? bexp(v[k] - mx) : 0.f; }
        return mx + blog(s);
    };
    fold_k(xs);
    __syncthreads();
    float m = 0.f, S = 1.f, s0 = 0.f, c0 = 0.f, s1 = 0.f, c1 = 0.f, lqx = 0.f;
    double sq = 0.;
    int jb = 0, bi = 0, nacc = 0;
    uint32_t o[4] = {0u, 0u, 0u, 0u};
    auto choose = [&](uint64_t t) {
        philox(0x10000000u, (uint32_t)wl, (uint32_t)t, (uint32_t)(t >> 32), k0, k1, o);
        const float u = (float)(o[0] >> 8) * 0x1p-24f;
        int kk = K - 1;
#pragma unroll
        for (int k = KMAX - 2; k >= 0; --k) { kk = (k < K - 1 && u < cws[k]) ? k : kk; }
        kst[col] = kk;
    };
    auto draw_normals = [&](uint64_t t) {
        constexpr int NZ = (D + NPART - 2) / (NPART - 1);
        float zq[4 * ((NZ + 3) / 4)];
        const uint32_t base = 0x10000000u + ((uint32_t)part << 20);
#pragma unroll
        for (int c = 0; c < (NZ + 3) / 4; ++c) { normal4(base + 1u + (uint32_t)c, (uint32_t)wl, (uint32_t)t, (uint32_t)(t >> 32), k0, k1, zq + 4 * c); }
#pragma unroll
        for (int i = 0; i < NZ; ++i) {
            const int j = part - 1 + (NPART - 1) * i;
            if (j < D) { zs[j * TW + col] = zq[i]; }
        }
    };
    if (part == 0) {
        lqx = logq();
        m = St.m[wl];
        S = St.S[wl];
        sq = St.sq[wl];
        if (St.B > 0) {
            s0 = St.sb[wl];
            c0 = St.cb[wl];
            s1 = St.sb[W + wl];
            c1 = St.cb[W + wl];
            jb = St.jb[wl];
            bi = St.bi[wl];
        }
        choose(t0);
    }
    else { draw_normals(t0); }
    __syncthreads();
    for (int st = 0; st < nl; ++st) {
        const float *Ubase = T.Ut, *Umbase = T.Um;
        asm volatile("" : "+s"(Ubase), "+s"(Umbase));
        const uint64_t t = t0 + (uint64_t)st;
        {
            const int k = kst[col];
            float zv[(D + NPART - 1) / NPART];
#pragma unroll
            for (int i = 0; i < (D + NPART - 1) / NPART; ++i) {
                const int j = part + NPART * i;
                zv[i] = j < D ? zs[j * TW + col] : 0.f;
            }
#pragma unroll
            for (int i = 0; i < (D + NPART - 1) / NPART; ++i) {
                const int j = part + NPART * i;
                if (j < D) { xps[j * TW + col] = fmaf(sgs[k * DP + j], zv[i], mus[k * DP + j]); }
            }
        }
        __syncthreads();
        fold_k(xps);
        {
            const int kq[2] = {wv, wv};
            const int cq[2] = {0, 1};
            float qv[2];
            target_q<D, true>(Ubase, Umbase, kq, cq, xps, lane, qv);
            if (lane < 32) {
                qn[wv * TW + lane] = qv[0];
                qn[wv * TW + 32 + lane] = qv[1];
            }
        }
        __syncthreads();
        if (part == 0) {
            const float lqn = logq();
            float mn = -__builtin_huge_valf();
            float e[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                e[k] = T.c[k] - 0.5f * qn[k * TW + col];
                mn = fmaxf(mn, e[k]);
            }
            float Sn = 0.f;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) { Sn += bexp(e[k] - mn); }
            const float ua = ((float)(o[1] >> 8) + 0.5f) * 0x1p-24f;
            const bool acc = live && (blog(ua) < (mn + blog(Sn)) - (m + blog(S)) + (lqx - lqn));
            const float x0 = acc ? xps[col] : xs[col];
            if (acc) {
                double s2 = 0.;
#pragma unroll
                for (int j0 = 0; j0 < D; j0 += 8) {
                    float v[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) { v[j] = xps[(j0 + j) * TW + col]; }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        xs[(j0 + j) * TW + col] = v[j];
                        s2 = fma((double)v[j], (double)v[j], s2);
                    }
                }
#pragma unroll
                for (int k = 0; k < KMAX; ++k) { qx[k * TW + col] = qn[k * TW + col]; }
                m = mn;
                S = Sn;
                sq = s2;
                lqx = lqn;
                ++nacc;
            }
            if (live && St.B > 0) { accumulate(St, w, true, sq, x0, s0, c0, s1, c1, jb, bi); }
            if (st + 1 < nl) { choose(t + 1); }
        }
        else if (st + 1 < nl) { draw_normals(t + 1); }
        __syncthreads();
    }
    if (part == 0) { anyacc[col] = nacc; }
    __syncthreads();
    if (live && anyacc[col] > 0) {
        for (int j = part; j < D; j += NPART) { St.x[(int64_t)j * W + w] = xs[j * TW + col]; }
        St.q[(int64_t)part * W + w] = qx[part * TW + col];
    }
    if (part == 0 && live) {
        St.m[w] = m;
        St.S[w] = S;
        St.sq[w] = sq;
        if (St.B > 0) {
            St.sb[w] = s0;
            St.cb[w] = c0;
            St.sb[W + w] = s1;
            St.cb[W + w] = c1;
            St.jb[w] = jb;
            St.bi[w] = bi;
        }
        nacc_ind[w] += nacc;
    }
}

template <int D>
size_t indep_lds()
{
    return sizeof(float) * ((size_t)3 * D * TW + (size_t)3 * KMAX * TW + (size_t)3 * KMAX * (D + 1) + 2 * KMAX) + sizeof(int) * 3 * TW;
}

template <class R>
__global__ void k_lp_snapshot(State<R> St, int d, float *dst)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (int64_t)d * St.W) { dst[e] = (float)St.x[e]; }
}

__global__ void k_lp_pairs(const float *snap, int nsnap, int d, int64_t W, int B, uint64_t iter, uint32_t k0, uint32_t k1, float same_frac,
                           const float *fixed_ctx, float *ctx, float *tgt)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= B) { return; }
    uint32_t o[4];
    philox(0x30000000u, (uint32_t)p, (uint32_t)iter, (uint32_t)(iter >> 32), k0, k1, o);
    const int s1 = (int)(o[0] % (uint32_t)nsnap);
    const int64_t w1 = (int64_t)(o[1] % (uint64_t)W);
    int s2 = (int)(o[2] % (uint32_t)nsnap);
    const bool same = (float)(o[3] >> 24) * 0x1p-8f < same_frac;
    const int64_t w2 = same ? w1 : (int64_t)((o[3] & 0xffffffu) % (uint64_t)W);
    if (same && nsnap > 1 && s2 == s1) { s2 = (s1 + 1) % nsnap; }
    const float *a = snap + (int64_t)s1 * d * W;
    const float *b = snap + (int64_t)s2 * d * W;
    for (int j = 0; j < d; ++j) {
        ctx[(int64_t)j * B + p] = fixed_ctx ? fixed_ctx[j] : a[(int64_t)j * W + w1];
        tgt[(int64_t)j * B + p] = b[(int64_t)j * W + w2];
    }
}
} // namespace

size_t rw_lds_bytes(int d, int esize) { return (size_t)esize * ((size_t)KMAX * d * (d + 1) + (size_t)KMAX * d + (size_t)d * WG); }

template <class R>
hipError_t launch_lp_init(const Target<R> &T, const State<R> &S, uint32_t k0, uint32_t k1, hipStream_t s)
{
    hipLaunchKernelGGL(k_lp_init<R>, dim3((unsigned)((S.W + WG - 1) / WG)), dim3(WG), sizeof(float) * (size_t)T.d * WG, s, T, S, k0, k1);
    return hipGetLastError();
}

template <class R>
hipError_t launch_lp_rw(const Target<R> &T, const State<R> &S, R step, uint64_t t0, int64_t nsteps, uint32_t k0, uint32_t k1, hipStream_t s)
{
    const size_t lds = rw_lds_bytes(T.d, sizeof(R));
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_lp_rw<R>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL(k_lp_rw<R>, dim3((unsigned)((S.W + WG - 1) / WG)), dim3(WG), rw_lds_bytes(T.d, sizeof(R)), s, T, S, step, t0, nsteps,
                       k0, k1);
    return hipGetLastError();
}

namespace
{
template <int MODE, int D, int H>
hipError_t launch_learned_t(const NetDev &N, const Target<float> &T, const State<float> &S, const float *xin, float *out, uint64_t t,
                            int nl, uint32_t k0, uint32_t k1, int32_t *nacc_ind, hipStream_t s)
{
    const size_t lds = learned_lds<D, H>();
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_lp_learned<MODE, D, H>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL((k_lp_learned<MODE, D, H>), dim3((unsigned)((S.W + TW - 1) / TW)), dim3(TT), lds, s, N, T, S, xin, out, t, nl, k0,
                       k1, nacc_ind);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_learned_m(const NetDev &N, const Target<float> &T, const State<float> &S, const float *xin, float *out, uint64_t t,
                            int nl, uint32_t k0, uint32_t k1, int32_t *nacc_ind, hipStream_t s)
{
#define MCIL_CASE(DD, HH)                                                                                                                  \
    if (N.d == DD && N.H == HH) { return launch_learned_t<MODE, DD, HH>(N, T, S, xin, out, t, nl, k0, k1, nacc_ind, s); }
    MCIL_CASE(8, 64)
    MCIL_CASE(8, 128)
    MCIL_CASE(16, 64)
    MCIL_CASE(16, 128)
    MCIL_CASE(32, 64)
    MCIL_CASE(32, 128)
#undef MCIL_CASE
    return hipErrorInvalidValue;
}
} // namespace

bool learned_supported(int d, int H, int K) { return (d == 8 || d == 16 || d == 32) && (H == 64 || H == 128) && K >= 1 && K <= KMAX; }

hipError_t launch_lp_learned(const NetDev &N, const Target<float> &T, const State<float> &S, uint64_t t0, int nl, uint32_t k0, uint32_t k1,
                             int32_t *nacc_ind, hipStream_t s)
{
    if (!learned_supported(N.d, N.H, N.K) || T.d != N.d || nl < 1) { return hipErrorInvalidValue; }
    return launch_learned_m<1>(N, T, S, nullptr, nullptr, t0, nl, k0, k1, nacc_ind, s);
}

hipError_t launch_lp_state(const Target<float> &T, const State<float> &S, hipStream_t s)
{
    hipLaunchKernelGGL(k_lp_state<float>, dim3((unsigned)((S.W + WG - 1) / WG)), dim3(WG), 0, s, T, S);
    return hipGetLastError();
}

namespace
{
template <int D>
hipError_t launch_indep_t(const IndepMix &M, const Target<float> &T, const State<float> &S, uint64_t t0, int nl, uint32_t k0, uint32_t k1,
                          int32_t *nacc_ind, hipStream_t s)
{
    const size_t lds = indep_lds<D>();
    hipLaunchKernelGGL(k_lp_indep<D>, dim3((unsigned)((S.W + TW - 1) / TW)), dim3(TT), lds, s, M, T, S, t0, nl, k0, k1, nacc_ind);
    return hipGetLastError();
}
} // namespace

hipError_t launch_lp_indep(const IndepMix &M, const Target<float> &T, const State<float> &S, uint64_t t0, int nl, uint32_t k0, uint32_t k1,
                           int32_t *nacc_ind, hipStream_t s)
{
    if (M.K < 1 || M.K > KMAX || nl < 1 || S.W < 1) { return hipErrorInvalidValue; }
    switch (T.d) {
    case 8: return launch_indep_t<8>(M, T, S, t0, nl, k0, k1, nacc_ind, s);
    case 16: return launch_indep_t<16>(M, T, S, t0, nl, k0, k1, nacc_ind, s);
    case 32: return launch_indep_t<32>(M, T, S, t0, nl, k0, k1, nacc_ind, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lp_network(const NetDev &N, const Target<float> &T, const State<float> &S, const float *xin, float *out, hipStream_t s)
{
    if (!learned_supported(N.d, N.H, N.K)) { return hipErrorInvalidValue; }
    return launch_learned_m<0>(N, T, S, xin, out, 0, 0, 0, 0, nullptr, s);
}

template <class R>
hipError_t launch_lp_snapshot(const State<R> &S, int d, float *dst, hipStream_t s)
{
    const int64_t n = (int64_t)d * S.W;
    hipLaunchKernelGGL(k_lp_snapshot<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, d, dst);
    return hipGetLastError();
}

hipError_t launch_lp_pairs(const float *snap, int nsnap, int d, int64_t W, int B, uint64_t iter, uint32_t k0, uint32_t k1, float same_frac,
                           const float *fixed_ctx, float *ctx, float *tgt, hipStream_t s)
{
    hipLaunchKernelGGL(k_lp_pairs, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, snap, nsnap, d, W, B, iter, k0, k1, same_frac, fixed_ctx, ctx, tgt);
    return hipGetLastError();
}

template hipError_t launch_lp_init<double>(const Target<double> &, const State<double> &, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_lp_init<float>(const Target<float> &, const State<float> &, uint32_t, uint32_t, hipStream_t);
template hipError_t launch_lp_rw<double>(const Target<double> &, const State<double> &, double, uint64_t, int64_t, uint32_t, uint32_t,
                                         hipStream_t);
template hipError_t launch_lp_rw<float>(const Target<float> &, const State<float> &, float, uint64_t, int64_t, uint32_t, uint32_t,
                                        hipStream_t);
template hipError_t launch_lp_snapshot<double>(const State<double> &, int, float *, hipStream_t);
template hipError_t launch_lp_snapshot<float>(const State<float> &, int, float *, hipStream_t);
} // namespace mcil
