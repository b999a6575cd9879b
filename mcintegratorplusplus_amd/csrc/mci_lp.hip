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
// Learned move (FP32; k_lp_propose, the network forward, k_lp_accept): the network maps the
// current x to a K-component diagonal Gaussian mixture q(. | x); propose x' ~ q(. | x), evaluate
// the network at x' and accept with min(1, p(x') q(x | x') / (p(x) q(x' | x))), so the target stays
// exact whatever the network computes.  The network's output at the current x is kept per
// walker, so one forward pass (at x') is needed per learned move.
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
using mcif::blog;
using mcif::sym24;

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t o[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
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

// One launch for a whole run of the chain that mixes random-walk moves with the independence
// proposal q(x') = sum_k pi_k N(x'; mu_k, diag sigma_k^2) given by the network at a fixed input.
// Step t is a learned move for every walker when the step's common word says so (probability
// pmix), so the branch is uniform.  A learned move: component k ~ pi, x' = mu_k + sigma_k z,
// accept with min(1, p(x') q(x) / (p(x) q(x'))); q and the target's U_k live in LDS.
constexpr int WGM = 512;

__device__ __forceinline__ float indep_logq(const float *Qs, int Kq, int d, const float xr[DMAX])
{
    constexpr float HALF_LOG_2PI = 0.91893853320467274f;
    float comp[KMAX];
    float cmax = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k >= Kq) { comp[k] = -__builtin_huge_valf(); continue; }
        const float *row = Qs + k * (1 + 3 * d); // [logc][mu d][sigma d][1/sigma d]
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < DMAX; ++j) {
            if (j < d) {
                const float z = (xr[j] - row[1 + j]) * row[1 + 2 * d + j];
                acc = fmaf(z, z, acc);
            }
        }
        comp[k] = row[0] - 0.5f * acc - (float)d * HALF_LOG_2PI;
        cmax = fmaxf(cmax, comp[k]);
    }
    float sc = 0.f;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { sc += (k < Kq) ? __expf(comp[k] - cmax) : 0.f; }
    return cmax + __logf(sc);
}

// the same for a position in LDS (stride ld)
__device__ __forceinline__ float indep_logq_lds(const float *Qs, int Kq, int d, const float *x, int ld)
{
    constexpr float HALF_LOG_2PI = 0.91893853320467274f;
    float comp[KMAX];
    float cmax = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        if (k >= Kq) { comp[k] = -__builtin_huge_valf(); continue; }
        const float *row = Qs + k * (1 + 3 * d);
        float acc = 0.f;
        for (int j = 0; j < d; ++j) {
            const float z = (x[j * ld] - row[1 + j]) * row[1 + 2 * d + j];
            acc = fmaf(z, z, acc);
        }
        comp[k] = row[0] - 0.5f * acc - (float)d * HALF_LOG_2PI;
        cmax = fmaxf(cmax, comp[k]);
    }
    float sc = 0.f;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) { sc += (k < Kq) ? __expf(comp[k] - cmax) : 0.f; }
    return cmax + __logf(sc);
}

__global__ __launch_bounds__(WGM) void k_lp_mix_indep(Target<float> T, State<float> St, const float *Q, int Kq, float step, uint64_t t0,
                                                      int64_t nsteps, float pmix, uint32_t k0, uint32_t k1, int32_t *nacc_ind)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int d = T.d, tid = threadIdx.x;
    float *Ps = reinterpret_cast<float *>(lds_raw); // [KMAX][d][d + 1]
    float *vs = Ps + KMAX * d * (d + 1);             // [KMAX][d]
    float *Us = vs + KMAX * d;                       // [KMAX][d][d]
    float *Qs = Us + KMAX * d * d;                   // [Kq][1 + 3 d], then cumulative weights [Kq]
    float *xs = Qs + Kq * (1 + 3 * d) + Kq;          // [d][WGM]
    const int64_t w = (int64_t)blockIdx.x * WGM + tid;
    const bool live = w < St.W;
    const int64_t wl = live ? w : St.W - 1;
    load_target_lds<float, WGM>(T, Ps, vs);
    for (int e = tid; e < KMAX * d * d; e += WGM) { Us[e] = T.U[e]; }
    for (int e = tid; e < Kq * (1 + 3 * d) + Kq; e += WGM) { Qs[e] = Q[e]; }
    for (int i = 0; i < d; ++i) { xs[i * WGM + tid] = St.x[(int64_t)i * St.W + wl]; }
    __syncthreads();
    const float *cwq = Qs + Kq * (1 + 3 * d);
    Target<float> TL = T;
    TL.U = Us;
    Lane<float> L;
    lane_load(T, St, wl, L);
    for (int64_t s = 0; s < nsteps; ++s) {
        const uint64_t t = t0 + (uint64_t)s;
        const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32);
        uint32_t oc[4];
        philox(0x20000000u, 0xffffffffu, tl, th, k0, k1, oc);
        const bool learned = (float)(oc[0] >> 8) * 0x1p-24f < pmix;
        if (!learned) {
            (void)rw_step<float, WGM>(T, Ps, vs, xs, tid, (uint32_t)wl, L, step, t, k0, k1);
        }
        else {
            uint32_t o[4];
            philox(0x10000000u, (uint32_t)wl, tl, th, k0, k1, o);
            const float uc = (float)(o[0] >> 8) * 0x1p-24f;
            int kc = 0;
            for (int k = 0; k < Kq - 1; ++k) { kc += uc >= cwq[k] ? 1 : 0; }
            const float *row = Qs + kc * (1 + 3 * d);
            const float lq_old = indep_logq_lds(Qs, Kq, d, xs + tid, WGM);
            Normals N;
            N.avail = 0;
            N.has_spare = false;
            N.call = 1;
            float xn[DMAX];
#pragma unroll
            for (int j = 0; j < DMAX; ++j) {
                xn[j] = 0.f;
                if (j < d) {
                    const float z = next_normal(N, 0x10000000u, (uint32_t)wl, tl, th, k0, k1);
                    xn[j] = fmaf(row[1 + d + j], z, row[1 + j]);
                }
            }
            const float lq_new = indep_logq(Qs, Kq, d, xn);
            float qn[KMAX], mn, Sn;
            target_eval_reg(TL, Us, xn, qn, mn, Sn);
            const float ua = ((float)(o[1] >> 8) + 0.5f) * 0x1p-24f;
            const bool acc = __logf(ua) < (mn + __logf(Sn)) - (L.m + __logf(L.S)) + (lq_old - lq_new);
            if (acc) {
                double sq = 0.;
#pragma unroll
                for (int j = 0; j < DMAX; ++j) {
                    if (j < d) {
                        xs[j * WGM + tid] = xn[j];
                        sq = fma((double)xn[j], (double)xn[j], sq);
                    }
                }
#pragma unroll
                for (int k = 0; k < KMAX; ++k) { L.q[k] = qn[k]; }
                L.m = mn;
                L.S = Sn;
                L.sq = sq;
            }
            L.nacc_ind += acc ? 1 : 0;
        }
        after_step<float, WGM>(T, St, Ps, xs, tid, w, live, L);
    }
    if (live) {
        lane_store<float, WGM>(St, d, w, xs, tid, L);
        nacc_ind[w] += L.nacc_ind;
    }
}

// ---- the learned move, fused: one launch per step for all walkers ----
// A workgroup of 256 threads takes 64 walkers (column c = tid % 64, part = tid / 64).  The
// network's mixture at the current x of every walker is kept in a BF16 cache [R][W] (rows per
// component k: d means, d log-scales (already transformed), then K logits); the proposal q(. | x)
// is defined by these rounded values.  A step:
//  1. draw the component from the cached logits and x' = mu + exp(ls) z; log q(x' | x);
//  2. the network at x' on the matrix cores (64 x 64 tiles of v_mfma_f32_32x32x16_bf16, weights
//     read from L2 as [out][in] rows, activations walker-major in LDS), each 64-row chunk of the
//     last layer consumed at once: rounded to BF16 into LDS and folded into log q(x | x');
//  3. log p(x') (components split over the four parts), the MH test, and for accepted walkers the
//     new position, target state and cached mixture.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int TW = 64;    // walkers per workgroup
constexpr int TT = 256;   // threads per workgroup

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// rows r0 .. r0 + 63 of Wt [M][K] (BF16, row = output unit) times act [64 walkers][lda] (BF16, LDS):
// wave (wm, wn) accumulates rows r0 + 32 wm .., walkers 32 wn ..
__device__ __forceinline__ f32x16 mfma_chunk(const uint16_t *Wt, int K, int M, int r0, const uint16_t *act, int lda)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    const int row = r0 + wm * 32 + (lane & 31);
    // the weight fragments of the whole reduction first (K <= 128: at most 8), then the products
    bf16x8 a[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = 16 * q + 8 * (lane >> 5);
        a[q] = bf16x8{};
        if (16 * q < K && row < M) { a[q] = *reinterpret_cast<const bf16x8 *>(Wt + (int64_t)row * K + k); }
    }
    f32x16 acc = {};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (16 * q < K) {
            const int k = 16 * q + 8 * (lane >> 5);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(act + (wn * 32 + (lane & 31)) * lda + k);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q], b, acc, 0, 0, 0);
        }
    }
    return acc;
}

// hidden layer: out[c][r] = bf16(silu(W act + b)) for rows 0 .. M-1
__device__ void hidden_layer(const uint16_t *Wt, const float *bias, int K, int M, const uint16_t *act, int lda, uint16_t *out, int ldo)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wm = w >> 1, wn = w & 1;
    for (int r0 = 0; r0 < M; r0 += 64) {
        const f32x16 acc = mfma_chunk(Wt, K, M, r0, act, lda);
        const int col = wn * 32 + (lane & 31);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = r0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            if (row < M) {
                const float v = acc[i] + bias[row];
                out[col * ldo + row] = f2bf(v / (1.0f + __expf(-v)));
            }
        }
    }
}

template <int MODE> // 0: fill the cache at the current positions, 1: one learned step
__global__ __launch_bounds__(TT) void k_lp_learned(NetDev N, Target<float> T, State<float> St, uint16_t *cache, uint64_t t, uint32_t k0,
                                                   uint32_t k1, int32_t *nacc_ind)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int d = N.d, K = N.K, H = N.H, R = N.O;
    const int tid = threadIdx.x, col = tid & 63, part = tid >> 6;
    const int64_t W = St.W;
    const int64_t w = (int64_t)blockIdx.x * TW + col;
    const bool live = w < W;
    const int64_t wl = live ? w : W - 1;
    const int ldx = N.dp + 8, ldh = H + 8;
    // LDS carve-up (byte offsets multiple of 16)
    float *xs = reinterpret_cast<float *>(lds_raw);            // [DMAX][64] current x
    float *xps = xs + DMAX * TW;                                // [DMAX][64] proposal
    float *part_f = xps + DMAX * TW;                            // [KMAX][4][64] partial log-densities
    float *part_r = part_f + KMAX * 4 * TW;                     // [KMAX][4][64]
    float *qn = part_r + KMAX * 4 * TW;                         // [KMAX][64]
    float *misc = qn + KMAX * TW;                               // [4][64]: log q(x'|x), lse of logits, ...
    int *kc = reinterpret_cast<int *>(misc + 4 * TW);           // [64]
    int *accf = kc + TW;                                        // [64]
    uint16_t *Xb = reinterpret_cast<uint16_t *>(accf + TW);     // [64][ldx]
    uint16_t *H1 = Xb + TW * ldx;                               // [64][ldh]
    uint16_t *H2 = H1 + TW * ldh;                               // [64][ldh]  (before: normals; after: last-layer tile)
    uint16_t *Pn = H2 + TW * ldh;                               // [R][64] mixture at x' (MODE 1)
    float *zs = reinterpret_cast<float *>(H2);                  // [DMAX][64]
    const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32);
    constexpr float HALF_LOG_2PI = 0.91893853320467274f;

    // current positions
    for (int j = part; j < DMAX; j += 4) { xs[j * TW + col] = (j < d) ? St.x[(int64_t)j * W + wl] : 0.f; }
    if (MODE == 0) {
        for (int j = part; j < N.dp; j += 4) { Xb[col * ldx + j] = f2bf(j < d ? xs[j * TW + col] * mcim::IN_SCALE : 0.f); }
    }
    __syncthreads();
    if (MODE == 1) {
        // 1. component and proposal from the cached mixture at x
        if (part == 0) {
            float amax = -__builtin_huge_valf();
            for (int k = 0; k < K; ++k) { amax = fmaxf(amax, bf2f(cache[(int64_t)(2 * d * K + k) * W + wl])); }
            float sa = 0.f;
            for (int k = 0; k < K; ++k) { sa += __expf(bf2f(cache[(int64_t)(2 * d * K + k) * W + wl]) - amax); }
            uint32_t o[4];
            philox(0x10000000u, (uint32_t)wl, tl, th, k0, k1, o);
            const float uc = (float)(o[0] >> 8) * 0x1p-24f * sa;
            int kk = K - 1;
            float cum = 0.f;
            for (int k = 0; k < K - 1; ++k) {
                cum += __expf(bf2f(cache[(int64_t)(2 * d * K + k) * W + wl]) - amax);
                if (uc < cum) { kk = k; break; }
            }
            kc[col] = kk;
            misc[1 * TW + col] = amax + __logf(sa); // lse of the logits at x
        }
        {
            Normals Nm;
            Nm.avail = 0;
            Nm.has_spare = false;
            Nm.call = 1;
            const uint32_t base = 0x10000000u + ((uint32_t)(part + 1) << 20);
            for (int j = part; j < d; j += 4) { zs[j * TW + col] = next_normal(Nm, base, (uint32_t)wl, tl, th, k0, k1); }
        }
        __syncthreads();
        {
            const int kk = kc[col];
            for (int j = part; j < DMAX; j += 4) {
                float v = 0.f;
                if (j < d) {
                    const float mu = bf2f(cache[(int64_t)(2 * d * kk + j) * W + wl]);
                    const float ls = bf2f(cache[(int64_t)(2 * d * kk + d + j) * W + wl]);
                    v = fmaf(__expf(ls), zs[j * TW + col], mu);
                }
                xps[j * TW + col] = v;
                if (j < N.dp) { Xb[col * ldx + j] = f2bf(v * mcim::IN_SCALE); }
            }
        }
        __syncthreads();
        // log q(x' | x): part p sums coordinates j = p, p + 4, ...
        for (int k = 0; k < K; ++k) {
            float acc = 0.f;
            for (int j = part; j < d; j += 4) {
                const float mu = bf2f(cache[(int64_t)(2 * d * k + j) * W + wl]);
                const float ls = bf2f(cache[(int64_t)(2 * d * k + d + j) * W + wl]);
                const float z = (xps[j * TW + col] - mu) * __expf(-ls);
                acc += -ls - 0.5f * z * z;
            }
            part_f[(k * 4 + part) * TW + col] = acc;
        }
        __syncthreads();
        if (part == 0) {
            const float lsa = misc[1 * TW + col];
            float comp[KMAX];
            float cmax = -__builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                comp[k] = -__builtin_huge_valf();
                if (k < K) {
                    const float sm = ((part_f[(k * 4 + 0) * TW + col] + part_f[(k * 4 + 1) * TW + col]) + part_f[(k * 4 + 2) * TW + col]) +
                                     part_f[(k * 4 + 3) * TW + col];
                    comp[k] = bf2f(cache[(int64_t)(2 * d * K + k) * W + wl]) - lsa + sm - (float)d * HALF_LOG_2PI;
                    cmax = fmaxf(cmax, comp[k]);
                }
            }
            float sc = 0.f;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) { sc += (k < K) ? __expf(comp[k] - cmax) : 0.f; }
            misc[0 * TW + col] = cmax + __logf(sc); // log q(x' | x)
        }
    }
    // 2. the network at Xb (x' for a step, x for a fill)
    hidden_layer(N.W1t, N.b1, N.dp, H, Xb, ldx, H2, ldh);
    __syncthreads();
    hidden_layer(N.W2t, N.b2, H, H, H2, ldh, H1, ldh);
    __syncthreads();
    // H1 now holds layer 2; the last layer's tile goes to Cs, which overlaps H2
    float *Ct = reinterpret_cast<float *>(H2);
    for (int r0 = 0; r0 < R; r0 += 64) {
        const f32x16 acc = mfma_chunk(N.W3t, H, N.Op, r0, H1, ldh);
        {
            const int lane = tid & 63, wv = tid >> 6, wm = wv >> 1, wn = wv & 1;
            const int c = wn * 32 + (lane & 31);
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int rr = wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                Ct[rr * (TW + 1) + c] = acc[i] + ((r0 + rr < N.Op) ? N.b3[r0 + rr] : 0.f);
            }
        }
        __syncthreads();
        // transform and round: means x MU_SCALE, log-scales LS_MAX tanh(o / LS_MAX), logits as they are
        for (int rr = part; rr < 64; rr += 4) {
            const int r = r0 + rr;
            if (r >= R) { break; }
            const float v = Ct[rr * (TW + 1) + col];
            float val = v;
            if (r < 2 * d * K) { val = ((r % (2 * d)) < d) ? mcim::MU_SCALE * v : mcim::LS_MAX * tanhf(v / mcim::LS_MAX); }
            const uint16_t b = f2bf(val);
            if (MODE == 0) {
                if (live) { cache[(int64_t)r * W + w] = b; }
            }
            else { Pn[r * TW + col] = b; }
        }
        if (MODE == 1) {
            __syncthreads();
            // fold this chunk's components into log q(x | x'): part covers 8 coordinate pairs
            // (component k occupies rows 2 d k .. 2 d k + 2 d - 1, so pair g = (component g / d,
            // coordinate g % d) and a chunk of 64 rows holds 32 consecutive pairs; d is a multiple of 8)
            const int g0 = (r0 >> 1) + part * 8;
            const int k = g0 / d, j0 = g0 % d;
            if (k < K) {
                float acc = 0.f;
                for (int q = 0; q < 8; ++q) {
                    const int j = j0 + q;
                    const float mu = bf2f(Pn[(2 * d * k + j) * TW + col]);
                    const float ls = bf2f(Pn[(2 * d * k + d + j) * TW + col]);
                    const float z = (xs[j * TW + col] - mu) * __expf(-ls);
                    acc += -ls - 0.5f * z * z;
                }
                part_r[(k * 4 + (j0 >> 3)) * TW + col] = acc;
            }
        }
        __syncthreads();
    }
    if (MODE == 0) { return; }
    // 3. log q(x | x') and the target at x'
    if (part == 0) {
        float amax = -__builtin_huge_valf();
        for (int k = 0; k < K; ++k) { amax = fmaxf(amax, bf2f(Pn[(2 * d * K + k) * TW + col])); }
        float sa = 0.f;
        for (int k = 0; k < K; ++k) { sa += __expf(bf2f(Pn[(2 * d * K + k) * TW + col]) - amax); }
        const float lsa = amax + __logf(sa);
        const int nsub = d / 8;
        float comp[KMAX];
        float cmax = -__builtin_huge_valf();
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            comp[k] = -__builtin_huge_valf();
            if (k < K) {
                float sm = 0.f;
                for (int sp = 0; sp < nsub; ++sp) { sm += part_r[(k * 4 + sp) * TW + col]; }
                comp[k] = bf2f(Pn[(2 * d * K + k) * TW + col]) - lsa + sm - (float)d * HALF_LOG_2PI;
                cmax = fmaxf(cmax, comp[k]);
            }
        }
        float sc = 0.f;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) { sc += (k < K) ? __expf(comp[k] - cmax) : 0.f; }
        misc[2 * TW + col] = cmax + __logf(sc); // log q(x | x')
    }
    {
        // components 2 part, 2 part + 1 of the target
        float xr[DMAX];
#pragma unroll
        for (int j = 0; j < DMAX; ++j) { xr[j] = xps[j * TW + col]; }
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int k = 2 * part + kk;
            const float *Uk = T.U + (int64_t)k * T.d * T.d;
            const float *bk = T.Um + k * T.d;
            float qk = 0.f;
#pragma unroll
            for (int i = 0; i < DMAX; ++i) {
                if (i < T.d) {
                    float r = -bk[i];
#pragma unroll
                    for (int j = i; j < DMAX; ++j) {
                        if (j < T.d) { r = fmaf(Uk[i * T.d + j], xr[j], r); }
                    }
                    qk = fmaf(r, r, qk);
                }
            }
            qn[k * TW + col] = qk;
        }
    }
    __syncthreads();
    if (part == 0) {
        float mn = -__builtin_huge_valf();
        float e[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            e[k] = T.c[k] - 0.5f * qn[k * TW + col];
            mn = fmaxf(mn, e[k]);
        }
        float Sn = 0.f;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) { Sn += __expf(e[k] - mn); }
        uint32_t o[4];
        philox(0x10000000u, (uint32_t)wl, tl, th, k0, k1, o);
        const float ua = ((float)(o[1] >> 8) + 0.5f) * 0x1p-24f;
        const float logp_old = St.m[wl] + __logf(St.S[wl]);
        const bool acc = live && (__logf(ua) < (mn + __logf(Sn)) - logp_old + (misc[2 * TW + col] - misc[0 * TW + col]));
        accf[col] = acc ? 1 : 0;
        if (acc) {
            double sq = 0.;
            for (int j = 0; j < d; ++j) {
                const float v = xps[j * TW + col];
                St.x[(int64_t)j * W + w] = v;
                sq = fma((double)v, (double)v, sq);
            }
#pragma unroll
            for (int k = 0; k < KMAX; ++k) { St.q[(int64_t)k * W + w] = qn[k * TW + col]; }
            St.m[w] = mn;
            St.S[w] = Sn;
            St.sq[w] = sq;
            nacc_ind[w] += 1;
        }
        if (live && St.B > 0) {
            float s0 = St.sb[w], c0 = St.cb[w], s1 = St.sb[W + w], c1 = St.cb[W + w];
            int jb = St.jb[w], bi = St.bi[w];
            accumulate(St, w, true, St.sq[w], acc ? xps[col] : xs[col], s0, c0, s1, c1, jb, bi);
            St.sb[w] = s0;
            St.cb[w] = c0;
            St.sb[W + w] = s1;
            St.cb[W + w] = c1;
            St.jb[w] = jb;
            St.bi[w] = bi;
        }
    }
    __syncthreads();
    if (accf[col]) {
        for (int r = part; r < R; r += 4) { cache[(int64_t)r * W + w] = Pn[r * TW + col]; }
    }
}

template <class R>
__global__ void k_lp_snapshot(State<R> St, int d, float *dst)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (int64_t)d * St.W) { dst[e] = (float)St.x[e]; }
}

__global__ void k_lp_pairs(const float *snap, int nsnap, int d, int64_t W, int B, uint64_t iter, uint32_t k0, uint32_t k1, float same_frac,
                           float *ctx, float *tgt)
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
        ctx[(int64_t)j * B + p] = a[(int64_t)j * W + w1];
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

size_t mix_lds_bytes(int d, int Kq)
{
    return sizeof(float) * ((size_t)KMAX * d * (d + 1) + (size_t)KMAX * d + (size_t)KMAX * d * d + (size_t)Kq * (1 + 3 * d) + Kq +
                            (size_t)d * WGM);
}

hipError_t launch_lp_mix_indep(const Target<float> &T, const State<float> &S, const float *Q, int Kq, float step, uint64_t t0, int64_t nsteps,
                               float pmix, uint32_t k0, uint32_t k1, int32_t *nacc_ind, hipStream_t s)
{
    const size_t lds = mix_lds_bytes(T.d, Kq);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_lp_mix_indep), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL(k_lp_mix_indep, dim3((unsigned)((S.W + WGM - 1) / WGM)), dim3(WGM), lds, s, T, S, Q, Kq, step, t0, nsteps, pmix, k0, k1,
                       nacc_ind);
    return hipGetLastError();
}

size_t learned_lds_bytes(const NetDev &N)
{
    const size_t f = sizeof(float) * ((size_t)2 * DMAX * TW + (size_t)2 * KMAX * 4 * TW + (size_t)KMAX * TW + 4 * TW) + sizeof(int) * 2 * TW;
    const size_t hb = sizeof(uint16_t) * ((size_t)TW * (N.dp + 8) + (size_t)2 * TW * (N.H + 8) + (size_t)N.O * TW);
    return f + hb;
}

hipError_t launch_lp_learned(int mode, const NetDev &N, const Target<float> &T, const State<float> &S, uint16_t *cache, uint64_t t,
                             uint32_t k0, uint32_t k1, int32_t *nacc_ind, hipStream_t s)
{
    const size_t lds = learned_lds_bytes(N);
    const dim3 grid((unsigned)((S.W + TW - 1) / TW));
    if (mode == 0) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_lp_learned<0>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) { return e; }
        hipLaunchKernelGGL(k_lp_learned<0>, grid, dim3(TT), lds, s, N, T, S, cache, t, k0, k1, nacc_ind);
    }
    else {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_lp_learned<1>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) { return e; }
        hipLaunchKernelGGL(k_lp_learned<1>, grid, dim3(TT), lds, s, N, T, S, cache, t, k0, k1, nacc_ind);
    }
    return hipGetLastError();
}

template <class R>
hipError_t launch_lp_snapshot(const State<R> &S, int d, float *dst, hipStream_t s)
{
    const int64_t n = (int64_t)d * S.W;
    hipLaunchKernelGGL(k_lp_snapshot<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, d, dst);
    return hipGetLastError();
}

hipError_t launch_lp_pairs(const float *snap, int nsnap, int d, int64_t W, int B, uint64_t iter, uint32_t k0, uint32_t k1, float same_frac,
                           float *ctx, float *tgt, hipStream_t s)
{
    hipLaunchKernelGGL(k_lp_pairs, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, snap, nsnap, d, W, B, iter, k0, k1, same_frac, ctx, tgt);
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
