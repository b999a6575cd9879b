// mci_batch.hip — batched Gaussian-mixture integrals on FP32 walkers.
//
// Every walker is one lane.  A walker of integral i (dimension d, K <= 4 isotropic components
// with log-weights lw_k = log w_k - d/2 log(2 pi s_k^2) and h_k = 1/(2 s_k^2)) keeps
//   x[d]  in LDS as [coordinate][lane] (conflict-free for the random per-lane coordinate),
//   r_k = |x - mu_k|^2, the log-sum-exp shift m = max_k e_k with e_k = lw_k - h_k r_k,
//   S = sum_k exp(e_k - m) in registers, and q = |x|^2 of the FP32 coordinates in FP64 (its
//   per-step update is exact to one FP64 rounding, so the FP32 block sums of q are accurate
//   to FP32 rounding of q itself).
// A single-coordinate move x_j -> x_j' changes r_k by (x_j' - x_j)(x_j' + x_j - 2 mu_kj), so a
// step costs K loads of mu_kj, K + 1 exponentials and no logarithm: the Metropolis test
//   u <= p(x')/p(x) = exp(m' - m) S'/S   is evaluated as   u S < exp(m' - m) S'.
// At every block end r_k, m, S and q are recomputed from x, which bounds the drift of the
// incremental updates to one block.
//
// exp and log are this file's own FP32 routines (bexp, blog) so that the oracle, which
// evaluates the same operations in the same order, reproduces every walker bit for bit.
// FP32 divide and sqrt are IEEE correctly rounded (built with
// -fhip-fp32-correctly-rounded-divide-sqrt), as on the CPU.
//
// Random words: Philox4x32-10, key = seed, counter (c0, walker id, step low, step high):
//   c0 = 0               step words: [0] coordinate (Lemire on [0, d)), [1] proposal, [2] u
//   c0 = 0x20000000 + k/4  Lemire rejection words k of that step
//   c0 = 0x40000000 + n  (step 0) initial draw: call 0 picks the component, calls 1.. feed
//                        Marsaglia's polar method two attempts per call
// Sampling walkers use id = walker index; calibration lanes use id = 2^31 | (CALR i + lane).
#include "mci_batch.hpp"
#include "mci_f32math.hpp"
#include "../../include/mci_amd.h"

#include <float.h>

namespace mcib
{
namespace
{
using mcif::bexp;
using mcif::blog;
using mcif::xor3;
using mcif::sym24;

// Philox4x32-10 with the round keys precomputed (they are the same for every lane and step)
struct Keys
{
    uint32_t k0[10], k1[10];
};

__device__ __forceinline__ Keys make_keys(uint32_t k0, uint32_t k1)
{
    Keys K;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        K.k0[r] = k0 + (uint32_t)r * 0x9E3779B9u;
        K.k1[r] = k1 + (uint32_t)r * 0xBB67AE85u;
    }
    return K;
}

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, const Keys &K, uint32_t o[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, K.k0[r]);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, K.k1[r]);
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

// Kahan-compensated add of y when on (selects, no branch)
__device__ __forceinline__ void kahan_if(float &s, float &c, float y, bool on)
{
    const float yy = y - c;
    const float tt = s + yy;
    const float cc = (tt - s) - yy;
    s = on ? tt : s;
    c = on ? cc : c;
}

// walker state.  Components k >= K are inert: lw = -inf and h = 0 make e_k = -inf and
// exp(e_k - m) = 0 exactly, so a wave can run every lane with the wave's largest K.
struct BW
{
    int d, K;
    uint32_t thr;      // Lemire threshold 2^32 mod d
    float lw[MAXK], h[MAXK];
    const float *mu;   // [K][d]
    float r[MAXK], m, S;
    double q;          // |x|^2 of the FP32 coordinates, kept in FP64
    float step;
    uint32_t id;
    uint64_t t;
};

__device__ __forceinline__ void bw_load(BW &L, const BatchParams &P, int i, uint32_t id)
{
    L.d = P.ndim[i];
    L.K = P.ncomp[i];
    L.thr = (uint32_t)(0u - (uint32_t)L.d) % (uint32_t)L.d;
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        L.lw[k] = (k < L.K) ? P.lw[i * MAXK + k] : -__builtin_huge_valf();
        L.h[k] = (k < L.K) ? P.h[i * MAXK + k] : 0.f;
        L.r[k] = 0.f;
    }
    L.mu = P.mu + P.mu_off[i];
    L.id = id;
    L.t = 0;
    L.m = 0.f;
    L.S = 1.f;
    L.q = 0.;
}

// exact draw from the mixture into this lane's LDS column
__device__ void bw_draw(const BW &L, float *xs, int lane, const BatchParams &P, const Keys &K, int i)
{
    uint32_t o[4];
    philox(0x40000000u, L.id, 0u, 0u, K, o);
    const float uc = (float)(o[0] >> 8) * 0x1p-24f;
    int kc = 0;
#pragma unroll
    for (int k = 0; k < MAXK - 1; ++k) { kc += (k < L.K - 1 && uc >= P.cw[i * MAXK + k]) ? 1 : 0; }
    const float sig = P.sig[i * MAXK + kc];
    const float *muk = L.mu + (int64_t)kc * L.d;
    uint32_t call = 1;
    int j = 0;
    while (j < L.d) {
        philox(0x40000000u + call, L.id, 0u, 0u, K, o);
        ++call;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            if (j < L.d) {
                const float v1 = sym24(o[2 * a]);
                const float v2 = sym24(o[2 * a + 1]);
                const float s = fmaf(v1, v1, v2 * v2);
                if (s < 1.0f) {
                    const float f = sqrtf(-2.0f * blog(s) / s);
                    xs[j * WAVE + lane] = fmaf(sig, v1 * f, muk[j]);
                    ++j;
                    if (j < L.d) {
                        xs[j * WAVE + lane] = fmaf(sig, v2 * f, muk[j]);
                        ++j;
                    }
                }
            }
        }
    }
}

// r_k, m, S and q recomputed from x
__device__ void bw_refresh(BW &L, const float *xs, int lane)
{
    double q = 0.;
    for (int j = 0; j < L.d; ++j) {
        const double v = (double)xs[j * WAVE + lane];
        q = fma(v, v, q);
    }
    L.q = q;
    float e[MAXK];
    float m = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        if (k < L.K) {
            const float *mk = L.mu + (int64_t)k * L.d;
            float rk = 0.f;
            for (int j = 0; j < L.d; ++j) {
                const float dd = xs[j * WAVE + lane] - mk[j];
                rk = fmaf(dd, dd, rk);
            }
            L.r[k] = rk;
            e[k] = fmaf(-L.h[k], rk, L.lw[k]);
            m = fmaxf(m, e[k]);
        }
    }
    float S = 0.f;
#pragma unroll
    for (int k = 0; k < MAXK; ++k) {
        if (k < L.K) { S = S + bexp(e[k] - m); }
    }
    L.m = m;
    L.S = S;
}

// the random words and mixture means one step needs, fetched one step ahead so that the global
// loads of mu_kj are in flight while the previous step computes
struct Pre
{
    uint32_t w1, w2;   // proposal word, acceptance word
    int j;             // coordinate
    float mu[MAXK];    // mu_kj
};

// words of step t (Philox, Lemire index with the rejection words on a wave vote) and the loads
// of mu_kj for the first KW components; every lane of the wave calls it together
template <int KW>
__device__ __forceinline__ void bw_fetch(const BW &L, uint64_t t, Pre &P, const Keys &K)
{
    uint32_t o[4];
    const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32);
    philox(0u, L.id, tl, th, K, o);
    uint64_t prod = (uint64_t)o[0] * (uint32_t)L.d;
    if (__builtin_expect(__any((uint32_t)prod < L.thr), 0)) {
        uint32_t low = (uint32_t)prod;
        uint32_t kx = 0;
        while (low < L.thr) {
            const uint32_t k = 0x80000000u + (kx++);
            uint32_t e[4];
            philox(k >> 2, L.id, tl, th, K, e);
            prod = (uint64_t)e[k & 3u] * (uint32_t)L.d;
            low = (uint32_t)prod;
        }
    }
    P.j = (int)(prod >> 32);
    P.w1 = o[1];
    P.w2 = o[2];
#pragma unroll
    for (int k = 0; k < KW; ++k) {
        const int kk = k < L.K ? k : L.K - 1; // inert components read a valid row
        P.mu[k] = L.mu[(int64_t)kk * L.d + P.j];
    }
}

// one Metropolis step with a single-coordinate uniform move (SRRDVecMove, veclen 1) from the
// prefetched words of step L.t; fetches step L.t + 1 into cur.  Every lane of the wave calls it.
template <int KW>
__device__ __forceinline__ bool bw_step(BW &L, float *xs, int lane, Pre &cur, const Keys &K)
{
    const int j = cur.j;
    const float xo = xs[j * WAVE + lane];
    const float xn = fmaf(L.step, sym24(cur.w1), xo);
    const float ua = (float)(cur.w2 >> 8) * 0x1p-24f;
    float mk[KW];
#pragma unroll
    for (int k = 0; k < KW; ++k) { mk[k] = cur.mu[k]; }
    bw_fetch<KW>(L, L.t + 1, cur, K);
    const float dx = xn - xo, sx = xn + xo;
    float rn[KW], en[KW];
    float mn = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < KW; ++k) {
        rn[k] = fmaf(dx, sx - (mk[k] + mk[k]), L.r[k]);
        en[k] = fmaf(-L.h[k], rn[k], L.lw[k]);
        mn = fmaxf(mn, en[k]);
    }
    float Sn = 0.f;
#pragma unroll
    for (int k = 0; k < KW; ++k) { Sn = Sn + bexp(en[k] - mn); }
    const float E = bexp(mn - L.m);
    const bool acc = ua * L.S < E * Sn;
    if (acc) { xs[j * WAVE + lane] = xn; }
#pragma unroll
    for (int k = 0; k < KW; ++k) { L.r[k] = acc ? rn[k] : L.r[k]; }
    L.m = acc ? mn : L.m;
    L.S = acc ? Sn : L.S;
    // x_j'^2 - x_j^2 of FP32 values, exact products in FP64, one rounding
    const double xn64 = (double)xn, xo64 = (double)xo;
    const double dq = fma(xn64, xn64, -(xo64 * xo64));
    L.q = acc ? L.q + dq : L.q;
    ++L.t;
    return acc;
}

__device__ __forceinline__ int wave_max_int(int v)
{
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const int u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return v;
}

__global__ void k_batch_zero(unsigned int *q, int n)
{
    if ((int)threadIdx.x < n) { q[threadIdx.x] = 0u; }
}

// MCI::findMRT2Step (src/MCIntegrator.cpp:87-168) for one integral with its CALR lanes as CALR
// ranks: MIN_STAT = max(100, sqrt(40000 d)/CALR), the rank-averaged acceptance rate (tree sum
// over the group's lanes), MIN_CONS 5, TOL 0.05, factor in [0.5, 2], step in [FLT_MIN, FLT_MAX]
template <int KW>
__device__ void calibrate_loop(const CalibArgs &A, BW &L, float *xs, int lane, const Keys &K, double &step, double &rate)
{
    const int N = A.nfind;
    const int64_t min_stat = (int64_t)fmax(100., sqrt(40000. * L.d) / (double)CALR);
    int cons = 0, counter = 0;
    Pre cur;
    bw_fetch<KW>(L, L.t, cur, K);
    while ((N < 0 && cons < 5) || counter < N) {
        L.step = (float)step;
        int64_t acc = 0;
        for (int64_t s = 0; s < min_stat; ++s) { acc += bw_step<KW>(L, xs, lane, cur, K) ? 1 : 0; }
        double v = (acc > 0) ? (double)acc / ((double)acc + (double)(min_stat - acc)) : 0.;
        for (int o = CALR / 2; o > 0; o >>= 1) { v = v + __shfl_down(v, o, CALR); }
        v = __shfl(v, 0, CALR);
        rate = v / (double)CALR;
        if (fabs(rate - A.target) < 0.05) { ++cons; }
        else { cons = 0; }
        const double fact = fmin(2., fmax(0.5, rate / A.target));
        step = step * fact;
        if (step > (double)FLT_MAX) { step = (double)FLT_MAX; }
        if (step < (double)FLT_MIN) { step = (double)FLT_MIN; }
        ++counter;
        if (N < 0 && counter >= -N) { break; }
    }
}

// one integral per group of CALR lanes (WAVE / CALR integrals per wave; the host sorts each
// launch's list by component count so that a wave's groups mostly take the same path)
__global__ __launch_bounds__(WAVE) void k_batch_calibrate(CalibArgs A, int nlist)
{
    extern __shared__ float cal_xs[];
    const int lane = threadIdx.x;
    const int gi = blockIdx.x * (WAVE / CALR) + lane / CALR;
    if (gi >= nlist) { return; } // whole groups leave together
    const int i = A.ilist[gi];
    const int l = lane % CALR;
    double step = A.step0[i];
    double rate = 0.;
    if (A.nfind != 0) {
        const Keys K = make_keys(A.P.key0, A.P.key1);
        BW L;
        bw_load(L, A.P, i, 0x80000000u | (uint32_t)(i * CALR + l));
        bw_draw(L, cal_xs, lane, A.P, K, i);
        bw_refresh(L, cal_xs, lane);
        switch (L.K) {
        case 1: calibrate_loop<1>(A, L, cal_xs, lane, K, step, rate); break;
        case 2: calibrate_loop<2>(A, L, cal_xs, lane, K, step, rate); break;
        case 3: calibrate_loop<3>(A, L, cal_xs, lane, K, step, rate); break;
        default: calibrate_loop<4>(A, L, cal_xs, lane, K, step, rate); break;
        }
    }
    if (l == 0) {
        A.step[i] = step;
        A.rate[i] = rate;
    }
}

// one item: ndecorr steps, then T steps with Kahan block sums of |x|^2 flushed as FP64 means
template <int KW>
__device__ void sample_item(const SampleArgs &A, BW &L, float *xs, int lane, const Keys &K, int64_t g, bool live, int64_t T,
                            int64_t Tmax)
{
    const int64_t B = live ? T / A.nblocks : 1;
    Pre cur;
    bw_fetch<KW>(L, L.t, cur, K);
    for (int64_t s = 0; s < A.ndecorr; ++s) { (void)bw_step<KW>(L, xs, lane, cur, K); }
    float Sb = 0.f, Cb = 0.f;
    int64_t jb = 0, b = 0;
    int32_t nacc = 0;
    for (int64_t s = 0; s < Tmax; ++s) {
        const bool acc = bw_step<KW>(L, xs, lane, cur, K);
        const bool on = s < T;
        nacc += (on && acc) ? 1 : 0;
        kahan_if(Sb, Cb, (float)L.q, on);
        jb += on ? 1 : 0;
        if (on && jb == B) {
            A.bmeans[b * A.nwalkers + g] = ((double)Sb - (double)Cb) * (1. / (double)B);
            Sb = 0.f;
            Cb = 0.f;
            jb = 0;
            ++b;
            bw_refresh(L, xs, lane);
        }
    }
    if (live) { A.acc[g] = nacc; }
}

// persistent sampling: each wave takes 64-walker items from its launch's item list until the list
// is exhausted (launches are split by dimension so that each sizes its LDS to its own items)
__global__ __launch_bounds__(WAVE) void k_batch_sample(SampleArgs A)
{
    extern __shared__ float smp_xs[];
    const int lane = threadIdx.x;
    const Keys K = make_keys(A.P.key0, A.P.key1);
    for (;;) {
        unsigned int item = 0u;
        if (lane == 0) { item = atomicAdd(A.queue, 1u); }
        item = __shfl(item, 0);
        if ((int64_t)item >= A.nitems) { break; }
        const int64_t g = (int64_t)A.items[item] * WAVE + lane;
        const bool live = g < A.nwalkers;
        const int i = A.wint[live ? g : A.nwalkers - 1];
        BW L;
        bw_load(L, A.P, i, (uint32_t)g);
        L.step = (float)A.step[i];
        bw_draw(L, smp_xs, lane, A.P, K, i);
        bw_refresh(L, smp_xs, lane);
        const int64_t T = live ? A.wsteps[i] : 0;
        int64_t Tmax = T;
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const int64_t v = __shfl_xor(Tmax, o);
            Tmax = v > Tmax ? v : Tmax;
        }
        switch (wave_max_int(L.K)) {
        case 1: sample_item<1>(A, L, smp_xs, lane, K, g, live, T, Tmax); break;
        case 2: sample_item<2>(A, L, smp_xs, lane, K, g, live, T, Tmax); break;
        case 3: sample_item<3>(A, L, smp_xs, lane, K, g, live, T, Tmax); break;
        default: sample_item<4>(A, L, smp_xs, lane, K, g, live, T, Tmax); break;
        }
    }
}

// the deterministic sum of n values (chunks of 256 summed by a pairwise tree, chunks in
// order), with f applied to each value; every lane of the wave returns it
template <class F>
__device__ double wave_det_sum(int64_t n, int lane, F f)
{
    double total = 0.;
    bool first = true;
    for (int64_t c0 = 0; c0 < n; c0 += 256) {
        double e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t idx = c0 + q * WAVE + lane;
            e[q] = idx < n ? f(idx) : 0.;
        }
        double c = (e[0] + e[2]) + (e[1] + e[3]);
        for (int o = WAVE / 2; o > 0; o >>= 1) { c = c + __shfl_down(c, o); }
        c = __shfl(c, 0);
        total = first ? c : total + c;
        first = false;
    }
    return total;
}

// per integral: avg = sum of walker means / R; err = standard error of the R walker means
// (spread) or sqrt(sum of walker errors^2)/R (MPIMCI, src/MPIMCI.cpp:85-92)
__global__ __launch_bounds__(WAVE) void k_batch_combine(CombineArgs A)
{
    const int i = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t off = A.woff[i], R = A.wcount[i];
    const double *wa = A.wavg + off, *we = A.werr + off;
    const double avg = wave_det_sum(R, lane, [&](int64_t r) { return wa[r]; }) / (double)R;
    double err;
    if (A.mode == MCI_BATCH_ERR_SPREAD && R >= 2) {
        const double ss = wave_det_sum(R, lane, [&](int64_t r) {
            const double dv = wa[r] - avg;
            return dv * dv;
        });
        err = sqrt(ss / ((double)R * (double)(R - 1)));
    }
    else {
        err = sqrt(wave_det_sum(R, lane, [&](int64_t r) { return we[r] * we[r]; })) / (double)R;
    }
    int64_t na = 0;
    for (int64_t r = lane; r < R; r += WAVE) { na += A.acc[off + r]; }
    for (int o = WAVE / 2; o > 0; o >>= 1) { na += __shfl_down(na, o); }
    if (lane == 0) {
        A.avg[i] = avg;
        A.err[i] = err;
        A.accrate[i] = (double)na / ((double)R * (double)A.wsteps[i]);
    }
}
} // namespace

size_t batch_lds_bytes(int dmax) { return sizeof(float) * (size_t)WAVE * (size_t)(dmax > 0 ? dmax : 1); }

int batch_sample_waves_per_cu(int dmax)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_sample, WAVE, batch_lds_bytes(dmax)) != hipSuccess) { return 0; }
    return n;
}

hipError_t launch_batch_zero(unsigned int *q, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_zero, dim3(1), dim3(WAVE), 0, s, q, n);
    return hipGetLastError();
}

hipError_t launch_batch_calibrate(const CalibArgs &a, int nlist, int dmax, hipStream_t s)
{
    const unsigned g = (unsigned)((nlist + WAVE / CALR - 1) / (WAVE / CALR));
    hipLaunchKernelGGL(k_batch_calibrate, dim3(g), dim3(WAVE), batch_lds_bytes(dmax), s, a, nlist);
    return hipGetLastError();
}

hipError_t launch_batch_sample(const SampleArgs &a, int dmax, int nwaves, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_sample, dim3((unsigned)nwaves), dim3(WAVE), batch_lds_bytes(dmax), s, a);
    return hipGetLastError();
}

hipError_t launch_batch_combine(const CombineArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_combine, dim3((unsigned)a.nint), dim3(WAVE), 0, s, a);
    return hipGetLastError();
}
} // namespace mcib
