// mci_fast.hip — specialised Metropolis kernels for the configuration family
//   single-coordinate uniform moves (SRRDVecMove, veclen 1, one step-size type)
//   on one diagonal Gaussian (MCI_PDF_GAUSS_DIAG), unbounded domain, Philox streams,
//   observables from {X2, X2_SUM, EXP_X2} with Block or Simple accumulators (nskip 1).
//
// Layout: one wave (64 walkers) per workgroup.  The walker coordinates and, for X2, the
// per-coordinate block sums live in LDS as [coordinate][lane]: lane l touches bank l only, so
// the random per-lane coordinate index never conflicts.  Per step a walker reads one
// coordinate, and touches LDS again only when the move is accepted.
//
// Block sums by dwell time.  With x_i piecewise constant, the block sum of x_i^2 over steps
// j = 0..B-1 is  B x_i(start)^2 + sum over accepted changes at step j of (x_new^2 - x_old^2)(B - j);
// |x|^2 and exp(-|x|^2/c) are handled the same way.  This is the same block mean the
// reference accumulates step by step (BlockAccumulator.cpp:20-30), rounded differently.
//
// FP64 acceptance: u <= exp(x) is decided by exp in FP32 whenever u lies outside a 4e-5
// relative band around it, else by the FP64 exp — the decision is that of the FP64 test.
// Random words per step come from one Philox4x32-10 call with the counter layout of the
// generic kernels (word 0 index, 1 proposal, 2-3 acceptance), so decisions and positions are
// identical to mci_kernels.hip's k_sample.
//
// FP32 precision: coordinates, proposal (24-bit symmetric grid), log-ratio and exp in FP32,
// acceptance uniform on a 2^-24 grid, block sums as Kahan-compensated FP32 pairs, flushed to
// FP64 for the FP64 estimators.
#include "mci_fast.hpp"

#include <float.h>

namespace mcid
{
namespace
{
constexpr int WAVE = 64;

struct FastArgs
{
    int64_t W;
    int ndim;
    int64_t nsteps;
    uint32_t key0, key1;
    uint64_t *ctr;         // [W]
    double *x64;           // [ndim][W] (FP64 state, also FP32 mode's persistent copy)
    const double *a;       // [ndim] p_i = a_i x_i^2
    const double *steps;   // [W] (one type)
    int64_t *acc, *rej;    // [W]
    int blocksize;         // B (Simple: nsteps)
    int64_t nblocks;
    double *d_x2;          // X2 data [nblocks][ndim][W] or nullptr
    double *d_x2sum;       // X2_SUM data [nblocks][1][W] or nullptr
    double *d_exp;         // EXP_X2 data [nblocks][1][W] or nullptr
    double x2sum_div;      // X2_SUM divisor
    double exp_div;        // EXP_X2 divisor
    int nvecs;
    const int *skip;       // nonzero *skip: return at once (finished calibration loop)
};

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t &o0, uint32_t &o1, uint32_t &o2, uint32_t &o3)
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
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

// Lemire index on [0, n) from word w0 with rejection words from the extra stream
__device__ __forceinline__ int draw_vidx(uint32_t w0, int n, uint32_t wid, uint64_t t, uint32_t k0, uint32_t k1)
{
    const uint32_t range = (uint32_t)n;
    uint64_t prod = (uint64_t)w0 * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        const uint32_t thr = (0u - range) % range;
        uint32_t kx = 0;
        while (low < thr) {
            const uint32_t k = 0x80000000u + (kx++);
            uint32_t o[4];
            philox(k >> 2, wid, (uint32_t)t, (uint32_t)(t >> 32), k0, k1, o[0], o[1], o[2], o[3]);
            prod = (uint64_t)o[k & 3u] * range;
            low = (uint32_t)prod;
        }
    }
    return (int)(prod >> 32);
}

template <bool ACC>
__global__ __launch_bounds__(WAVE) void k_fast64(FastArgs A)
{
    extern __shared__ double lds[];
    const int d = A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    if (A.skip && *A.skip) { return; }
    const bool live = w < A.W;
    double *xs = lds;                 // [d][64]
    double *Ss = lds + d * WAVE;      // [d][64] block sums of x_i^2 (X2)
    double *as = lds + (ACC ? 2 : 1) * d * WAVE;  // [d]
    for (int i = lane; i < d; i += WAVE) { as[i] = A.a[i]; }
    double r = 0.0;                   // |x|^2
    if (live) {
        for (int i = 0; i < d; ++i) {
            const double v = A.x64[(int64_t)i * A.W + w];
            xs[i * WAVE + lane] = v;
            r = r + v * v;
        }
    }
    __syncthreads();
    if (!live) { return; }
    const double step = A.steps[w];
    uint64_t t = A.ctr[w];
    const uint32_t wid = (uint32_t)w;
    int64_t nacc = 0;
    const bool hx2 = ACC && A.d_x2;
    const bool hsum = ACC && A.d_x2sum;
    const bool hexp = ACC && A.d_exp;
    const int B = A.blocksize;
    double Sr = 0.0, Se = 0.0, e = 0.0;
    int j = 0;
    int64_t b = 0;
    if (ACC) {
        if (hx2) { for (int i = 0; i < d; ++i) { const double v = xs[i * WAVE + lane]; Ss[i * WAVE + lane] = (double)B * (v * v); } }
        Sr = (double)B * r;
        if (hexp) { e = exp(-(r / A.exp_div)); Se = (double)B * e; }
    }
    // Two steps per iteration.  Both steps' LDS operands are read before either write-back;
    // when the second step moves the same coordinate as the first, the first step's new
    // values are forwarded by select (what an LDS round trip would return).  The Philox words
    // of the next two steps are generated meanwhile: two independent calls give the wave
    // enough independent instructions to issue while the decision chains wait.
    // The common path has no divergent branch: the decision is formed with selects and the
    // LDS state is always written back (old value on rejection; adding +0.0 to a sum is
    // exact, so every stored number equals the branched form).
    double mult = (double)B;  // B - j: steps of the block the new value counts for

    // decision and write-back of one step; returns whether a block was flushed after it
    auto step_core = [&](int slot, double xo, double av, double So, uint32_t w1, uint32_t w2, uint32_t w3,
                         double &xout, double &Sout) -> bool {
        // (2 w1 + 1 - 2^32) 2^-32 and the top 53 bits of (w2:w3) 2^-53, both exact
        const double rr = ((double)(int32_t)(w1 - 0x80000000u) + 0.5) * 0x1p-31;
        const double u = (double)w2 * 0x1p-32 + (double)(w3 >> 11) * 0x1p-53;
        const double xn = xo + step * rr;
        const double pn = (xn * xn) * av, po = (xo * xo) * av;
        const double ex = pn - po;
        // u <= exp(-ex): FP32 exp outside a 4e-5 band, FP64 exp inside it
        const double f = (double)__expf((float)(-ex));
        bool acc = (ex <= 0.0) || (u < f * (1.0 - 4e-5));
        const bool band = !acc && !(u > f * (1.0 + 4e-5));
        if (__builtin_expect(__any(band), 0)) {
            if (band) { acc = u <= exp(-ex); }
        }
        nacc += acc ? 1 : 0;
        xout = acc ? xn : xo;
        xs[slot] = xout;
        Sout = So;
        if (ACC) {
            const double dx2 = acc ? (xn * xn - xo * xo) : 0.0;
            if (hx2) { Sout = So + dx2 * mult; Ss[slot] = Sout; }
            r = r + dx2;
            Sr = Sr + dx2 * mult;
            if (hexp && acc) { const double en = exp(-(r / A.exp_div)); Se = Se + (en - e) * mult; e = en; }
            mult = mult - 1.0;
            if (++j == B) {
                const double norm = 1. / B;
                double rx = 0.0;
                for (int i = 0; i < d; ++i) {
                    const double v2 = xs[i * WAVE + lane];
                    const double q = v2 * v2;
                    rx = rx + q;
                    if (hx2) {
                        A.d_x2[(b * d + i) * A.W + w] = Ss[i * WAVE + lane] * norm;
                        Ss[i * WAVE + lane] = (double)B * q;
                    }
                }
                if (hsum) { A.d_x2sum[b * A.W + w] = (Sr * norm) / A.x2sum_div; }
                if (hexp) { A.d_exp[b * A.W + w] = Se * norm; }
                r = rx;  // refresh |x|^2 exactly at every block boundary
                Sr = (double)B * r;
                if (hexp) { e = exp(-(r / A.exp_div)); Se = (double)B * e; }
                j = 0;
                mult = (double)B;
                ++b;
                return true;
            }
        }
        return false;
    };

    uint32_t a0, a1, a2, a3, b0, b1, b2, b3;
    philox(0u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, a0, a1, a2, a3);
    philox(0u, wid, (uint32_t)(t + 1), (uint32_t)((t + 1) >> 32), A.key0, A.key1, b0, b1, b2, b3);
    int64_t s = 0;
    for (; s + 1 < A.nsteps; s += 2) {
        const int va = draw_vidx(a0, A.nvecs, wid, t, A.key0, A.key1);
        const int vb = draw_vidx(b0, A.nvecs, wid, t + 1, A.key0, A.key1);
        const int sa = va * WAVE + lane, sb = vb * WAVE + lane;
        const double xoa = xs[sa], ava = as[va];
        const double xob_l = xs[sb], avb = as[vb];
        const double Soa = hx2 ? Ss[sa] : 0.0;
        const double Sob_l = hx2 ? Ss[sb] : 0.0;
        const uint32_t pa1 = a1, pa2 = a2, pa3 = a3, pb1 = b1, pb2 = b2, pb3 = b3;
        t += 2;
        philox(0u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, a0, a1, a2, a3);
        philox(0u, wid, (uint32_t)(t + 1), (uint32_t)((t + 1) >> 32), A.key0, A.key1, b0, b1, b2, b3);
        double xa, Sa;
        const bool flushed = step_core(sa, xoa, ava, Soa, pa1, pa2, pa3, xa, Sa);
        const bool same = (vb == va);
        const double xob = same ? xa : xob_l;
        double Sob = same ? Sa : Sob_l;
        if (hx2 && flushed) { Sob = Ss[sb]; }  // a block ended between the two steps
        double xb, Sb;
        step_core(sb, xob, avb, Sob, pb1, pb2, pb3, xb, Sb);
    }
    if (s < A.nsteps) {  // odd step count: the last step alone (its words are in a0..a3)
        const int va = draw_vidx(a0, A.nvecs, wid, t, A.key0, A.key1);
        const int sa = va * WAVE + lane;
        double xa, Sa;
        step_core(sa, xs[sa], as[va], hx2 ? Ss[sa] : 0.0, a1, a2, a3, xa, Sa);
        ++t;
    }
    for (int i = 0; i < d; ++i) { A.x64[(int64_t)i * A.W + w] = xs[i * WAVE + lane]; }
    A.ctr[w] = t;
    A.acc[w] = nacc;
    A.rej[w] = A.nsteps - nacc;
}

// Kahan-compensated FP32 accumulation: s + y with compensation c
__device__ __forceinline__ void kahan(float &s, float &c, float y)
{
    const float yy = y - c;
    const float tt = s + yy;
    c = (tt - s) - yy;
    s = tt;
}

template <bool ACC>
__global__ __launch_bounds__(WAVE) void k_fast32(FastArgs A)
{
    extern __shared__ float lds32[];
    const int d = A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    if (A.skip && *A.skip) { return; }
    const bool live = w < A.W;
    float *xs = lds32;                  // [d][64]
    float *Ss = lds32 + d * WAVE;       // [d][64] Kahan sums of x_i^2
    float *Cs = lds32 + 2 * d * WAVE;   // [d][64] their compensations
    float *as = lds32 + (ACC ? 3 : 1) * d * WAVE;   // [d]
    for (int i = lane; i < d; i += WAVE) { as[i] = (float)A.a[i]; }
    float r = 0.f, rc = 0.f;            // |x|^2, Kahan-compensated
    if (live) {
        for (int i = 0; i < d; ++i) {
            const float v = (float)A.x64[(int64_t)i * A.W + w];
            xs[i * WAVE + lane] = v;
            kahan(r, rc, v * v);
        }
    }
    __syncthreads();
    if (!live) { return; }
    const float step = (float)A.steps[w];
    uint64_t t = A.ctr[w];
    const uint32_t wid = (uint32_t)w;
    int64_t nacc = 0;
    const bool hx2 = ACC && A.d_x2;
    const bool hsum = ACC && A.d_x2sum;
    const bool hexp = ACC && A.d_exp;
    const int B = A.blocksize;
    const float inv_ediv = (float)(1.0 / A.exp_div);
    float Sr = 0.f, Cr = 0.f, Se = 0.f, Ce = 0.f, e = 0.f;
    int j = 0;
    int64_t b = 0;
    if (ACC) {
        if (hx2) {
            for (int i = 0; i < d; ++i) { const float v = xs[i * WAVE + lane]; Ss[i * WAVE + lane] = (float)B * (v * v); Cs[i * WAVE + lane] = 0.f; }
        }
        Sr = (float)B * r;
        if (hexp) { e = __expf(-r * inv_ediv); Se = (float)B * e; }
    }
    uint32_t w0, w1, w2, w3;
    philox(0u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, w0, w1, w2, w3);
    float mult = (float)B;
    for (int64_t s = 0; s < A.nsteps; ++s) {
        const int v = draw_vidx(w0, A.nvecs, wid, t, A.key0, A.key1);
        const int slot = v * WAVE + lane;
        const float xo = xs[slot];
        const float asv = as[v];
        float So = 0.f, Co = 0.f;
        if (hx2) { So = Ss[slot]; Co = Cs[slot]; }
        const float rr = (float)(2 * (int32_t)(w1 >> 8) + 1 - 16777216) * 0x1p-24f;
        const float u = (float)(w2 >> 8) * 0x1p-24f;
        ++t;
        philox(0u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, w0, w1, w2, w3);
        const float xn = fmaf(step, rr, xo);
        const float dx2 = (xn - xo) * (xn + xo);
        const float ex = asv * dx2;
        const bool acc = u <= __expf(-ex);
        nacc += acc ? 1 : 0;
        xs[slot] = acc ? xn : xo;
        if (ACC) {
            const float da = acc ? dx2 : 0.f;
            if (hx2) {
                if (acc) { kahan(So, Co, da * mult); }
                Ss[slot] = So;
                Cs[slot] = Co;
            }
            if (acc) {
                kahan(r, rc, da);
                kahan(Sr, Cr, da * mult);
                if (hexp) { const float en = __expf(-r * inv_ediv); kahan(Se, Ce, (en - e) * mult); e = en; }
            }
            mult = mult - 1.f;
            if (++j == B) {
                const double norm = 1. / B;
                float rx = 0.f, rxc = 0.f;
                for (int i = 0; i < d; ++i) {
                    const float v2 = xs[i * WAVE + lane];
                    const float q = v2 * v2;
                    kahan(rx, rxc, q);
                    if (hx2) {
                        A.d_x2[(b * d + i) * A.W + w] = ((double)Ss[i * WAVE + lane] - (double)Cs[i * WAVE + lane]) * norm;
                        Ss[i * WAVE + lane] = (float)B * q;
                        Cs[i * WAVE + lane] = 0.f;
                    }
                }
                if (hsum) { A.d_x2sum[b * A.W + w] = (((double)Sr - (double)Cr) * norm) / A.x2sum_div; }
                if (hexp) { A.d_exp[b * A.W + w] = ((double)Se - (double)Ce) * norm; }
                r = rx;
                rc = rxc;
                Sr = (float)B * r;
                Cr = 0.f;
                if (hexp) { e = __expf(-r * inv_ediv); Se = (float)B * e; Ce = 0.f; }
                j = 0;
                mult = (float)B;
                ++b;
            }
        }
    }
    for (int i = 0; i < d; ++i) { A.x64[(int64_t)i * A.W + w] = (double)xs[i * WAVE + lane]; }
    A.ctr[w] = t;
    A.acc[w] = nacc;
    A.rej[w] = A.nsteps - nacc;
}

size_t lds_bytes(int prec, int ndim, bool acc)
{
    if (prec == MCI_PREC_FP64) { return sizeof(double) * ((size_t)(acc ? 2 : 1) * ndim * WAVE + ndim); }
    return sizeof(float) * ((size_t)(acc ? 3 : 1) * ndim * WAVE + ndim);
}

size_t lds_limit()
{
    static size_t lim = 0;
    if (lim == 0) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) { lim = p.sharedMemPerBlock; }
        if (lim == 0) { lim = 65536; }
    }
    return lim;
}
} // namespace

bool fast_eligible(int prec, int rng, int movetype, int veclen, int ntypes, int periodic, int npdf, int pdfkind)
{
    (void)prec;
    return rng == MCI_RNG_PHILOX && movetype == MCI_MOVE_VEC && veclen == 1 && ntypes == 1 && !periodic && npdf == 1 &&
           pdfkind == MCI_PDF_GAUSS_DIAG;
}

bool fast_obs_eligible(const ObsDev *obs, int nobs, int ndim, int prec)
{
    int B = -1;
    int seen = 0;
    for (int i = 0; i < nobs; ++i) {
        const ObsDev &o = obs[i];
        if (o.nskip != 1 || o.acc == ACC_FULL) { return false; }
        if (o.kind != MCI_OBS_X2 && o.kind != MCI_OBS_X2_SUM && o.kind != MCI_OBS_EXP_X2) { return false; }
        const int bit = 1 << o.kind;
        if (seen & bit) { return false; }
        seen |= bit;
        const int bs = (int)(o.acc == ACC_SIMPLE ? -2 : o.blocksize);
        if (B == -1) { B = bs; }
        else if (B != bs) { return false; }
    }
    return lds_bytes(prec, ndim, nobs > 0) <= lds_limit();
}

int fast_sample(FastState &fs, const Chain &c, int prec, int64_t nsteps, const ObsDev *obs, int nobs, hipStream_t s, const int *skip)
{
    (void)fs;
    FastArgs A;
    A.skip = skip;
    A.W = c.W;
    A.ndim = c.ndim;
    A.nsteps = nsteps;
    A.key0 = c.key0;
    A.key1 = c.key1;
    A.ctr = c.ctr;
    A.x64 = c.xold;
    A.a = c.pdf[0].a;
    A.steps = c.steps;
    A.acc = c.acc;
    A.rej = c.rej;
    A.nvecs = c.nvecs;
    A.blocksize = 1;
    A.nblocks = 0;
    A.d_x2 = A.d_x2sum = A.d_exp = nullptr;
    A.x2sum_div = 1.0;
    A.exp_div = (double)c.ndim;
    const bool acc = obs != nullptr && nobs > 0;
    for (int i = 0; i < nobs && obs; ++i) {
        const ObsDev &o = obs[i];
        A.blocksize = (o.acc == ACC_SIMPLE) ? (int)nsteps : o.blocksize;
        A.nblocks = o.nstore;
        if (o.kind == MCI_OBS_X2) { A.d_x2 = o.data; }
        if (o.kind == MCI_OBS_X2_SUM) { A.d_x2sum = o.data; A.x2sum_div = o.p0; }
        if (o.kind == MCI_OBS_EXP_X2) { A.d_exp = o.data; A.exp_div = o.p0; }
    }
    if (!A.a) { return MCI_E_INVALID_ARGUMENT; } // the fast path needs explicit weights
    const dim3 grid((unsigned)((c.W + WAVE - 1) / WAVE));
    const size_t sh = lds_bytes(prec, c.ndim, acc);
    if (prec == MCI_PREC_FP64) {
        if (acc) { hipLaunchKernelGGL(k_fast64<true>, grid, dim3(WAVE), sh, s, A); }
        else { hipLaunchKernelGGL(k_fast64<false>, grid, dim3(WAVE), sh, s, A); }
    }
    else {
        if (acc) { hipLaunchKernelGGL(k_fast32<true>, grid, dim3(WAVE), sh, s, A); }
        else { hipLaunchKernelGGL(k_fast32<false>, grid, dim3(WAVE), sh, s, A); }
    }
    return hipGetLastError() == hipSuccess ? MCI_OK : MCI_E_DEVICE;
}

void fast_release(FastState &fs) { (void)fs; }
} // namespace mcid
