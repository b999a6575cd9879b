// mci_device.hpp — device-side building blocks of the Metropolis chain on gfx950.
//
// One walker per lane.  Walker state lives in HBM as structure-of-arrays with the
// walker index fastest ([dim][walker]), so every per-dimension access of a wave is
// one coalesced 512-byte transaction.  All FP64 arithmetic here is written in the
// reference's expression order and compiled with -ffp-contract=off, so a walker fed
// the reference's random stream reproduces the reference chain bit for bit.
//
// Reference map: doStepMRT2 src/MCIntegrator.cpp:326-360; trial moves
// include/mci/SRRDAllMove.hpp:67-80, SRRDVecMove.hpp:75-96; acceptance
// SamplingFunctionInterface.hpp:49-57; accumulation AccumulatorInterface.cpp:99-114.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mci_amd.h"
#include "../../include/mci_amd/device_functor.hpp"
#include "mci_f32math.hpp"

namespace mcid
{
using mcif::xor3;
constexpr int MAX_PDF = 4;
constexpr int MAX_OBS = 16;
constexpr int MAX_TYPES = 16;
constexpr int MT_N = 312;
constexpr int MT_M = 156;

// accumulator storage types (Factories.hpp:29-43: blocksize 0 / 1 / >1)
enum { ACC_SIMPLE = 0, ACC_FULL = 1, ACC_BLOCK = 2 };

struct PdfDev
{
    int kind, nproto, off, K;
    int uid;           // MCI_PDF_DEVICE: the functor's index in the module built with it
    const double *a;   // per-dimension weights (GAUSS_DIAG, EXP_DIAG), the functor's parameters (DEVICE), or nullptr (= 1)
    const double *gc;  // GMM_ISO: log w_k - ndim log s_k
    const double *gb;  // GMM_ISO: 1/(2 s_k^2)
    const double *gmu; // GMM_ISO: means [K][ndim]
};

struct ObsDev
{
    int kind, nobs, voff, nskip, acc, blocksize;
    int uid;          // MCI_OBS_DEVICE: the functor's index in the module built with it
    const double *par; // MCI_OBS_DEVICE: its parameters
    double p0;        // first parameter (divisor, constant or index)
    int64_t nstore;   // rows of data: 1 (simple), NAccu (full), nblocks (block)
    double *data;     // [nstore][nobs][W]
};

// Everything a kernel needs to advance W walkers; passed by value.
struct Chain
{
    int ndim;
    int64_t W;
    int rng;
    uint32_t key0, key1;        // Philox key = seed
    uint64_t *mt;               // [MT_N][W] mt19937_64 state (reference streams)
    int32_t *mtp;               // [W] next state index
    uint64_t *ctr;              // [W] Philox step counter
    double *xold, *xnew;        // [ndim][W]
    double *pold, *pnew;        // [nproto][W]
    int nproto;
    double *steps;              // [ntypes][W]
    int ntypes;
    int typeEnds[MAX_TYPES];
    int movetype, veclen, nvecs;
    int ibits;                  // index bits in the native shared word (shared_index_bits), -1: none
    int srrd;                   // the move's distribution (MCI_SRRD_*)
    double *nsv;                // [2][W] saved second normals of the move's distribution
    int32_t *nsf;               // [W] bit s: nsv row s available
    int ms_nsteps;              // MultiStepMove around the move above: inner steps (0: none)
    int ms_on;
    int ms_npdf, ms_nproto;     // its own sampling functions and their proto values
    PdfDev ms_pdf[MAX_PDF];
    double *ms_pold, *ms_pnew;  // [ms_nproto][W]
    // the caller's own trial move (movetype MCI_MOVE_DEVICE): its parameters and proto values
    const double *mv_par;
    int mv_nproto;
    double *mv_pold, *mv_pnew;  // [mv_nproto][W]
    double *ms_x0;              // [ndim][W] the position the move starts from
    int lm_on, lm_K;            // the independence half of a mixed move (0: off), mixture components
    double lm_pmix;             // its probability per step
    const double *lm_mu, *lm_isig, *lm_sig; // [K][ndim] means, 1 / sigma, sigma
    const double *lm_lc, *lm_cw;             // [K] log pi_k + sum_j log(1 / sigma_kj) - ndim log(2 pi) / 2, cumulative weights
    double *rec_obs;            // [nrec][rec_nobsdim][W] observable values every rec_fobs steps (file output) or null
    double *rec_x;              // [nrec][ndim][W] positions every rec_fx steps or null
    int64_t rec_fobs, rec_fx;
    int rec_nobsdim;
    double *cb_buf;             // callback record of walker 0: [cb_cap][ndim + 2] (position, accepted, rejected) or null
    int64_t *cb_n;              // [1] calls so far (also past cb_cap)
    int64_t cb_cap;
    int periodic;
    const double *lb, *ub;      // [ndim]
    int npdf;
    PdfDev pdf[MAX_PDF];
    int nobs;
    ObsDev obs[MAX_OBS];
    double *obsval;             // [nobsdim][W] last observable values
    int64_t *acc, *rej;         // [W] counters of the last sample() call
};

__device__ __forceinline__ double &at(double *p, int64_t W, int i, int64_t w) { return p[(int64_t)i * W + w]; }

// --------------------------------------------------------------------------
// random streams
// --------------------------------------------------------------------------
struct Lane
{
    int64_t w;
    int32_t mtp;       // MT: next index
    uint64_t t;        // Philox: step counter
    uint32_t k, kx;    // Philox: word index within the step, extra-word index
    int nchanged;      // changed coordinates [cfirst, cfirst + nchanged) (== ndim: all)
    int cfirst;
    bool accepted;
    int64_t acc, rej;
    double sv[2];      // saved normals of the move's distribution (nsv)
    int sf;            // their availability bits (nsf)
};

// mt19937_64 with the twist done lazily, one element per output: element k of a round
// uses x[k+1] (not yet twisted this round) and x[(k+156) % 312] (twisted iff k >= 156),
// exactly the values the batch twist of libstdc++ (bits/random.tcc) uses, so the
// output stream is identical.  State index 0 after seeding = batch index 312.
// the new value of element k from x[k], x[k+1] and x[(k+156) % 312] as the twist reads them
__device__ __forceinline__ uint64_t mt_twist(uint64_t xk, uint64_t xk1, uint64_t xkm)
{
    const uint64_t y = (xk & 0xffffffff80000000ULL) | (xk1 & 0x7fffffffULL);
    return xkm ^ (y >> 1) ^ ((y & 1ULL) ? 0xb5026f5aa96619e9ULL : 0ULL);
}

__device__ __forceinline__ uint64_t mt_temper(uint64_t z)
{
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71d67fffeda60000ULL;
    z ^= (z << 37) & 0xfff7eee000000000ULL;
    z ^= (z >> 43);
    return z;
}

__device__ __forceinline__ uint64_t mt_next(const Chain &c, Lane &L)
{
    const int64_t W = c.W;
    const int k = L.mtp;
    const int k1 = (k + 1 == MT_N) ? 0 : k + 1;
    const int km = (k < MT_N - MT_M) ? k + MT_M : k + MT_M - MT_N;
    const uint64_t xk = c.mt[(int64_t)k * W + L.w];
    const uint64_t xk1 = c.mt[(int64_t)k1 * W + L.w];
    const uint64_t xkm = c.mt[(int64_t)km * W + L.w];
    const uint64_t nv = mt_twist(xk, xk1, xkm);
    c.mt[(int64_t)k * W + L.w] = nv;
    L.mtp = k1;
    return mt_temper(nv);
}

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                              uint32_t out[4])
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
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The native stream (key = seed).  Steps go in pairs: block j of step pair t / 2 is Philox(counter =
// (j, walker, t / 2)), and main word k of step t is word 2 (t mod 2) + (k mod 2) of block k / 2, so
// one block holds words 0 and 1 of two consecutive steps.  Main word 0 is the step's shared word
// (draw_accept); the move's draws start at word 1.  Side streams keyed by the step itself: Lemire
// rejections (block 0x20000000 + kx / 4, word kx mod 4) and the acceptance uniform's low word
// (block 0x30000000, word 0).
__device__ __forceinline__ uint32_t ph_block_word(const Chain &c, const Lane &L, uint32_t blk, uint64_t ct, uint32_t word)
{
    uint32_t o[4];
    philox4x32_10(blk, (uint32_t)L.w, (uint32_t)ct, (uint32_t)(ct >> 32), c.key0, c.key1, o);
    return o[word];
}

__device__ __forceinline__ uint32_t ph_word(const Chain &c, const Lane &L, uint32_t k)
{
    return ph_block_word(c, L, k >> 1, L.t >> 1, 2u * (uint32_t)(L.t & 1u) + (k & 1u));
}

__device__ __forceinline__ void step_begin(Lane &L) { L.k = 1; L.kx = 0; }
__device__ __forceinline__ void step_end(const Chain &c, Lane &L) { if (c.rng == MCI_RNG_PHILOX) { ++L.t; } }

// generate_canonical<double,53>(mt19937_64): double(word)/2^64, clamped below 1
__device__ __forceinline__ double canonical53(uint64_t w)
{
    double r = (double)w / 18446744073709551616.0;
    if (r >= 1.0) { r = 0x1.fffffffffffffp-1; }
    return r;
}

// symmetric uniform on (-1,1) (TrialMoveInterface.hpp:120-124): (u*(b-a)) + a
__device__ __forceinline__ double draw_sym(const Chain &c, Lane &L)
{
    if (c.rng == MCI_RNG_MT19937_64) {
        const double u = canonical53(mt_next(c, L));
        return (u * 2.0) + (-1.0);
    }
    const uint32_t v = ph_word(c, L, L.k++);
    return (double)(2 * (int64_t)v + 1 - 4294967296LL) * 0x1p-32;
}

// uniform on [0,1) (MCIntegrator.cpp:343 acceptance draw; :365 random moves)
__device__ __forceinline__ double draw_unit(const Chain &c, Lane &L)
{
    if (c.rng == MCI_RNG_MT19937_64) {
        const double u = canonical53(mt_next(c, L));
        return (u * 1.0) + 0.0;
    }
    const uint64_t a = ph_word(c, L, L.k++);
    const uint64_t b = ph_word(c, L, L.k++);
    return (double)(((a << 32) | b) >> 11) * 0x1p-53;
}

// the acceptance uniform of a Metropolis step (MCIntegrator.cpp:343).  Native: with a single-vector
// move over 2^m vectors (c.ibits = m <= 11) the shared word's top m bits are the move's index and
// u = ((shared word : side-stream word) << m) >> 11, times 2^-53 (the m zero bits shifted in fall
// below the 53 kept); otherwise m = 0.
__device__ __forceinline__ double draw_accept(const Chain &c, Lane &L)
{
    if (c.rng == MCI_RNG_MT19937_64) { return draw_unit(c, L); }
    const uint32_t m = c.ibits > 0 ? (uint32_t)c.ibits : 0u;
    const uint64_t hi = ph_word(c, L, 0), lo = ph_block_word(c, L, 0x30000000u, L.t, 0);
    return (double)((((hi << 32) | lo) << m) >> 11) * 0x1p-53;
}

// uniform integer on [0, n) (SRRDVecMove.hpp:78): Lemire's method as in libstdc++
// (64-bit engine, 128-bit product) or on 32-bit Philox words
__device__ __forceinline__ int draw_index(const Chain &c, Lane &L, int n)
{
    if (c.rng == MCI_RNG_MT19937_64) {
        const uint64_t range = (uint64_t)n;
        uint64_t wd = mt_next(c, L);
        uint64_t low = wd * range;
        uint64_t high = __umul64hi(wd, range);
        if (low < range) {
            const uint64_t thr = (0 - range) % range;
            while (low < thr) {
                wd = mt_next(c, L);
                low = wd * range;
                high = __umul64hi(wd, range);
            }
        }
        return (int)high;
    }
    const uint32_t range = (uint32_t)n;
    uint64_t prod = (uint64_t)ph_word(c, L, L.k++) * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        const uint32_t thr = (0u - range) % range;
        while (low < thr) {
            const uint32_t kx = L.kx++;
            prod = (uint64_t)ph_block_word(c, L, 0x20000000u + (kx >> 2), L.t, kx & 3u) * range;
            low = (uint32_t)prod;
        }
    }
    return (int)(prod >> 32);
}

// --------------------------------------------------------------------------
// the move distributions (SRRDType, Factories.hpp:118-131) with createSymRRD's default parameters
// (TrialMoveInterface.hpp:107-183), drawn as libstdc++ (GCC 11) draws them: every uniform is one
// canonical double (draw_unit), the normal is Marsaglia's polar method with the pair's second
// value kept per walker (two generators at most: slots 0 and 1), the gamma Marsaglia-Tsang
// boosted by u^(1/alpha) for alpha < 1, the one-sided ones signed by one fair bernoulli draw after
// the value (SymmetrizedPRRD, TrialMoveInterface.hpp:80-98).  With the reference stream the
// draws are the reference's up to the last place of the device's log, tan, pow and exp.
// --------------------------------------------------------------------------
__device__ inline double nd_draw(const Chain &c, Lane &L, int s)
{
    double ret;
    if (L.sf & (1 << s)) {
        L.sf &= ~(1 << s);
        ret = L.sv[s];
    }
    else {
        double x, y, r2;
        do {
            x = 2.0 * draw_unit(c, L) - 1.0;
            y = 2.0 * draw_unit(c, L) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2 * log(r2) / r2);
        L.sv[s] = x * mult;
        L.sf |= 1 << s;
        ret = y * mult;
    }
    return ret * 1.0 + 0.0;
}

__device__ inline double gamma_draw(const Chain &c, Lane &L, int s, double alpha, double beta)
{
    const double malpha = alpha < 1.0 ? alpha + 1.0 : alpha;
    const double a1 = malpha - 1.0 / 3.0;
    const double a2 = 1.0 / sqrt(9.0 * a1);
    double u, v, n;
    do {
        do {
            n = nd_draw(c, L, s);
            v = 1.0 + a2 * n;
        } while (v <= 0.0);
        v = v * v * v;
        u = draw_unit(c, L);
    } while (u > 1.0 - 0.0331 * n * n * n * n && (log(u) > (0.5 * n * n + a1 * (1.0 - v + log(v)))));
    if (alpha == malpha) { return a1 * v * beta; }
    do { u = draw_unit(c, L); } while (u == 0.0);
    return pow(u, 1.0 / alpha) * a1 * v * beta;
}

__device__ inline double srrd_draw(const Chain &c, Lane &L)
{
    double val;
    switch (c.srrd) {
    case MCI_SRRD_UNIFORM: return draw_sym(c, L);
    case MCI_SRRD_GAUSSIAN: return nd_draw(c, L, 0);
    case MCI_SRRD_STUDENT: {
        const double z = nd_draw(c, L, 0);
        return z * sqrt(1.0 / gamma_draw(c, L, 1, 0.5, 2.0));
    }
    case MCI_SRRD_CAUCHY: {
        double u;
        do { u = draw_unit(c, L); } while (u == 0.5);
        return 0.0 + 1.0 * tan(3.141592653589793 * u);
    }
    case MCI_SRRD_EXPONENTIAL: val = -log(1.0 - draw_unit(c, L)) / 1.0; break;
    case MCI_SRRD_GAMMA: val = gamma_draw(c, L, 0, 1.0, 1.0); break;
    case MCI_SRRD_WEIBULL: val = 1.0 * pow(-log(1.0 - draw_unit(c, L)), 1.0 / 1.0); break;
    case MCI_SRRD_LOGNORMAL: val = exp(1.0 * nd_draw(c, L, 0) + 0.0); break;
    case MCI_SRRD_CHISQ: val = 2 * gamma_draw(c, L, 0, 0.5, 1.0); break;
    case MCI_SRRD_FISHER: {
        const double gx = gamma_draw(c, L, 0, 0.5, 1.0);
        const double gy = gamma_draw(c, L, 1, 0.5, 1.0);
        val = (gx * 1.0) / (gy * 1.0);
        break;
    }
    default: return __builtin_nan("");
    }
    return ((draw_unit(c, L) - 0.0) < 0.5 * (1.0 - 0.0)) ? val : -val;
}

// --------------------------------------------------------------------------
// sampling functions (device functors; arithmetic of include/mci_amd.h)
// --------------------------------------------------------------------------
__device__ __forceinline__ double pdf_a(const PdfDev &p, int i) { return p.a ? p.a[i] : 1.0; }

__device__ inline double gmm_logp(const Chain &c, const PdfDev &p, const double *pv, int64_t W, int64_t w)
{
    double m = -__builtin_huge_val();
    for (int k = 0; k < p.K; ++k) {
        const double t = p.gc[k] - p.gb[k] * pv[(int64_t)k * W + w];
        if (t > m) { m = t; }
    }
    double s = 0.0;
    for (int k = 0; k < p.K; ++k) { s = s + exp((p.gc[k] - p.gb[k] * pv[(int64_t)k * W + w]) - m); }
    return m + log(s);
}

// protoFunction(x) -> proto values (written to pv + p.off rows)
__device__ inline void pdf_proto(const Chain &c, const PdfDev &p, const double *x, double *pv, int64_t w)
{
    const int64_t W = c.W;
    double *q = pv + (int64_t)p.off * W;
    switch (p.kind) {
    case MCI_PDF_GAUSS_ISO: {
        double s = x[w] * x[w];
        for (int i = 1; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; s = s + v * v; }
        q[w] = s;
        break;
    }
    case MCI_PDF_GAUSS_DIAG:
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; q[(int64_t)i * W + w] = (v * v) * pdf_a(p, i); }
        break;
    case MCI_PDF_EXP_DIAG:
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; q[(int64_t)i * W + w] = fabs(v) * pdf_a(p, i); }
        break;
    case MCI_PDF_GMM_ISO:
        for (int k = 0; k < p.K; ++k) {
            double s = 0.0;
            for (int i = 0; i < c.ndim; ++i) { const double d = x[(int64_t)i * W + w] - p.gmu[k * c.ndim + i]; s = s + d * d; }
            q[(int64_t)k * W + w] = s;
        }
        q[(int64_t)p.K * W + w] = gmm_logp(c, p, q, W, w);
        break;
#ifdef MCI_RTC_USER
    case MCI_PDF_DEVICE: // the caller's protoFunction (built at run time with the functor's source)
        mci_user_pdf_proto(p.uid, c.ndim, MciVec{const_cast<double *>(x) + w, W}, p.a, MciVec{q + w, W});
        break;
#endif
    }
}

__device__ inline double pdf_acceptance(const Chain &c, const PdfDev &p, const double *pold, const double *pnew, int64_t w)
{
    const int64_t W = c.W;
    const double *po = pold + (int64_t)p.off * W;
    const double *pn = pnew + (int64_t)p.off * W;
    switch (p.kind) {
    case MCI_PDF_GAUSS_ISO:
        return exp(-pn[w] + po[w]);
    case MCI_PDF_GAUSS_DIAG:
    case MCI_PDF_EXP_DIAG: {
        double e = 0.0;
        for (int i = 0; i < p.nproto; ++i) { e = e + po[(int64_t)i * W + w]; }
        double s = 0.0;
        for (int i = 0; i < p.nproto; ++i) { s = s + pn[(int64_t)i * W + w]; }
        e = e - s;
        return exp(e);
    }
    case MCI_PDF_GMM_ISO:
        return exp(pn[(int64_t)p.K * W + w] - po[(int64_t)p.K * W + w]);
#ifdef MCI_RTC_USER
    case MCI_PDF_DEVICE: // the caller's acceptanceFunction
        return mci_user_pdf_acceptance(p.uid, p.nproto, MciVec{const_cast<double *>(po) + w, W}, MciVec{const_cast<double *>(pn) + w, W}, p.a);
#endif
    }
    return 0.0;
}

// SamplingFunctionInterface::computeAcceptance (+ the updatedAcceptance of each kind)
__device__ inline double pdf_compute_acceptance(const Chain &c, const PdfDev &p, const Lane &L, double *pold, double *pnew)
{
    const int64_t W = c.W, w = L.w;
#ifdef MCI_RTC_USER
    // the caller's updatedAcceptance when it defines one (else its protoFunction and
    // acceptanceFunction, the interface's default), for the changed coordinates
    if (p.kind == MCI_PDF_DEVICE && L.nchanged < c.ndim) {
        double *po = pold + (int64_t)p.off * W + w, *pn = pnew + (int64_t)p.off * W + w;
        return mci_user_pdf_updated(p.uid, c.ndim, p.nproto, MciVec{c.xnew + w, W}, MciChanged{L.cfirst, L.nchanged}, MciVec{po, W},
                                    MciVec{pn, W}, p.a);
    }
#endif
    if (L.nchanged < c.ndim) {
        double *po = pold + (int64_t)p.off * W;
        double *pn = pnew + (int64_t)p.off * W;
        if (p.kind == MCI_PDF_GAUSS_DIAG || p.kind == MCI_PDF_EXP_DIAG) {
            double e = 0.0;
            for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) {
                const double v = c.xnew[(int64_t)j * W + w];
                const double f = (p.kind == MCI_PDF_GAUSS_DIAG) ? (v * v) * pdf_a(p, j) : fabs(v) * pdf_a(p, j);
                pn[(int64_t)j * W + w] = f;
                e = e + (f - po[(int64_t)j * W + w]);
            }
            return exp(-e);
        }
        if (p.kind == MCI_PDF_GMM_ISO) {
            for (int k = 0; k < p.K; ++k) {
                double dr = 0.0;
                for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) {
                    const double mu = p.gmu[k * c.ndim + j];
                    const double dn = c.xnew[(int64_t)j * W + w] - mu, dd = c.xold[(int64_t)j * W + w] - mu;
                    dr = dr + (dn * dn - dd * dd);
                }
                pn[(int64_t)k * W + w] = po[(int64_t)k * W + w] + dr;
            }
            pn[(int64_t)p.K * W + w] = gmm_logp(c, p, pn, W, w);
            return exp(pn[(int64_t)p.K * W + w] - po[(int64_t)p.K * W + w]);
        }
    }
    pdf_proto(c, p, c.xnew, pnew, w);
    return pdf_acceptance(c, p, pold, pnew, w);
}

// samplingFunction(proto values) of each kind (exp of minus the proto sum, or of the mixture's log p)
__device__ inline double pdf_value(const Chain &c, const PdfDev &p, const double *pv, int64_t w)
{
    const int64_t W = c.W;
    const double *q = pv + (int64_t)p.off * W;
    switch (p.kind) {
    case MCI_PDF_GAUSS_ISO: return exp(-q[w]);
    case MCI_PDF_GAUSS_DIAG:
    case MCI_PDF_EXP_DIAG: {
        double s = 0.0;
        for (int i = 0; i < p.nproto; ++i) { s = s + q[(int64_t)i * W + w]; }
        return exp(-s);
    }
    case MCI_PDF_GMM_ISO: return exp(q[(int64_t)p.K * W + w]);
#ifdef MCI_RTC_USER
    case MCI_PDF_DEVICE: return mci_user_pdf_sampling(p.uid, p.nproto, MciVec{const_cast<double *>(q) + w, W}, p.a);
#endif
    }
    return 0.0;
}

// --------------------------------------------------------------------------
// domain (OrthoPeriodicDomain.cpp:38-61; UnboundDomain is a no-op)
// --------------------------------------------------------------------------
// The reference wraps by repeated addition/subtraction of the box length; that is kept
// (bit-identical) for coordinates within 64 box lengths.  Farther coordinates are first
// reduced with fmod so every loop is bounded; non-finite coordinates are left unchanged.
__device__ __forceinline__ void wrap(const Chain &c, int i, double &x)
{
    const double lo = c.lb[i], hi = c.ub[i], len = hi - lo;
    if (!isfinite(x)) { return; }
    if (x < lo - 64.0 * len || x > hi + 64.0 * len) {
        double r = fmod(x - lo, len);
        if (r < 0.0) { r = r + len; }
        x = lo + r;
    }
    int it = 0;
    while (x < lo && it < 130) { x = x + len; ++it; }
    while (x > hi && it < 260) { x = x - len; ++it; }
}

// --------------------------------------------------------------------------
// the Metropolis step
// --------------------------------------------------------------------------
// SRRDAllMove / SRRDVecMove trial moves (uniform SRRD)
__device__ inline double srrd_move(const Chain &c, Lane &L, bool outer)
{
    const int64_t W = c.W, w = L.w;
    if (c.movetype == MCI_MOVE_ALL) {
        int xidx = 0;
        for (int t = 0; t < c.ntypes; ++t) {
            const double st = c.steps[(int64_t)t * W + w];
            while (xidx < c.typeEnds[t]) {
                double &xn = c.xnew[(int64_t)xidx * W + w];
                xn = xn + st * srrd_draw(c, L);
                ++xidx;
            }
        }
        L.nchanged = c.ndim;
        L.cfirst = 0;
        return 1.;
    }
    // the chain's own move takes a power-of-two index from the shared word's top bits
    const int ib = outer ? c.ibits : -1;
    const int vidx = ib > 0 ? (int)(ph_word(c, L, 0) >> (32 - ib)) : ib == 0 ? 0 : draw_index(c, L, c.nvecs);
    const int xidx = vidx * c.veclen;
    int t = 0;
    while (t < c.ntypes) {
        if (xidx < c.typeEnds[t]) { break; }
        ++t;
    }
    const double st = c.steps[(int64_t)t * W + w];
    for (int i = 0; i < c.veclen; ++i) {
        double &xn = c.xnew[(int64_t)(xidx + i) * W + w];
        xn = xn + st * srrd_draw(c, L);
    }
    L.nchanged = c.veclen;
    L.cfirst = xidx;
    return 1.;
}

// the state the reference's callback sees (MCIntegrator.cpp:265, 346-347, 373): walker 0's committed
// position and its counters, recorded for the host to replay the calls in order
__device__ inline void cb_record(const Chain &c, const Lane &L)
{
    if (c.cb_buf == nullptr || L.w != 0) { return; }
    const int64_t n = c.cb_n[0];
    if (n < c.cb_cap) {
        double *r = c.cb_buf + n * (c.ndim + 2);
        for (int i = 0; i < c.ndim; ++i) { r[i] = c.xold[(int64_t)i * c.W]; }
        r[c.ndim] = (double)L.acc;
        r[c.ndim + 1] = (double)L.rej;
    }
    c.cb_n[0] = n + 1;
}

// log q(y) of the independence mixture at column w of y [ndim][W]: the log-sum-exp over components of
// lc_k - |(y - mu_k) / sigma_k|^2 / 2
__device__ inline double lm_logq(const Chain &c, const double *y, int64_t w)
{
    const int64_t W = c.W;
    double e[8], mx = -__builtin_huge_val();
    for (int k = 0; k < c.lm_K; ++k) {
        double s = 0.0;
        for (int j = 0; j < c.ndim; ++j) {
            const double z = (y[(int64_t)j * W + w] - c.lm_mu[k * c.ndim + j]) * c.lm_isig[k * c.ndim + j];
            s = s + z * z;
        }
        e[k] = c.lm_lc[k] - 0.5 * s;
        mx = e[k] > mx ? e[k] : mx;
    }
    double s = 0.0;
    for (int k = 0; k < c.lm_K; ++k) { s = s + exp(e[k] - mx); }
    return mx + log(s);
}

// The independence move of a mixed trial move: a component by its cumulative weight (one uniform),
// standard normals by Marsaglia's polar method (pairs, nothing kept between steps), x' = mu_k +
// sigma_k z for every coordinate; the move's acceptance factor q(x) / q(x') (the reference's
// trial-move hook, TrialMoveInterface.hpp:63-69, MCIntegrator.cpp:343), so that the chain keeps the
// target exact whatever the mixture is.
__device__ inline double lm_move(const Chain &c, Lane &L)
{
    const int64_t W = c.W, w = L.w;
    const double uc = draw_unit(c, L);
    int k = c.lm_K - 1;
    for (int kk = 0; kk < c.lm_K - 1; ++kk) {
        if (uc < c.lm_cw[kk]) { k = kk; break; }
    }
    for (int j = 0; j < c.ndim; j += 2) {
        double x, y, r2;
        do {
            x = 2.0 * draw_unit(c, L) - 1.0;
            y = 2.0 * draw_unit(c, L) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2 * log(r2) / r2);
        c.xnew[(int64_t)j * W + w] = c.lm_mu[k * c.ndim + j] + c.lm_sig[k * c.ndim + j] * (y * mult);
        if (j + 1 < c.ndim) { c.xnew[(int64_t)(j + 1) * W + w] = c.lm_mu[k * c.ndim + j + 1] + c.lm_sig[k * c.ndim + j + 1] * (x * mult); }
    }
    L.nchanged = c.ndim;
    L.cfirst = 0;
    return exp(lm_logq(c, c.xold, w) - lm_logq(c, c.xnew, w));
}

// MultiStepMove::trialMove (MultiStepMove.cpp:6-47): ms_nsteps Metropolis steps of the contained
// move under the move's own sampling functions (none: every step accepted), from xold; xnew ends at
// the last position, xold is restored, every coordinate counts as changed, and the factor
// oldPDF / newPDF of the own sampling functions is the move's acceptance factor
__device__ inline double trial_move(const Chain &c, Lane &L)
{
    // a mixed move: first the uniform that picks the independence move with probability lm_pmix
    if (c.lm_on && draw_unit(c, L) < c.lm_pmix) { return lm_move(c, L); }
#ifdef MCI_RTC_USER
    if (c.movetype == MCI_MOVE_DEVICE) {
        // the caller's trialMove: it moves xnew (which holds xold), names the changed coordinates
        // (one ascending range) and returns its acceptance factor
        const int64_t W = c.W, w = L.w;
        MciRng rng{&c, &L};
        MciMoved mv{0, c.ndim};
        const double f = mci_user_move(c.ndim, MciVec{c.xnew + w, W}, mv, MciVec{c.mv_pold + w, W}, MciVec{c.mv_pnew + w, W},
                                       MciVec{c.steps + w, W}, rng, c.mv_par);
        L.cfirst = mv.first;
        L.nchanged = mv.n;
        return f;
    }
#endif
    if (!c.ms_on) { return srrd_move(c, L, true); }
    const int64_t W = c.W, w = L.w;
    for (int i = 0; i < c.ndim; ++i) { c.ms_x0[(int64_t)i * W + w] = c.xold[(int64_t)i * W + w]; }
    for (int p = 0; p < c.ms_npdf; ++p) { pdf_proto(c, c.ms_pdf[p], c.xold, c.ms_pold, w); }
    for (int i = 0; i < c.ms_nproto; ++i) { c.ms_pnew[(int64_t)i * W + w] = c.ms_pold[(int64_t)i * W + w]; }
    double oldPDF = 1.;
    for (int p = 0; p < c.ms_npdf; ++p) { oldPDF *= pdf_value(c, c.ms_pdf[p], c.ms_pold, w); }
    for (int s = 0; s < c.ms_nsteps; ++s) {
        const double moveAcc = srrd_move(c, L, false);
        double pdfAcc = 1.;
        for (int p = 0; p < c.ms_npdf; ++p) { pdfAcc *= pdf_compute_acceptance(c, c.ms_pdf[p], L, c.ms_pold, c.ms_pnew); }
        const bool acc = (draw_unit(c, L) <= pdfAcc * moveAcc);
        double *pf = acc ? c.ms_pnew : c.ms_pold, *pt = acc ? c.ms_pold : c.ms_pnew;
        double *xf = acc ? c.xnew : c.xold, *xt = acc ? c.xold : c.xnew;
        for (int i = 0; i < c.ms_nproto; ++i) { pt[(int64_t)i * W + w] = pf[(int64_t)i * W + w]; }
        for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) { xt[(int64_t)j * W + w] = xf[(int64_t)j * W + w]; }
    }
    double newPDF = 1.;
    for (int p = 0; p < c.ms_npdf; ++p) { newPDF *= pdf_value(c, c.ms_pdf[p], c.ms_pold, w); }
    for (int i = 0; i < c.ndim; ++i) { c.xold[(int64_t)i * W + w] = c.ms_x0[(int64_t)i * W + w]; }
    L.nchanged = c.ndim;
    L.cfirst = 0;
    return oldPDF / newPDF;
}

// MCI::initializeSampling (MCIntegrator.cpp:252-271)
__device__ inline void init_sampling(const Chain &c, Lane &L)
{
    const int64_t W = c.W, w = L.w;
    L.acc = 0;
    L.rej = 0;
    for (int i = 0; i < c.ndim; ++i) { c.xnew[(int64_t)i * W + w] = c.xold[(int64_t)i * W + w]; }
    L.nchanged = c.ndim;
    L.cfirst = 0;
    L.accepted = true;
    for (int p = 0; p < c.npdf; ++p) { pdf_proto(c, c.pdf[p], c.xold, c.pold, w); }
    for (int i = 0; i < c.nproto; ++i) { c.pnew[(int64_t)i * W + w] = c.pold[(int64_t)i * W + w]; }
#ifdef MCI_RTC_USER
    if (c.movetype == MCI_MOVE_DEVICE && c.mv_nproto > 0) { // the trial move's initializeProtoValues
        mci_user_move_proto(c.ndim, MciVec{c.xold + w, W}, c.mv_par, MciVec{c.mv_pold + w, W});
        for (int i = 0; i < c.mv_nproto; ++i) { c.mv_pnew[(int64_t)i * W + w] = c.mv_pold[(int64_t)i * W + w]; }
    }
#endif
    cb_record(c, L);
}

// MCI::doStepMRT2 (MCIntegrator.cpp:326-360)
__device__ inline void step_mrt2(const Chain &c, Lane &L)
{
    const int64_t W = c.W, w = L.w;
    step_begin(L);
    const double moveAcc = trial_move(c, L);
    if (c.periodic) {
        for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) { wrap(c, j, c.xnew[(int64_t)j * W + w]); }
    }
    double pdfAcc = 1.;
    for (int p = 0; p < c.npdf; ++p) { pdfAcc = pdfAcc * pdf_compute_acceptance(c, c.pdf[p], L, c.pold, c.pnew); }
    const double u = draw_accept(c, L);
    L.accepted = (u <= pdfAcc * moveAcc);
    if (L.accepted) { ++L.acc; } else { ++L.rej; }
    cb_record(c, L);
    step_end(c, L);
    // commit or roll back the coordinates and proto values that can differ
    if (L.accepted) {
        for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) { c.xold[(int64_t)j * W + w] = c.xnew[(int64_t)j * W + w]; }
        for (int i = 0; i < c.nproto; ++i) { c.pold[(int64_t)i * W + w] = c.pnew[(int64_t)i * W + w]; }
    }
    else {
        for (int j = L.cfirst; j < L.cfirst + L.nchanged; ++j) { c.xnew[(int64_t)j * W + w] = c.xold[(int64_t)j * W + w]; }
        for (int i = 0; i < c.nproto; ++i) { c.pnew[(int64_t)i * W + w] = c.pold[(int64_t)i * W + w]; }
    }
#ifdef MCI_RTC_USER
    if (c.movetype == MCI_MOVE_DEVICE) { // the trial move's newToOld / oldToNew
        double *mf = L.accepted ? c.mv_pnew : c.mv_pold, *mt = L.accepted ? c.mv_pold : c.mv_pnew;
        for (int i = 0; i < c.mv_nproto; ++i) { mt[(int64_t)i * W + w] = mf[(int64_t)i * W + w]; }
    }
#endif
}

// MCI::doStepRandom (MCIntegrator.cpp:362-376): no sampling function
__device__ inline void step_random(const Chain &c, Lane &L)
{
    const int64_t W = c.W, w = L.w;
    step_begin(L);
    for (int i = 0; i < c.ndim; ++i) { c.xnew[(int64_t)i * W + w] = draw_unit(c, L); }
    step_end(c, L);
    for (int i = 0; i < c.ndim; ++i) {
        double &x = c.xnew[(int64_t)i * W + w];
        if (c.periodic) { x = c.lb[i] + x * (c.ub[i] - c.lb[i]); }
        else { x = -(double)3.402823466e+38f + x * ((double)3.402823466e+38f + (double)3.402823466e+38f); }
    }
    L.nchanged = c.ndim;
    L.cfirst = 0;
    L.accepted = true;
    ++L.acc;
    cb_record(c, L);
    for (int i = 0; i < c.ndim; ++i) { c.xold[(int64_t)i * W + w] = c.xnew[(int64_t)i * W + w]; }
}

// --------------------------------------------------------------------------
// observables (ObservableFunctionInterface::observableFunction of each kind)
// --------------------------------------------------------------------------
__device__ inline void obs_eval(const Chain &c, const ObsDev &o, int64_t w)
{
    const int64_t W = c.W;
    const double *x = c.xold;
    double *out = c.obsval + (int64_t)o.voff * W;
    switch (o.kind) {
    case MCI_OBS_X2:
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; out[(int64_t)i * W + w] = v * v; }
        break;
    case MCI_OBS_X2_SUM: {
        double s = 0.0;
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; s = s + v * v; }
        out[w] = s / o.p0;
        break;
    }
    case MCI_OBS_X:
        for (int i = 0; i < c.ndim; ++i) { out[(int64_t)i * W + w] = x[(int64_t)i * W + w]; }
        break;
    case MCI_OBS_X_SUM: {
        double s = 0.0;
        for (int i = 0; i < c.ndim; ++i) { s = s + x[(int64_t)i * W + w]; }
        out[w] = s;
        break;
    }
    case MCI_OBS_CONST:
        out[w] = o.p0;
        break;
    case MCI_OBS_EXP_X2: {
        double s = 0.0;
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; s = s + v * v; }
        out[w] = exp(-(s / o.p0));
        break;
    }
    case MCI_OBS_INDICATOR:
        out[w] = (x[(int64_t)(int)o.p0 * W + w] > 0.0) ? 1.0 : 0.0;
        break;
    case MCI_OBS_GAUSS_X2: {
        double s = 0.0;
        for (int i = 0; i < c.ndim; ++i) { const double v = x[(int64_t)i * W + w]; s = s + v * v; }
        out[w] = exp(-s) * s * o.p0;
        break;
    }
    case MCI_OBS_PDF_LOG: { // the committed proto values of sampling function p0
        const PdfDev &p = c.pdf[(int)o.p0];
        const double *pv = c.pold + (int64_t)p.off * W + w;
        if (p.kind == MCI_PDF_GMM_ISO) { out[w] = pv[(int64_t)p.K * W]; }
#ifdef MCI_RTC_USER
        else if (p.kind == MCI_PDF_DEVICE) { out[w] = log(mci_user_pdf_sampling(p.uid, p.nproto, MciVec{const_cast<double *>(pv), W}, p.a)); }
#endif
        else {
            double s = 0.0;
            for (int j = 0; j < p.nproto; ++j) { s = s + pv[(int64_t)j * W]; }
            out[w] = -s;
        }
        break;
    }
    case MCI_OBS_PDF_PROTO: {
        const PdfDev &p = c.pdf[(int)o.p0];
        for (int j = 0; j < o.nobs; ++j) { out[(int64_t)j * W + w] = c.pold[(int64_t)(p.off + j) * W + w]; }
        break;
    }
    case MCI_OBS_DEP_SQUARE: { // the last values of an earlier observable (evaluated before this one)
        const double *v = c.obsval + (int64_t)c.obs[(int)o.p0].voff * W;
        for (int j = 0; j < o.nobs; ++j) { const double y = v[(int64_t)j * W + w]; out[(int64_t)j * W + w] = y * y; }
        break;
    }
#ifdef MCI_RTC_USER
    case MCI_OBS_DEVICE: // the caller's observableFunction (built at run time with the functor's source)
        mci_user_observable(o.uid, c.ndim, MciVec{const_cast<double *>(x) + w, W}, o.par, MciVec{out + w, W});
        break;
#endif
    }
}

// One accumulation step of every observable (AccumulatorInterface::accumulate and the
// storage of Simple/Block/FullAccumulator).  Observable values are a function of the
// current position, so they are re-evaluated exactly when the position changed since the
// last evaluation (the reference's _nchanged bookkeeping) — the stored numbers are equal.
__device__ inline void accumulate(const Chain &c, const Lane &L, int64_t s, uint32_t &dirty)
{
    const int64_t W = c.W, w = L.w;
    if (L.accepted) { dirty = 0xffffffffu; }
    for (int i = 0; i < c.nobs; ++i) {
        const ObsDev &o = c.obs[i];
        if (s % o.nskip != 0) { continue; }
        if (dirty & (1u << i)) {
            obs_eval(c, o, w);
            dirty &= ~(1u << i);
        }
        const int64_t a = s / o.nskip;
        const double *v = c.obsval + (int64_t)o.voff * W;
        if (o.acc == ACC_FULL) {
            for (int j = 0; j < o.nobs; ++j) { o.data[(a * o.nobs + j) * W + w] = v[(int64_t)j * W + w]; }
        }
        else {
            const int64_t b = (o.acc == ACC_BLOCK) ? a / o.blocksize : 0;
            for (int j = 0; j < o.nobs; ++j) {
                double &d = o.data[(b * o.nobs + j) * W + w];
                d = d + v[(int64_t)j * W + w];
            }
        }
    }
}

// AccumulatorInterface::finalize of Simple (x 1/NAccu) and Block (x 1/blocksize)
__device__ inline void finalize_accu(const Chain &c, int64_t w, int64_t nsteps)
{
    const int64_t W = c.W;
    for (int i = 0; i < c.nobs; ++i) {
        const ObsDev &o = c.obs[i];
        if (o.acc == ACC_FULL) { continue; }
        double normf;
        if (o.acc == ACC_SIMPLE) { normf = 1. / (double)(1 + (nsteps - 1) / o.nskip); }
        else { normf = 1. / o.blocksize; }
        for (int64_t r = 0; r < o.nstore * o.nobs; ++r) {
            double &d = o.data[r * W + w];
            d = d * normf;
        }
    }
}

__device__ inline void zero_accu(const Chain &c, int64_t w)
{
    const int64_t W = c.W;
    for (int i = 0; i < c.nobs; ++i) {
        const ObsDev &o = c.obs[i];
        if (o.acc == ACC_FULL) { continue; }
        for (int64_t r = 0; r < o.nstore * o.nobs; ++r) { o.data[r * W + w] = 0.0; }
    }
}

__device__ inline void load_lane(const Chain &c, Lane &L, int64_t w)
{
    L.w = w;
    L.mtp = (c.rng == MCI_RNG_MT19937_64) ? c.mtp[w] : 0;
    L.t = (c.rng == MCI_RNG_PHILOX) ? c.ctr[w] : 0;
    L.k = 0;
    L.kx = 0;
    L.nchanged = c.ndim;
    L.cfirst = 0;
    L.accepted = true;
    L.acc = 0;
    L.rej = 0;
    L.sf = 0;
    L.sv[0] = L.sv[1] = 0.0;
    if (c.nsf != nullptr) {
        L.sf = c.nsf[w];
        L.sv[0] = c.nsv[w];
        L.sv[1] = c.nsv[c.W + w];
    }
}

__device__ inline void store_lane(const Chain &c, const Lane &L)
{
    if (c.rng == MCI_RNG_MT19937_64) { c.mtp[L.w] = L.mtp; }
    else { c.ctr[L.w] = L.t; }
    if (c.nsf != nullptr) {
        c.nsf[L.w] = L.sf;
        c.nsv[L.w] = L.sv[0];
        c.nsv[c.W + L.w] = L.sv[1];
    }
    c.acc[L.w] = L.acc;
    c.rej[L.w] = L.rej;
}
} // namespace mcid

#ifdef MCI_RTC_USER
// a uniform on [0, 1) from the chain's stream, as MCI's generator through
// uniform_real_distribution<double>(0, 1) gives it (reference streams) or the native stream's draw
__device__ inline double mci_uniform(MciRng &rng)
{
    return mcid::draw_unit(*static_cast<const mcid::Chain *>(rng.c), *static_cast<mcid::Lane *>(rng.L));
}
#endif
