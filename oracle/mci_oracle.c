/*
 * mci_oracle.c — CPU oracle of MCIntegrator++'s Metropolis sampling path.
 *
 * TEST INFRASTRUCTURE ONLY: used as the checker by tests/, the smoke check and
 * the cpu_baseline leg of bench.py.  Nothing in mcintegratorplusplus_amd/
 * links, loads or calls this file.
 *
 * A plain-C restatement of the reference algorithm.  Each function names the
 * reference file:line it restates (paths relative to the reference root).
 * Build with -O2 -ffp-contract=off (oracle/Makefile) so every floating-point
 * operation rounds exactly as in the reference built with the same flags.
 *
 * Parity pins (tests/test_oracle_*.py):
 *  - mt19937_64 / generate_canonical / uniform_int: the C++ standard's
 *    known answer and a g++-compiled libstdc++ probe (oracle/stdrng_probe.cpp).
 *  - Philox4x32-10: Random123 known-answer vectors.
 *  - chain, accumulators, estimators, calibration, decorrelation, integrate:
 *    golden vectors written by the reference itself (oracle/_ref, built from
 *    the reference sources by oracle/Makefile) and closed forms.
 */
#include "mci_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* errors                                                                     */
/* ------------------------------------------------------------------------- */
static _Thread_local char g_err[512] = "";

const char *orc_last_error(void) { return g_err; }

static mci_status fail(mci_status code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define TRY(expr) do { mci_status _s = (expr); if (_s != MCI_OK) return _s; } while (0)

/* ------------------------------------------------------------------------- */
/* mt19937_64 — libstdc++ mersenne_twister_engine (bits/random.tcc) with the  */
/* parameters of std::mt19937_64 ([rand.predef]); used as in                 */
/* include/mci/MCIntegrator.hpp:32-34 and src/MCIntegrator.cpp:627.           */
/* ------------------------------------------------------------------------- */
#define MT_N 312
#define MT_M 156
#define MT_A 0xb5026f5aa96619e9ULL
#define MT_UPPER 0xffffffff80000000ULL
#define MT_LOWER 0x000000007fffffffULL

typedef struct { uint64_t x[MT_N]; int p; } mt64;

static void mt64_seed(mt64 *g, uint64_t s)
{
    g->x[0] = s;
    for (int i = 1; i < MT_N; ++i) {
        uint64_t v = g->x[i - 1];
        v ^= v >> 62;
        v *= 6364136223846793005ULL;
        v += (uint64_t)i;
        g->x[i] = v;
    }
    g->p = MT_N;
}

static void mt64_twist(mt64 *g)
{
    int k;
    for (k = 0; k < MT_N - MT_M; ++k) {
        uint64_t y = (g->x[k] & MT_UPPER) | (g->x[k + 1] & MT_LOWER);
        g->x[k] = g->x[k + MT_M] ^ (y >> 1) ^ ((y & 1ULL) ? MT_A : 0ULL);
    }
    for (; k < MT_N - 1; ++k) {
        uint64_t y = (g->x[k] & MT_UPPER) | (g->x[k + 1] & MT_LOWER);
        g->x[k] = g->x[k + (MT_M - MT_N)] ^ (y >> 1) ^ ((y & 1ULL) ? MT_A : 0ULL);
    }
    uint64_t y = (g->x[MT_N - 1] & MT_UPPER) | (g->x[0] & MT_LOWER);
    g->x[MT_N - 1] = g->x[MT_M - 1] ^ (y >> 1) ^ ((y & 1ULL) ? MT_A : 0ULL);
    g->p = 0;
}

static uint64_t mt64_next(mt64 *g)
{
    if (g->p >= MT_N) mt64_twist(g);
    uint64_t z = g->x[g->p++];
    z ^= (z >> 29) & 0x5555555555555555ULL;
    z ^= (z << 17) & 0x71d67fffeda60000ULL;
    z ^= (z << 37) & 0xfff7eee000000000ULL;
    z ^= (z >> 43);
    return z;
}

uint64_t orc_mt64_nth(uint64_t seed, int64_t n)
{
    mt64 g;
    mt64_seed(&g, seed);
    uint64_t v = 0;
    for (int64_t i = 0; i < n; ++i) v = mt64_next(&g);
    return v;
}

void orc_mt64_words(uint64_t seed, int64_t n, uint64_t *out)
{
    mt64 g;
    mt64_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) out[i] = mt64_next(&g);
}

/* generate_canonical<double, 53>(mt19937_64): one engine word,
 * ret = double(word) / 2^64, clamped to nextafter(1, 0) (bits/random.tcc generate_canonical). */
double orc_canonical53(uint64_t w)
{
    double r = (double)w / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* uniform_int_distribution<int>(0, range-1) on a 64-bit engine: Lemire's nearly
 * divisionless method with a 128-bit product (bits/uniform_int_dist.h _S_nd). */
uint64_t orc_lemire64(const uint64_t *words, int64_t nwords, uint64_t range, int64_t *used)
{
    int64_t k = 0;
    unsigned __int128 prod = (unsigned __int128)words[k++] * range;
    uint64_t low = (uint64_t)prod;
    if (low < range) {
        uint64_t thr = (0 - range) % range;
        while (low < thr && k < nwords) {
            prod = (unsigned __int128)words[k++] * range;
            low = (uint64_t)prod;
        }
    }
    if (used) *used = k;
    return (uint64_t)(prod >> 64);
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11; Random123 philox.h round/bumpkey).    */
/* Native random mode of this project; no counterpart in the reference.       */
/* ------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------- */
/* deterministic cross-walker sum: chunks of 256 values (zero padded), pairwise */
/* tree inside a chunk (stride 128, 64, ..., 1), chunk sums added in order.    */
/* The device reductions use exactly this order.                              */
/* ------------------------------------------------------------------------- */
double orc_det_sum(const double *v, int64_t n)
{
    double total = 0.0;
    int first = 1;
    for (int64_t c0 = 0; c0 < n; c0 += 256) {
        double buf[256];
        for (int i = 0; i < 256; ++i) buf[i] = (c0 + i < n) ? v[c0 + i] : 0.0;
        for (int s = 128; s > 0; s >>= 1)
            for (int i = 0; i < s; ++i) buf[i] = buf[i] + buf[i + s];
        if (first) { total = buf[0]; first = 0; }
        else total = total + buf[0];
    }
    return total;
}

/* ------------------------------------------------------------------------- */
/* functor descriptors                                                        */
/* ------------------------------------------------------------------------- */
typedef struct {
    int kind, ndim, nparams, nproto;
    double *params;
    /* derived (GMM_ISO) */
    int K;
    double *gw, *gs, *gmu, *gc, *gb; /* weights, widths, means[K][ndim], log-weight terms, 1/(2 s^2) */
} pdf_desc;

typedef struct {
    int kind, ndim, nparams, nobs, updateable;
    double *params;
    int blocksize, nskip, flag_equil, estim;
    int dep;          /* MCI_OBS_DEP_SQUARE: index of the observable it squares, in its container */
} obs_desc;

/* accumulator types, Factories.hpp:29-43 */
enum { ACC_SIMPLE = 0, ACC_FULL = 1, ACC_BLOCK = 2 };

/* AccumulatorInterface state (AccumulatorInterface.hpp:20-60) + child storage */
typedef struct {
    int type, nobs, xndim, nskip, blocksize, updateable;
    double *obs_values;
    uint8_t *flags_xchanged;
    int nchanged;
    int64_t nsteps, stepidx;
    int skipidx, flag_final;
    double *data;
    double *shadow;   /* FP32 chain only: the same block means summed step by step in FP64 */
    int64_t nstore;   /* Simple: 1 row; Full: NAccu; Block: nblocks */
    int bidx;
    int64_t storeidx;
} accu;

typedef struct {
    double *xold, *xnew;
    int nchanged;
    int *changedIdx;
    int accepted;
    double *pold, *pnew;   /* concatenated proto values of all pdfs */
    double *steps;         /* one per type */
    int64_t acc, rej;
    mt64 mt;
    uint64_t t;            /* Philox: steps taken */
    uint32_t k, kx;        /* Philox: next word within the step / extra words */
    uint32_t wid;
    double nsv[2];         /* saved second normals of the move's distribution (two normal generators at most) */
    int nsf;               /* bit s: nsv[s] available */
    double *ms_pold, *ms_pnew; /* MultiStepMove: proto values of its own sampling functions */
    double *ms_x0;         /* MultiStepMove: the position the move starts from */
} walker;

#define MAX_PDF 16
#define MAX_OBS 32

struct orc_ctx {
    int ndim;
    int64_t W;
    int rng, prec, coupling;
    int fp32_shadow;   /* FP32 chain: also sum the block means step by step in FP64 */
    uint64_t seed;
    uint64_t *wseeds;
    double target;
    int nfind;
    int64_t ndecorr;
    /* domain */
    int periodic;
    double *lb, *ub;
    /* trial move (with a MultiStepMove: its contained move) */
    int movetype, srrd, veclen, nvecs, ntypes;
    int *typeEnds;
    int ms_on, ms_nsteps;   /* MultiStepMove(ndim, nsteps) around the move above */
    int lm_K;               /* the independence half of a mixed move: mixture components (0: off) */
    double lm_pmix;
    double *lm;             /* mu [K][ndim], 1 / sigma [K][ndim], sigma [K][ndim], lc [K], cw [K] */
    int ms_npdf, ms_nproto;
    pdf_desc ms_pdf[MAX_PDF];
    /* functors */
    int npdf;
    pdf_desc pdf[MAX_PDF];
    int nproto_total;
    int nobs;
    obs_desc obs[MAX_OBS];
    /* walkers */
    walker *w;
    /* per-walker results of last integrate */
    double *wavg, *werr;
    int nobsdim_last;
    int64_t nwarn;
    /* accumulators of the last main sampling, [walker][obs] */
    accu *acc_last;
    int acc_last_nobs;
    /* trace */
    double *trace_x;
    uint8_t *trace_acc;
    int64_t trace_cap, trace_n;
    int tracing;
};

/* ------------------------------------------------------------------------- */
/* per-walker random streams                                                  */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* seed of walker w when only MCI::setSeed(seed) was called: w == 0 keeps the seed
 * (so one walker reproduces one reference MCI), others get splitmix64 draws. */
static uint64_t walker_seed(uint64_t seed, int64_t w)
{
    if (w == 0) return seed;
    uint64_t s = (seed ^ 0x6A09E667F3BCC909ULL) + (uint64_t)w * 0x9E3779B97F4A7C15ULL;
    return splitmix64(&s);
}

/* The native stream (Philox4x32-10, key = seed).  Steps go in pairs: block j of step pair
 * t >> 1 is Philox(counter = (j, walker, t >> 1)); main word k of step t is word
 * 2 (t & 1) + (k & 1) of block k >> 1, so one block holds words 0 and 1 of two consecutive steps.
 * Main word 0 of a step is its shared word (see draw_accept); the draws of the move start at
 * word 1.  Two side streams keyed by the step itself: the Lemire rejections (block 0x20000000 +
 * (kx >> 2), word kx & 3) and the low word of the acceptance uniform (block 0x30000000, word 0). */
static uint32_t ph_block_word(const struct orc_ctx *c, const walker *w, uint32_t blk, uint64_t ct, uint32_t word)
{
    uint32_t ctr[4] = { blk, w->wid, (uint32_t)ct, (uint32_t)(ct >> 32) };
    uint32_t key[2] = { (uint32_t)c->seed, (uint32_t)(c->seed >> 32) };
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[word];
}

static uint32_t ph_word(const struct orc_ctx *c, walker *w, uint32_t k)
{
    return ph_block_word(c, w, k >> 1, w->t >> 1, 2u * (uint32_t)(w->t & 1u) + (k & 1u));
}

static uint32_t ph_next(const struct orc_ctx *c, walker *w) { return ph_word(c, w, w->k++); }
static uint32_t ph_extra(const struct orc_ctx *c, walker *w)
{
    const uint32_t kx = w->kx++;
    return ph_block_word(c, w, 0x20000000u + (kx >> 2), w->t, kx & 3u);
}

static void step_begin(const struct orc_ctx *c, walker *w)
{
    (void)c;
    w->k = 1; /* main word 0 is the step's shared word */
    w->kx = 0;
}

static void step_end(const struct orc_ctx *c, walker *w)
{
    if (c->rng == MCI_RNG_PHILOX) ++w->t;
}

/* symmetric uniform in (-1,1): MT mode = uniform_real_distribution<double>(-1,1)
 * (TrialMoveInterface.hpp:120-124): (u*(b-a)) + a; native = (2w+1-2^32) * 2^-32 */
static double draw_sym(const struct orc_ctx *c, walker *w)
{
    if (c->rng == MCI_RNG_MT19937_64) {
        double u = orc_canonical53(mt64_next(&w->mt));
        return (u * 2.0) + (-1.0);
    }
    uint32_t v = ph_next(c, w);
    return (double)(2 * (int64_t)v + 1 - 4294967296LL) * 0x1p-32;
}

/* uniform in [0,1): MT mode = uniform_real_distribution<double>(0,1) (MCIntegrator.cpp:628,343);
 * native = 53 bits from two words */
static double draw_unit(const struct orc_ctx *c, walker *w)
{
    if (c->rng == MCI_RNG_MT19937_64) {
        double u = orc_canonical53(mt64_next(&w->mt));
        return (u * 1.0) + 0.0;
    }
    uint64_t a = ph_next(c, w), b = ph_next(c, w);
    uint64_t v = (a << 32) | b;
    return (double)(v >> 11) * 0x1p-53;
}

/* ------------------------------------------------------------------------- */
/* the move distributions (SRRDType, Factories.hpp:118-131) with createSymRRD's default parameters
 * (TrialMoveInterface.hpp:107-183).  The reference draws them with libstdc++'s distributions (GCC
 * 11.4, the system's; bits/random.h and bits/random.tcc), restated here: every uniform is
 * generate_canonical<double, 53> on the walker's engine (one word, as draw_unit), the normal is
 * Marsaglia's polar method keeping the pair's second value in the distribution object, the gamma is
 * Marsaglia-Tsang (boosted by u^(1/alpha) for alpha < 1), and the one-sided ones take a fair sign
 * from bernoulli_distribution(0.5) after the value (SymmetrizedPRRD, TrialMoveInterface.hpp:80-98).
 * Each distribution object of the reference is per walker (one MCI per walker), so is its saved
 * normal: slot 0 and slot 1 below.  The native (Philox) mode draws the same way on its own uniforms. */
/* ------------------------------------------------------------------------- */
static double canon(const struct orc_ctx *c, walker *w) { return draw_unit(c, w); }

/* normal_distribution<double>(0, 1)::operator() with its saved value in slot s */
static double nd_draw(const struct orc_ctx *c, walker *w, int s)
{
    double ret;
    if (w->nsf & (1 << s)) {
        w->nsf &= ~(1 << s);
        ret = w->nsv[s];
    }
    else {
        double x, y, r2;
        do {
            x = 2.0 * canon(c, w) - 1.0;
            y = 2.0 * canon(c, w) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2 * log(r2) / r2);
        w->nsv[s] = x * mult;
        w->nsf |= 1 << s;
        ret = y * mult;
    }
    return ret * 1.0 + 0.0;
}

/* gamma_distribution<double>(alpha, beta)::operator() with its normal in slot s */
static double gamma_draw(const struct orc_ctx *c, walker *w, int s, double alpha, double beta)
{
    const double malpha = alpha < 1.0 ? alpha + 1.0 : alpha;
    const double a1 = malpha - 1.0 / 3.0;
    const double a2 = 1.0 / sqrt(9.0 * a1);
    double u, v, n;
    do {
        do {
            n = nd_draw(c, w, s);
            v = 1.0 + a2 * n;
        } while (v <= 0.0);
        v = v * v * v;
        u = canon(c, w);
    } while (u > 1.0 - 0.0331 * n * n * n * n && (log(u) > (0.5 * n * n + a1 * (1.0 - v + log(v)))));
    if (alpha == malpha) return a1 * v * beta;
    do u = canon(c, w);
    while (u == 0.0);
    return pow(u, 1.0 / alpha) * a1 * v * beta;
}

/* one draw of the move's distribution */
static double srrd_draw(const struct orc_ctx *c, walker *w)
{
    double val;
    switch (c->srrd) {
    case MCI_SRRD_UNIFORM: return draw_sym(c, w);
    case MCI_SRRD_GAUSSIAN: return nd_draw(c, w, 0);
    case MCI_SRRD_STUDENT: { /* student_t_distribution(1): nd * sqrt(n / gd), gd = gamma(n/2, 2) */
        const double z = nd_draw(c, w, 0);
        return z * sqrt(1.0 / gamma_draw(c, w, 1, 0.5, 2.0));
    }
    case MCI_SRRD_CAUCHY: { /* cauchy_distribution(0, 1) */
        double u;
        do u = canon(c, w);
        while (u == 0.5);
        const double pi = 3.1415926535897932384626433832795029;
        return 0.0 + 1.0 * tan(pi * u);
    }
    case MCI_SRRD_EXPONENTIAL: val = -log(1.0 - canon(c, w)) / 1.0; break;
    case MCI_SRRD_GAMMA: val = gamma_draw(c, w, 0, 1.0, 1.0); break;
    case MCI_SRRD_WEIBULL: val = 1.0 * pow(-log(1.0 - canon(c, w)), 1.0 / 1.0); break;
    case MCI_SRRD_LOGNORMAL: val = exp(1.0 * nd_draw(c, w, 0) + 0.0); break;
    case MCI_SRRD_CHISQ: val = 2 * gamma_draw(c, w, 0, 0.5, 1.0); break;
    case MCI_SRRD_FISHER: { /* fisher_f_distribution(1, 1): (gd_x * n) / (gd_y * m), gd_x = gamma(m/2), gd_y = gamma(n/2) */
        const double gx = gamma_draw(c, w, 0, 0.5, 1.0);
        const double gy = gamma_draw(c, w, 1, 0.5, 1.0);
        val = (gx * 1.0) / (gy * 1.0);
        break;
    }
    default: return NAN;
    }
    /* SymmetrizedPRRD: bernoulli_distribution(0.5) on one canonical draw */
    return ((canon(c, w) - 0.0) < 0.5 * (1.0 - 0.0)) ? val : -val;
}

/* The shared word (main word 0 of a native step).  With a single-vector move over n = 2^m vectors
 * (m <= 11, not inside a MultiStepMove) its top m bits are the vector index; the acceptance
 * uniform is the 53 bits below them: u = ((word0 : low word) << m) >> 11, times 2^-53, the low word
 * from its side stream (the m <= 11 zero bits shifted in fall below the 53).  Otherwise m = 0 and
 * the index is drawn by draw_index.  Returns m, or -1 when the index is not shared. */
int orc_shared_index_bits(const orc_ctx *c)
{
    if (c->rng != MCI_RNG_PHILOX || c->movetype != MCI_MOVE_VEC || c->ms_on || c->nvecs < 1) return -1;
    const uint32_t n = (uint32_t)c->nvecs;
    if (n & (n - 1u)) return -1;
    int m = 0;
    while ((1u << m) < n) ++m;
    return m <= 11 ? m : -1;
}

/* the acceptance uniform of a Metropolis step: MT mode = uniform_real_distribution<double>(0,1)
 * (MCIntegrator.cpp:343); native = the shared word below its index bits, then the side-stream word */
static double draw_accept(const struct orc_ctx *c, walker *w)
{
    if (c->rng == MCI_RNG_MT19937_64) return draw_unit(c, w);
    const int ib = orc_shared_index_bits(c);
    const unsigned m = ib > 0 ? (unsigned)ib : 0u;
    const uint64_t hi = ph_word(c, w, 0), lo = ph_block_word(c, w, 0x30000000u, w->t, 0);
    const uint64_t v = ((hi << 32) | lo) << m;
    return (double)(v >> 11) * 0x1p-53;
}

/* uniform integer in [0, n): MT mode = uniform_int_distribution<int>(0, n-1) (SRRDVecMove.hpp:27,78);
 * native = Lemire on 32-bit words with rejection from the extra-word stream */
static int draw_index(const struct orc_ctx *c, walker *w, int n)
{
    if (c->rng == MCI_RNG_MT19937_64) {
        uint64_t range = (uint64_t)n;
        unsigned __int128 prod = (unsigned __int128)mt64_next(&w->mt) * range;
        uint64_t low = (uint64_t)prod;
        if (low < range) {
            uint64_t thr = (0 - range) % range;
            while (low < thr) {
                prod = (unsigned __int128)mt64_next(&w->mt) * range;
                low = (uint64_t)prod;
            }
        }
        return (int)(prod >> 64);
    }
    uint32_t range = (uint32_t)n;
    uint64_t prod = (uint64_t)ph_next(c, w) * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t thr = (0u - range) % range;
        while (low < thr) {
            prod = (uint64_t)ph_extra(c, w) * range;
            low = (uint32_t)prod;
        }
    }
    return (int)(prod >> 32);
}

/* ------------------------------------------------------------------------- */
/* sampling functions (SamplingFunctionInterface.hpp:49-57,92-97)             */
/* ------------------------------------------------------------------------- */
static double pdf_a(const pdf_desc *p, int i) { return (p->nparams >= p->ndim) ? p->params[i] : 1.0; }

/* log-density of the isotropic mixture: LSE_k (c_k - b_k r2_k), r2_k = sum_i (x_i - mu_ki)^2 */
static double gmm_iso_logp(const pdf_desc *p, const double *r2)
{
    double m = -INFINITY;
    for (int k = 0; k < p->K; ++k) {
        double t = p->gc[k] - p->gb[k] * r2[k];
        if (t > m) m = t;
    }
    double s = 0.0;
    for (int k = 0; k < p->K; ++k) s = s + exp((p->gc[k] - p->gb[k] * r2[k]) - m);
    return m + log(s);
}

/* protoFunction: fills proto[0..nproto) from x */
static void pdf_proto(const pdf_desc *p, const double *x, double *proto)
{
    switch (p->kind) {
    case MCI_PDF_GAUSS_ISO: { /* TestMCIFunctions.hpp:135-138 */
        double s = x[0] * x[0];
        for (int i = 1; i < p->ndim; ++i) s = s + x[i] * x[i];
        proto[0] = s;
        break;
    }
    case MCI_PDF_GAUSS_DIAG: /* TestMCIFunctions.hpp:159-164 */
        for (int i = 0; i < p->ndim; ++i) proto[i] = (x[i] * x[i]) * pdf_a(p, i);
        break;
    case MCI_PDF_EXP_DIAG: /* TestMCIFunctions.hpp:227-230 */
        for (int i = 0; i < p->ndim; ++i) proto[i] = fabs(x[i]) * pdf_a(p, i);
        break;
    case MCI_PDF_GMM_ISO: {
        for (int k = 0; k < p->K; ++k) {
            const double *mu = p->gmu + (size_t)k * p->ndim;
            double s = 0.0;
            for (int i = 0; i < p->ndim; ++i) { double d = x[i] - mu[i]; s = s + d * d; }
            proto[k] = s;
        }
        proto[p->K] = gmm_iso_logp(p, proto);
        break;
    }
    }
}

/* acceptanceFunction(protoold, protonew) */
static double pdf_acceptance(const pdf_desc *p, const double *po, const double *pn)
{
    switch (p->kind) {
    case MCI_PDF_GAUSS_ISO: /* TestMCIFunctions.hpp:145-148 */
        return exp(-pn[0] + po[0]);
    case MCI_PDF_GAUSS_DIAG:
    case MCI_PDF_EXP_DIAG: { /* TestMCIFunctions.hpp:171-176 (std::accumulate from 0.) */
        double e = 0.0;
        for (int i = 0; i < p->nproto; ++i) e = e + po[i];
        double s = 0.0;
        for (int i = 0; i < p->nproto; ++i) s = s + pn[i];
        e = e - s;
        return exp(e);
    }
    case MCI_PDF_GMM_ISO:
        return exp(pn[p->K] - po[p->K]);
    }
    return 0.0;
}

/* computeAcceptance (SamplingFunctionInterface.hpp:49-57) incl. updatedAcceptance overrides */
static double pdf_compute_acceptance(const pdf_desc *p, const walker *w, const double *po, double *pn)
{
    if (w->nchanged < p->ndim) {
        switch (p->kind) {
        case MCI_PDF_GAUSS_DIAG: { /* TestMCIFunctions.hpp:178-187 */
            double e = 0.0;
            for (int i = 0; i < w->nchanged; ++i) {
                int j = w->changedIdx[i];
                pn[j] = (w->xnew[j] * w->xnew[j]) * pdf_a(p, j);
                e = e + (pn[j] - po[j]);
            }
            return exp(-e);
        }
        case MCI_PDF_EXP_DIAG: { /* TestMCIFunctions.hpp:244-253 */
            double e = 0.0;
            for (int i = 0; i < w->nchanged; ++i) {
                int j = w->changedIdx[i];
                pn[j] = fabs(w->xnew[j]) * pdf_a(p, j);
                e = e + (pn[j] - po[j]);
            }
            return exp(-e);
        }
        case MCI_PDF_GMM_ISO: { /* incremental r2 update over the changed indices */
            for (int k = 0; k < p->K; ++k) {
                const double *mu = p->gmu + (size_t)k * p->ndim;
                double dr = 0.0;
                for (int i = 0; i < w->nchanged; ++i) {
                    int j = w->changedIdx[i];
                    double dn = w->xnew[j] - mu[j], dd = w->xold[j] - mu[j];
                    dr = dr + (dn * dn - dd * dd);
                }
                pn[k] = po[k] + dr;
            }
            pn[p->K] = gmm_iso_logp(p, pn);
            return exp(pn[p->K] - po[p->K]);
        }
        default: /* default updatedAcceptance = full recomputation (:92-97) */
            break;
        }
    }
    pdf_proto(p, w->xnew, pn);
    return pdf_acceptance(p, po, pn);
}

/* ------------------------------------------------------------------------- */
/* observables (ObservableFunctionInterface.hpp:48,61)                        */
/* ------------------------------------------------------------------------- */
static double obs_param(const obs_desc *o, int i, double dflt) { return (o->nparams > i) ? o->params[i] : dflt; }

/* what dependent observables read (DependentObservableInterface::registerDeps,
 * DependentObservableInterface.hpp:49-51): the sampling functions of the context with the walker's
 * committed proto values (what observationCallback receives, SamplingFunctionInterface.hpp:59-62,
 * 100-104), and the accumulators of the observable container being filled (their last values,
 * AccumulatorInterface.hpp:86) */
typedef struct {
    const struct orc_ctx *c;
    const walker *w;
    const accu *accs;
} obs_env;

static void obs_eval(const obs_desc *o, const double *x, double *out, const obs_env *env)
{
    int n = o->ndim;
    switch (o->kind) {
    case MCI_OBS_X2: /* TestMCIFunctions.hpp:446-451 */
        for (int i = 0; i < n; ++i) out[i] = x[i] * x[i];
        break;
    case MCI_OBS_X2_SUM: { /* TestMCIFunctions.hpp:422-427 (X2Sum), :270-273 (XSquared, c = 3) */
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = s + x[i] * x[i];
        out[0] = s / obs_param(o, 0, 1.0);
        break;
    }
    case MCI_OBS_X: /* TestMCIFunctions.hpp:345-348 */
        for (int i = 0; i < n; ++i) out[i] = x[i];
        break;
    case MCI_OBS_X_SUM: { /* TestMCIFunctions.hpp:398-404 */
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = s + x[i];
        out[0] = s;
        break;
    }
    case MCI_OBS_CONST: /* TestMCIFunctions.hpp:378-381 */
        out[0] = obs_param(o, 0, 1.0);
        break;
    case MCI_OBS_EXP_X2: {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = s + x[i] * x[i];
        out[0] = exp(-(s / obs_param(o, 0, (double)n)));
        break;
    }
    case MCI_OBS_INDICATOR: {
        int j = (int)obs_param(o, 0, 0.0);
        out[0] = (x[j] > 0.0) ? 1.0 : 0.0;
        break;
    }
    case MCI_OBS_GAUSS_X2: { /* TestMCIFunctions.hpp:285-293 */
        double s = 0.0;
        for (int i = 0; i < n; ++i) s = s + x[i] * x[i];
        out[0] = exp(-s) * s * obs_param(o, 0, 1. / sqrt(M_PI * M_PI * M_PI) / 3.);
        break;
    }
    case MCI_OBS_PDF_LOG: case MCI_OBS_PDF_PROTO: { /* the proto values of sampling function j at the accepted state */
        const int j = (int)obs_param(o, 0, 0.0);
        int off = 0;
        for (int p = 0; p < j; ++p) off += env->c->pdf[p].nproto;
        const pdf_desc *pd = &env->c->pdf[j];
        const double *pv = env->w->pold + off;
        if (o->kind == MCI_OBS_PDF_PROTO) {
            for (int i = 0; i < o->nobs; ++i) out[i] = pv[i];
        }
        else if (pd->kind == MCI_PDF_GMM_ISO) out[0] = pv[pd->K];
        else {
            double s = 0.0;
            for (int i = 0; i < pd->nproto; ++i) s = s + pv[i];
            out[0] = -s;
        }
        break;
    }
    case MCI_OBS_DEP_SQUARE: { /* the last values of an earlier observable of the container */
        const double *v = env->accs[o->dep].obs_values;
        for (int i = 0; i < o->nobs; ++i) out[i] = v[i] * v[i];
        break;
    }
    }
}

/* updatedObservable of the updateable kinds (X2, TestMCIFunctions.hpp:453-460) */
static void obs_update(const obs_desc *o, const double *x, const uint8_t *flags, double *out, const obs_env *env)
{
    if (o->kind == MCI_OBS_X2) {
        for (int i = 0; i < o->ndim; ++i)
            if (flags[i]) out[i] = x[i] * x[i];
        return;
    }
    obs_eval(o, x, out, env);
}

static int obs_nobs(int kind, int ndim)
{
    switch (kind) {
    case MCI_OBS_X2: case MCI_OBS_X: case MCI_OBS_PDF_PROTO: return ndim;
    case MCI_OBS_X2_SUM: case MCI_OBS_X_SUM: case MCI_OBS_CONST: case MCI_OBS_EXP_X2: case MCI_OBS_INDICATOR:
    case MCI_OBS_GAUSS_X2: case MCI_OBS_PDF_LOG: return 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* accumulators: AccumulatorInterface.cpp:5-137, Simple/Block/FullAccumulator */
/* ------------------------------------------------------------------------- */
static int64_t accu_naccu(const accu *a) { return (a->nsteps > 0) ? 1 + (a->nsteps - 1) / a->nskip : 0; }

static void accu_init_base(accu *a) /* AccumulatorInterface.cpp:24-32 */
{
    a->stepidx = 0;
    a->skipidx = a->nskip - 1;
    a->flag_final = 0;
    a->nchanged = a->xndim;
    if (a->updateable) memset(a->flags_xchanged, 1, (size_t)a->xndim);
}

static void accu_reset(accu *a) /* AccumulatorInterface.cpp:123-127 + child _reset */
{
    a->bidx = 0;
    a->storeidx = 0;
    if (a->data) memset(a->data, 0, sizeof(double) * (size_t)(a->nstore * a->nobs));
    accu_init_base(a);
}

static void accu_free(accu *a)
{
    free(a->obs_values); free(a->flags_xchanged); free(a->data); free(a->shadow);
    memset(a, 0, sizeof *a);
}

static mci_status accu_make(accu *a, const obs_desc *o, int xndim)
{
    memset(a, 0, sizeof *a);
    a->blocksize = o->blocksize;
    a->type = (o->blocksize == 0) ? ACC_SIMPLE : (o->blocksize == 1 ? ACC_FULL : ACC_BLOCK);
    a->nobs = o->nobs;
    a->xndim = xndim;
    a->nskip = o->nskip;
    a->updateable = o->updateable;
    a->obs_values = calloc((size_t)a->nobs, sizeof(double));
    a->flags_xchanged = calloc((size_t)xndim, 1);
    accu_init_base(a);
    return MCI_OK;
}

/* allocate (AccumulatorInterface.cpp:88-96 + BlockAccumulator.cpp:6-17, FullAccumulator.cpp:6-11,
 * SimpleAccumulator.cpp:6-11) */
static mci_status accu_allocate(accu *a, int64_t nsteps)
{
    free(a->data);
    a->data = NULL;
    a->nstore = 0;
    if (nsteps < 1) return fail(MCI_E_INVALID_ARGUMENT, "[AccumulatorInterface::allocate] Provided number of MC steps was < 1 .");
    a->nsteps = nsteps;
    int64_t naccu = accu_naccu(a);
    if (a->type == ACC_SIMPLE) a->nstore = 1;
    else if (a->type == ACC_FULL) a->nstore = naccu;
    else {
        if (naccu < a->blocksize)
            return fail(MCI_E_INVALID_ARGUMENT, "[BlockAccumulator::allocate] Requested number of accumulations is smaller than the requested block size.");
        if (naccu % a->blocksize != 0)
            return fail(MCI_E_INVALID_ARGUMENT, "[BlockAccumulator::allocate] Requested number of accumulations is not a multiple of the requested block size.");
        a->nstore = naccu / a->blocksize;
    }
    a->data = calloc((size_t)(a->nstore * a->nobs), sizeof(double));
    accu_reset(a);
    return MCI_OK;
}

static void accu_store(accu *a) /* child _accumulate */
{
    switch (a->type) {
    case ACC_SIMPLE: /* SimpleAccumulator.cpp:14-19 */
        for (int i = 0; i < a->nobs; ++i) a->data[i] = a->data[i] + a->obs_values[i];
        break;
    case ACC_BLOCK: /* BlockAccumulator.cpp:20-30 */
        for (int i = 0; i < a->nobs; ++i) a->data[a->storeidx + i] = a->data[a->storeidx + i] + a->obs_values[i];
        if (++a->bidx == a->blocksize) { a->bidx = 0; a->storeidx += a->nobs; }
        break;
    case ACC_FULL: /* FullAccumulator.cpp:14-18 */
        memcpy(a->data + a->storeidx, a->obs_values, sizeof(double) * (size_t)a->nobs);
        a->storeidx += a->nobs;
        break;
    }
}

/* accumulate (AccumulatorInterface.cpp:99-114, _processOld/_processFull/_processSelective :34-85) */
static void accu_accumulate(accu *a, const obs_desc *o, const walker *w, const obs_env *env)
{
    if (w->accepted || a->nchanged > 0) {
        if (a->updateable) {
            if (a->nchanged < a->xndim && w->accepted) {
                if (w->nchanged < a->xndim) {
                    for (int i = 0; i < w->nchanged; ++i) {
                        int j = w->changedIdx[i];
                        if (!a->flags_xchanged[j]) { a->flags_xchanged[j] = 1; ++a->nchanged; }
                    }
                } else {
                    a->nchanged = a->xndim;
                }
            }
            if (++a->skipidx == a->nskip) {
                a->skipidx = 0;
                if (a->nchanged < a->xndim) obs_update(o, w->xnew, a->flags_xchanged, a->obs_values, env);
                else obs_eval(o, w->xnew, a->obs_values, env);
                memset(a->flags_xchanged, 0, (size_t)a->xndim);
                a->nchanged = 0;
                accu_store(a);
            }
        } else {
            a->nchanged = a->xndim;
            if (++a->skipidx == a->nskip) {
                a->skipidx = 0;
                obs_eval(o, w->xnew, a->obs_values, env);
                a->nchanged = 0;
                accu_store(a);
            }
        }
    } else {
        if (++a->skipidx == a->nskip) { a->skipidx = 0; accu_store(a); }
    }
    ++a->stepidx;
}

/* finalize (AccumulatorInterface.cpp:117-121 + BlockAccumulator.cpp:33-41, SimpleAccumulator.cpp:22-28) */
static mci_status accu_finalize(accu *a)
{
    if (a->stepidx != a->nsteps)
        return fail(MCI_E_RUNTIME, "[AccumulatorInterface::finalize] Finalize was called, but number of accumulated steps do not match the planned amount.");
    if (!a->flag_final) {
        if (a->type == ACC_SIMPLE) {
            double normf = 1. / (double)accu_naccu(a);
            for (int i = 0; i < a->nobs; ++i) a->data[i] = a->data[i] * normf;
        } else if (a->type == ACC_BLOCK) {
            double normf = 1. / a->blocksize;
            for (int64_t i = 0; i < a->nstore * a->nobs; ++i) a->data[i] = a->data[i] * normf;
        }
    }
    a->flag_final = 1;
    return MCI_OK;
}

/* ------------------------------------------------------------------------- */
/* estimators: src/Estimators.cpp, src/MJBlocker.cpp                          */
/* ------------------------------------------------------------------------- */
static const double chi2_q95[64] = {
    /* 0.95 quantiles of the chi-squared distribution, df = 1..64, to 6 decimals
     * (MJBlocker.cpp:6-13; regenerated and checked with scipy.stats.chi2.ppf in tests) */
    3.841459, 5.991465, 7.814728, 9.487729, 11.070498, 12.591587, 14.067140, 15.507313,
    16.918978, 18.307038, 19.675138, 21.026070, 22.362032, 23.684791, 24.995790, 26.296228,
    27.587112, 28.869299, 30.143527, 31.410433, 32.670573, 33.924438, 35.172462, 36.415029,
    37.652484, 38.885139, 40.113272, 41.337138, 42.556968, 43.772972, 44.985343, 46.194260,
    47.399884, 48.602367, 49.801850, 50.998460, 52.192320, 53.383541, 54.572228, 55.758479,
    56.942387, 58.124038, 59.303512, 60.480887, 61.656233, 62.829620, 64.001112, 65.170769,
    66.338649, 67.504807, 68.669294, 69.832160, 70.993453, 72.153216, 73.311493, 74.468324,
    75.623748, 76.777803, 77.930524, 79.081944, 80.232098, 81.381015, 82.528727, 83.675261};

static int is_pow2(int64_t n) { return n > 0 && (((uint64_t)n & ((uint64_t)n - 1)) == 0); }

/* OneDimUncorrelatedEstimator, Estimators.cpp:37-57 */
static mci_status est_uncorr_1d(int64_t n, const double *x, double *avg, double *err)
{
    if (n < 2) return fail(MCI_E_INVALID_ARGUMENT, "[OneDimUncorrelatedEstimator] n must be larger than 1");
    double a = 0.0, e = 0.0;
    for (int64_t i = 0; i < n; ++i) a = a + x[i];
    for (int64_t i = 0; i < n; ++i) e = e + x[i] * x[i];
    const double norm = 1. / (double)n;
    a = a * norm;
    e = e * norm - a * a;
    if (e > 1.e-300) e = sqrt(e / ((double)n - 1.));
    else e = 0.;
    *avg = a;
    *err = e;
    return MCI_OK;
}

/* MultiDimUncorrelatedEstimator, Estimators.cpp:125-155 (stride = row length of x) */
static mci_status est_uncorr_nd(int64_t n, int ndim, const double *x, double *avg, double *err)
{
    if (n < 2) return fail(MCI_E_INVALID_ARGUMENT, "[MultiDimUncorrelatedEstimator] n must be larger than 1");
    for (int j = 0; j < ndim; ++j) { avg[j] = 0.; err[j] = 0.; }
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < ndim; ++j) {
            avg[j] = avg[j] + x[i * ndim + j];
            err[j] = err[j] + x[i * ndim + j] * x[i * ndim + j];
        }
    const double norm = 1. / (double)n;
    const double norm2 = 1. / ((double)n - 1.);
    for (int j = 0; j < ndim; ++j) {
        avg[j] = avg[j] * norm;
        err[j] = err[j] * norm - avg[j] * avg[j];
        if (err[j] > 1.e-300) err[j] = sqrt(err[j] * norm2);
        else err[j] = 0.;
    }
    return MCI_OK;
}

/* MultiDimBlockEstimator, Estimators.cpp:158-185 */
static mci_status est_block_nd(int64_t n, int ndim, const double *x, int64_t nblocks, double *avg, double *err)
{
    if (n < nblocks) return fail(MCI_E_INVALID_ARGUMENT, "MCI error MultiDimBlockEstimator() : n must be >= nblocks");
    const int64_t nper = n / nblocks;
    const double norm = 1. / (double)nper;
    double *av = calloc((size_t)(nblocks * ndim), sizeof(double));
    for (int64_t b = 0; b < nblocks; ++b)
        for (int64_t i = b * nper; i < (b + 1) * nper; ++i)
            for (int j = 0; j < ndim; ++j) av[b * ndim + j] = av[b * ndim + j] + x[i * ndim + j];
    for (int64_t i = 0; i < nblocks * ndim; ++i) av[i] = av[i] * norm;
    mci_status s = est_uncorr_nd(nblocks, ndim, av, avg, err);
    free(av);
    return s;
}

/* OneDimBlockEstimator, Estimators.cpp:60-77 */
static mci_status est_block_1d(int64_t n, const double *x, int64_t nblocks, double *avg, double *err)
{
    if (n < nblocks) return fail(MCI_E_INVALID_ARGUMENT, "[OneDimBlockEstimator] n must be >= nblocks");
    const int64_t nper = n / nblocks;
    const double norm = 1. / (double)nper;
    double *av = malloc(sizeof(double) * (size_t)nblocks);
    for (int64_t b = 0; b < nblocks; ++b) {
        double s = 0.0;
        for (int64_t i = b * nper; i < (b + 1) * nper; ++i) s = s + x[i];
        av[b] = s * norm;
    }
    mci_status st = est_uncorr_1d(nblocks, av, avg, err);
    free(av);
    return st;
}

/* calcErrDelta, Estimators.cpp:9-31 (central finite differences of order 2*mode) */
static double calc_err_delta(int mode, const double *e)
{
    const int im = 4;
    switch (mode) {
    case 1: return (-0.5 * e[im - 1] + 0.5 * e[im + 1]);
    case 2: return ((1. / 12.) * e[im - 2] - (2. / 3.) * e[im - 1] + (2. / 3.) * e[im + 1] - (1. / 12.) * e[im + 2]);
    case 3: return (-(1. / 60.) * e[im - 3] + (3. / 20.) * e[im - 2] - 0.75 * e[im - 1] + 0.75 * e[im + 1]
                    - (3. / 20.) * e[im + 2] + (1. / 60.) * e[im + 3]);
    case 4: return ((1. / 280.) * e[im - 4] - (4. / 105.) * e[im - 3] + 0.2 * e[im - 2] - 0.8 * e[im - 1]
                    + 0.8 * e[im + 1] - 0.2 * e[im + 2] + (4. / 105.) * e[im + 3] - (1. / 280.) * e[im + 4]);
    }
    return 0.;
}

#define FC_MIN 6
#define FC_MAX 50
#define FC_PLAT 4
#define FC_NAV (FC_MAX - FC_MIN + 1)
#define FC_NACCD (FC_NAV - 2 * FC_PLAT)

/* OneDimFCBlockerEstimator, Estimators.cpp:83-122 */
static mci_status est_fc_1d(int64_t n, const double *x, double *avg, double *err)
{
    if (n < FC_MAX) return fail(MCI_E_INVALID_ARGUMENT, "[OneDimFCBlockerEstimator] n must be >= 50");
    double av[FC_NAV], er[FC_NAV], accd[FC_NACCD];
    for (int i = 0; i < FC_NAV; ++i) TRY(est_block_1d(n, x, i + FC_MIN, &av[i], &er[i]));
    for (int i = 0; i < FC_NACCD; ++i) accd[i] = 0.;
    for (int i2 = FC_PLAT; i2 < FC_NACCD + FC_PLAT; ++i2)
        for (int i1 = 1; i1 <= FC_PLAT; ++i1) accd[i2 - FC_PLAT] = accd[i2 - FC_PLAT] + calc_err_delta(i1, er + i2 - 4);
    int imin = 0;
    for (int i2 = 1; i2 < FC_NACCD; ++i2)
        if (fabs(accd[i2]) < fabs(accd[imin])) imin = i2;
    imin += FC_PLAT;
    *avg = 0.2 * (av[imin - 2] + av[imin - 1] + av[imin] + av[imin + 1] + av[imin + 2]);
    *err = 0.2 * (er[imin - 2] + er[imin - 1] + er[imin] + er[imin + 1] + er[imin + 2]);
    return MCI_OK;
}

/* MultiDimFCBlockerEstimator, Estimators.cpp:191-246 */
static mci_status est_fc_nd(int64_t n, int ndim, const double *x, double *avg, double *err)
{
    if (n < FC_MAX) return fail(MCI_E_INVALID_ARGUMENT, "MCI error MultiDimFCBlockerEstimator() : n must be >= 50");
    double *av = malloc(sizeof(double) * FC_NAV * (size_t)ndim);
    double *er = malloc(sizeof(double) * FC_NAV * (size_t)ndim);
    double *accd = calloc((size_t)FC_NACCD * (size_t)ndim, sizeof(double));
    int *imin = calloc((size_t)ndim, sizeof(int));
    mci_status st = MCI_OK;
    for (int i1 = 0; i1 < FC_NAV && st == MCI_OK; ++i1) st = est_block_nd(n, ndim, x, i1 + FC_MIN, av + i1 * ndim, er + i1 * ndim);
    if (st == MCI_OK) {
        double eh[9];
        for (int i2 = FC_PLAT; i2 < FC_NACCD + FC_PLAT; ++i2)
            for (int j = 0; j < ndim; ++j) {
                for (int s = -4; s < 5; ++s) eh[s + 4] = er[(i2 + s) * ndim + j];
                for (int i1 = 1; i1 <= FC_PLAT; ++i1)
                    accd[(i2 - FC_PLAT) * ndim + j] = accd[(i2 - FC_PLAT) * ndim + j] + calc_err_delta(i1, eh);
            }
        for (int i2 = 1; i2 < FC_NACCD; ++i2)
            for (int j = 0; j < ndim; ++j)
                if (fabs(accd[i2 * ndim + j]) < fabs(accd[imin[j] * ndim + j])) imin[j] = i2;
        for (int j = 0; j < ndim; ++j) {
            int ix = imin[j] + FC_PLAT;
            avg[j] = 0.2 * (av[(ix - 2) * ndim + j] + av[(ix - 1) * ndim + j] + av[ix * ndim + j] + av[(ix + 1) * ndim + j] + av[(ix + 2) * ndim + j]);
            err[j] = 0.2 * (er[(ix - 2) * ndim + j] + er[(ix - 1) * ndim + j] + er[ix * ndim + j] + er[(ix + 1) * ndim + j] + er[(ix + 2) * ndim + j]);
        }
    }
    free(imin); free(accd); free(er); free(av);
    return st;
}

/* MJBlocker::estimate, MJBlocker.cpp:127-154 (Jonsson, PRE 98, 043304 (2018)).
 * Defined behaviour where the reference reads out of bounds (a level with zero
 * variance gives 0/0 in its test statistic): such a level contributes M = 0. */
static mci_status est_mj(int64_t n, int ndim, const double *xin, double *avg, double *err)
{
    if (n <= 0) return fail(MCI_E_INVALID_ARGUMENT, "[MJBlocker] ndata must be a natural number.");
    if (!is_pow2(n)) return fail(MCI_E_INVALID_ARGUMENT, "[MJBlocker] ndata must be a power of two.");
    int npow = 0;
    while (((int64_t)1 << npow) < n) ++npow;
    double *x = malloc(sizeof(double) * (size_t)(n * ndim));
    double *X = malloc(sizeof(double) * (size_t)(n * ndim));
    double *var = calloc((size_t)(npow * ndim) + 1, sizeof(double));
    double *gam = calloc((size_t)(npow * ndim) + 1, sizeof(double));
    double *M = calloc((size_t)(npow * ndim) + 1, sizeof(double));
    double *Ms = calloc((size_t)(npow * ndim) + 1, sizeof(double));
    memcpy(x, xin, sizeof(double) * (size_t)(n * ndim));
    /* _computeMean :34-44 */
    for (int j = 0; j < ndim; ++j) avg[j] = 0.;
    for (int64_t i = 0; i < n; ++i)
        for (int j = 0; j < ndim; ++j) avg[j] = avg[j] + x[i * ndim + j];
    for (int j = 0; j < ndim; ++j) avg[j] = avg[j] / (double)n;
    /* _initX :47-54 */
    for (int64_t i = 0; i < n * ndim; ++i) X[i] = x[i] - avg[i % ndim];
    int64_t nred = n;
    for (int k = 0; k < npow; ++k) {
        double *v = var + k * ndim, *g = gam + k * ndim;
        /* _gamma0 :59-68 */
        for (int j = 0; j < ndim; ++j) v[j] = 0.;
        for (int64_t i = 0; i < nred; ++i)
            for (int j = 0; j < ndim; ++j) v[j] = v[j] + X[i * ndim + j] * X[i * ndim + j];
        for (int j = 0; j < ndim; ++j) v[j] = v[j] / (double)nred;
        /* _gamma1 :71-80 */
        for (int j = 0; j < ndim; ++j) g[j] = 0.;
        for (int64_t i = 0; i < nred - 1; ++i)
            for (int j = 0; j < ndim; ++j) g[j] = g[j] + X[i * ndim + j] * X[(i + 1) * ndim + j];
        for (int j = 0; j < ndim; ++j) g[j] = g[j] / (double)nred;
        /* _transform :83-92 */
        for (int64_t i = 0; i < nred / 2; ++i)
            for (int j = 0; j < ndim; ++j) {
                x[i * ndim + j] = 0.5 * (x[(2 * i) * ndim + j] + x[(2 * i + 1) * ndim + j]);
                X[i * ndim + j] = x[i * ndim + j] - avg[j];
            }
        nred = nred / 2;
    }
    /* _generateM :106-123 */
    for (int i = 0; i < npow; ++i)
        for (int k = 0; k < ndim; ++k) {
            double g = gam[i * ndim + k], v = var[i * ndim + k];
            double m = (v == 0.0) ? 0.0 : pow(g / v, 2) * pow(2, npow - i);
            M[(npow - i - 1) * ndim + k] = m;
        }
    for (int i = 0; i < npow; ++i)
        for (int j = 0; j <= i; ++j)
            for (int k = 0; k < ndim; ++k) Ms[i * ndim + k] = Ms[i * ndim + k] + M[j * ndim + k];
    /* chi^2 search and error :140-153 */
    for (int l = 0; l < ndim; ++l) {
        int k;
        for (k = npow - 1; k >= 0; --k)
            if (Ms[k * ndim + l] < chi2_q95[k]) break;
        k = npow - (k + 1);
        if (k >= npow) k = npow - 1; /* unreachable with the M = 0 rule; kept as a guard */
        err[l] = sqrt(var[k * ndim + l] / pow(2, npow - k));
    }
    free(Ms); free(M); free(gam); free(var); free(X); free(x);
    return MCI_OK;
}

/* estimator dispatch: Factories.hpp:73-95, Estimators.cpp:250-293 */
static mci_status estimate_any(int estim, int64_t n, int ndim, const double *x, double *avg, double *err)
{
    switch (estim) {
    case MCI_ESTIM_NOOP: /* Estimators.cpp:289-293 */
        for (int j = 0; j < ndim; ++j) { avg[j] = x[j]; err[j] = 0.; }
        return MCI_OK;
    case MCI_ESTIM_UNCORRELATED:
        return (ndim > 1) ? est_uncorr_nd(n, ndim, x, avg, err) : est_uncorr_1d(n, x, avg, err);
    case MCI_ESTIM_FCBLOCKER:
        return (ndim > 1) ? est_fc_nd(n, ndim, x, avg, err) : est_fc_1d(n, x, avg, err);
    case MCI_ESTIM_MJBLOCKER:
        return est_mj(n, ndim, x, avg, err);
    case MCI_ESTIM_CORRELATED:
        if (!is_pow2(n)) return (ndim > 1) ? est_fc_nd(n, ndim, x, avg, err) : est_fc_1d(n, x, avg, err);
        return est_mj(n, ndim, x, avg, err);
    }
    return fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator.");
}

mci_status orc_estimate(int estimator, int64_t n, int ndim, const double *x, double *average, double *error)
{
    return estimate_any(estimator, n, ndim, x, average, error);
}

mci_status orc_block_estimate(int64_t n, int ndim, const double *x, int64_t nblocks, double *average, double *error)
{
    return (ndim > 1) ? est_block_nd(n, ndim, x, nblocks, average, error) : est_block_1d(n, x, nblocks, average, error);
}

/* ------------------------------------------------------------------------- */
/* context                                                                    */
/* ------------------------------------------------------------------------- */
static void walker_alloc(orc_ctx *c, walker *w)
{
    w->xold = calloc((size_t)c->ndim, sizeof(double));
    w->xnew = calloc((size_t)c->ndim, sizeof(double));
    w->changedIdx = calloc((size_t)c->ndim, sizeof(int));
    for (int i = 0; i < c->ndim; ++i) w->changedIdx[i] = i;
    w->nchanged = c->ndim;
    w->accepted = 1;
    w->steps = calloc(64, sizeof(double));
    w->pold = NULL;
    w->pnew = NULL;
    w->ms_x0 = calloc((size_t)c->ndim, sizeof(double));
    w->ms_pold = NULL;
    w->ms_pnew = NULL;
}

static void rebuild_ms_protos(orc_ctx *c)
{
    int np = 0;
    for (int i = 0; i < c->ms_npdf; ++i) np += c->ms_pdf[i].nproto;
    c->ms_nproto = np;
    for (int64_t k = 0; k < c->W; ++k) {
        walker *w = &c->w[k];
        free(w->ms_pold); free(w->ms_pnew);
        w->ms_pold = calloc((size_t)np + 1, sizeof(double));
        w->ms_pnew = calloc((size_t)np + 1, sizeof(double));
    }
}

static void rebuild_protos(orc_ctx *c)
{
    int np = 0;
    for (int i = 0; i < c->npdf; ++i) np += c->pdf[i].nproto;
    c->nproto_total = np;
    for (int64_t k = 0; k < c->W; ++k) {
        walker *w = &c->w[k];
        free(w->pold); free(w->pnew);
        w->pold = calloc((size_t)np + 1, sizeof(double));
        w->pnew = calloc((size_t)np + 1, sizeof(double));
    }
}

static void seed_walkers(orc_ctx *c)
{
    for (int64_t k = 0; k < c->W; ++k) {
        uint64_t s = c->wseeds ? c->wseeds[k] : walker_seed(c->seed, k);
        mt64_seed(&c->w[k].mt, s);
        c->w[k].t = 0;
        c->w[k].wid = (uint32_t)k;
    }
}

static void pdf_free(pdf_desc *p);

/* trial move construction: createSRRDAllMove / createSRRDVecMove (Factories.hpp:160-227),
 * TypedMoveInterface ctor (TypedMoveInterface.hpp:27-39), SRRDVecMove ctor checks (SRRDVecMove.hpp:39-50) */
static mci_status set_move(orc_ctx *c, int srrd, int veclen, int ntypes, const int *typeEnds, double init)
{
    if (srrd < MCI_SRRD_UNIFORM || srrd > MCI_SRRD_FISHER) return fail(MCI_E_DOMAIN, "[createSRRDAllMove] Unhandled SRRDType enumerator.");
    if (ntypes < 1) ntypes = 1;
    if (ntypes > 1 && typeEnds == NULL) return fail(MCI_E_INVALID_ARGUMENT, "[checkTrialMoveSanity] ntypes>1 requires passed typeEnds array.");
    if (veclen > 0 && c->ndim % veclen != 0)
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::setTrialMove] MCI's number of walkers must be a multiple of passed veclen.");
    if (ntypes > 64) return fail(MCI_E_INVALID_ARGUMENT, "[oracle] at most 64 step-size types");
    if (veclen > 0 && ntypes > 1)
        for (int i = 0; i < ntypes; ++i)
            if (typeEnds[i] % veclen != 0)
                return fail(MCI_E_INVALID_ARGUMENT, "[SRRDVecMove] All type end indices must be multiples of vector length.");
    free(c->typeEnds);
    c->typeEnds = malloc(sizeof(int) * (size_t)ntypes);
    if (ntypes > 1) memcpy(c->typeEnds, typeEnds, sizeof(int) * (size_t)ntypes);
    else c->typeEnds[0] = c->ndim;
    c->ntypes = ntypes;
    c->srrd = srrd;
    c->movetype = (veclen > 0) ? MCI_MOVE_VEC : MCI_MOVE_ALL;
    c->veclen = (veclen > 0) ? veclen : 0;
    c->nvecs = (veclen > 0) ? c->ndim / veclen : 0;
    for (int64_t k = 0; k < c->W; ++k) {
        for (int i = 0; i < ntypes; ++i) c->w[k].steps[i] = init;
        c->w[k].nsf = 0; /* a new move: fresh distribution objects */
    }
    c->ms_on = 0;
    c->ms_nsteps = 0;
    while (c->ms_npdf > 0) pdf_free(&c->ms_pdf[--c->ms_npdf]);
    c->lm_K = 0; /* a new trial move replaces a mixed one too */
    c->lm_pmix = 0.;
    return MCI_OK;
}

mci_status orc_create(int ndim, int64_t nwalkers, orc_ctx **out)
{
    if (ndim < 1) return fail(MCI_E_INVALID_ARGUMENT, "[ProtoFunctionInterface] Number of dimensions must be at least 1.");
    if (nwalkers < 1) return fail(MCI_E_INVALID_ARGUMENT, "number of walkers must be at least 1");
    orc_ctx *c = calloc(1, sizeof *c);
    c->ndim = ndim;
    c->W = nwalkers;
    c->rng = MCI_RNG_MT19937_64;
    c->prec = MCI_PREC_FP64;
    c->coupling = MCI_COUPLING_INDEPENDENT;
    c->seed = 5489u; /* a fixed default instead of std::random_device (MCIntegrator.cpp:627) */
    c->target = 0.5;
    c->nfind = -50;
    c->ndecorr = -10000;
    c->lb = calloc((size_t)ndim, sizeof(double));
    c->ub = calloc((size_t)ndim, sizeof(double));
    c->w = calloc((size_t)nwalkers, sizeof(walker));
    for (int64_t k = 0; k < nwalkers; ++k) walker_alloc(c, &c->w[k]);
    seed_walkers(c);
    set_move(c, MCI_SRRD_UNIFORM, 0, 1, NULL, 0.05); /* MCI ctor default: MoveType::All, DEFAULT_MRT2STEP 0.05 */
    rebuild_protos(c);
    *out = c;
    return MCI_OK;
}

static void free_last_accus(orc_ctx *c)
{
    if (c->acc_last) {
        for (int64_t i = 0; i < c->W * c->acc_last_nobs; ++i) accu_free(&c->acc_last[i]);
        free(c->acc_last);
        c->acc_last = NULL;
    }
}

mci_status orc_destroy(orc_ctx *c)
{
    if (!c) return MCI_OK;
    free_last_accus(c);
    for (int64_t k = 0; k < c->W; ++k) {
        walker *w = &c->w[k];
        free(w->xold); free(w->xnew); free(w->changedIdx); free(w->steps); free(w->pold); free(w->pnew);
    }
    for (int i = 0; i < c->npdf; ++i) {
        free(c->pdf[i].params); free(c->pdf[i].gw); free(c->pdf[i].gs); free(c->pdf[i].gmu);
        free(c->pdf[i].gc); free(c->pdf[i].gb);
    }
    for (int i = 0; i < c->nobs; ++i) free(c->obs[i].params);
    free(c->w); free(c->lb); free(c->ub); free(c->typeEnds); free(c->wseeds); free(c->wavg); free(c->werr);
    free(c->lm);
    free(c);
    return MCI_OK;
}

mci_status orc_set_seed(orc_ctx *c, uint64_t seed)
{
    c->seed = seed;
    free(c->wseeds);
    c->wseeds = NULL;
    seed_walkers(c);
    return MCI_OK;
}

mci_status orc_set_walker_seeds(orc_ctx *c, const uint64_t *seeds, int64_t n)
{
    if (n != c->W) return fail(MCI_E_INVALID_ARGUMENT, "need one seed per walker");
    free(c->wseeds);
    c->wseeds = malloc(sizeof(uint64_t) * (size_t)n);
    memcpy(c->wseeds, seeds, sizeof(uint64_t) * (size_t)n);
    seed_walkers(c);
    return MCI_OK;
}

mci_status orc_set_rng(orc_ctx *c, int k)
{
    if (k != MCI_RNG_MT19937_64 && k != MCI_RNG_PHILOX) return fail(MCI_E_DOMAIN, "unknown rng kind");
    c->rng = k;
    seed_walkers(c);
    return MCI_OK;
}

mci_status orc_set_precision(orc_ctx *c, int p)
{
    if (p != MCI_PREC_FP64 && p != MCI_PREC_FP32) return fail(MCI_E_DOMAIN, "unknown precision %d", p);
    c->prec = p;
    return MCI_OK;
}

mci_status orc_set_fp32_shadow(orc_ctx *c, int on)
{
    c->fp32_shadow = (on != 0);
    return MCI_OK;
}

mci_status orc_set_coupling(orc_ctx *c, int m)
{
    if (m != MCI_COUPLING_INDEPENDENT && m != MCI_COUPLING_SHARED) return fail(MCI_E_DOMAIN, "unknown coupling");
    c->coupling = m;
    return MCI_OK;
}

mci_status orc_set_target_acceptance_rate(orc_ctx *c, double r) { c->target = r; return MCI_OK; }
mci_status orc_set_nfind_mrt2_iterations(orc_ctx *c, int n) { c->nfind = n; return MCI_OK; }
mci_status orc_set_ndecorrelation_steps(orc_ctx *c, int64_t n) { c->ndecorr = n; return MCI_OK; }

/* ------------------------------------------------------------------------- */
/* domain: DomainInterface.hpp:12-17, UnboundDomain.hpp:25-35,                 */
/* OrthoPeriodicDomain.cpp:38-83                                              */
/* ------------------------------------------------------------------------- */
static const double DC_INF = (double)FLT_MAX;
static const double DC_INFX2 = (double)FLT_MAX + (double)FLT_MAX;

/* The reference wraps with unbounded while loops (OrthoPeriodicDomain.cpp:41-46); the rule
 * below is identical for coordinates within 64 box lengths and bounded beyond (fmod first),
 * non-finite coordinates are left unchanged.  The device uses the same rule. */
static void wrap1(double *x, double lo, double hi)
{
    const double len = hi - lo;
    if (!isfinite(*x)) return;
    if (*x < lo - 64.0 * len || *x > hi + 64.0 * len) {
        double r = fmod(*x - lo, len);
        if (r < 0.0) r = r + len;
        *x = lo + r;
    }
    int it = 0;
    while (*x < lo && it < 130) { *x = *x + len; ++it; }
    while (*x > hi && it < 260) { *x = *x - len; ++it; }
}

static void dom_apply_all(const orc_ctx *c, double *x)
{
    if (!c->periodic) return;
    for (int i = 0; i < c->ndim; ++i) wrap1(&x[i], c->lb[i], c->ub[i]);
}

static void dom_apply_wlk(const orc_ctx *c, walker *w)
{
    if (!c->periodic) return;
    for (int i = 0; i < w->nchanged; ++i) {
        int j = w->changedIdx[i];
        wrap1(&w->xnew[j], c->lb[j], c->ub[j]);
    }
}

static void dom_scale(const orc_ctx *c, double *x)
{
    for (int i = 0; i < c->ndim; ++i) {
        if (c->periodic) x[i] = c->lb[i] + x[i] * (c->ub[i] - c->lb[i]);
        else x[i] = -DC_INF + x[i] * DC_INFX2;
    }
}

static void dom_sizes(const orc_ctx *c, double *s)
{
    for (int i = 0; i < c->ndim; ++i) s[i] = c->periodic ? (c->ub[i] - c->lb[i]) : DC_INFX2;
}

static double dom_volume(const orc_ctx *c)
{
    if (!c->periodic) return 0.;
    double v = 1.;
    for (int i = 0; i < c->ndim; ++i) v = v * (c->ub[i] - c->lb[i]);
    return v;
}

mci_status orc_reset_domain(orc_ctx *c)
{
    c->periodic = 0;
    return MCI_OK;
}

mci_status orc_set_irange(orc_ctx *c, const double *lb, const double *ub)
{
    for (int i = 0; i < c->ndim; ++i)
        if (ub[i] <= lb[i])
            return fail(MCI_E_INVALID_ARGUMENT, "[OrthoPeriodicDomain::checkBounds] All upper bounds must be truly greater than their corresponding lower bounds.");
    memcpy(c->lb, lb, sizeof(double) * (size_t)c->ndim);
    memcpy(c->ub, ub, sizeof(double) * (size_t)c->ndim);
    c->periodic = 1;
    for (int64_t k = 0; k < c->W; ++k) dom_apply_all(c, c->w[k].xold);
    return MCI_OK;
}

mci_status orc_domain_is_finite(orc_ctx *c, int *f) { *f = dom_volume(c) != 0.; return MCI_OK; }

/* ------------------------------------------------------------------------- */
/* positions and step sizes (MCIntegrator.cpp:556-619)                        */
/* ------------------------------------------------------------------------- */
mci_status orc_set_x(orc_ctx *c, const double *x, int64_t n)
{
    if (n != c->ndim && n != c->ndim * c->W) return fail(MCI_E_INVALID_ARGUMENT, "x must have ndim or nwalkers*ndim entries");
    for (int64_t k = 0; k < c->W; ++k) {
        memcpy(c->w[k].xold, (n == c->ndim) ? x : x + k * c->ndim, sizeof(double) * (size_t)c->ndim);
        dom_apply_all(c, c->w[k].xold);
    }
    return MCI_OK;
}

mci_status orc_get_x(orc_ctx *c, double *x, int64_t n)
{
    if (n < c->ndim * c->W) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    for (int64_t k = 0; k < c->W; ++k) memcpy(x + k * c->ndim, c->w[k].xold, sizeof(double) * (size_t)c->ndim);
    return MCI_OK;
}

mci_status orc_center_x(orc_ctx *c) /* DomainInterface::getCenter DomainInterface.hpp:39-43 */
{
    for (int64_t k = 0; k < c->W; ++k) {
        for (int i = 0; i < c->ndim; ++i) c->w[k].xold[i] = 0.5;
        dom_scale(c, c->w[k].xold);
    }
    return MCI_OK;
}

static double trial_move(const orc_ctx *c, walker *w);

mci_status orc_move_x(orc_ctx *c) /* MCI::moveX MCIntegrator.cpp:606-612 */
{
    for (int64_t k = 0; k < c->W; ++k) {
        walker *w = &c->w[k];
        memcpy(w->xnew, w->xold, sizeof(double) * (size_t)c->ndim);
        step_begin(c, w);
        trial_move(c, w);
        step_end(c, w);
        if (w->nchanged < c->ndim) dom_apply_wlk(c, w);
        else dom_apply_all(c, w->xnew);
        memcpy(w->xold, w->xnew, sizeof(double) * (size_t)c->ndim);
    }
    return MCI_OK;
}

mci_status orc_new_random_x(orc_ctx *c) /* MCI::newRandomX MCIntegrator.cpp:614-619 */
{
    for (int64_t k = 0; k < c->W; ++k) {
        walker *w = &c->w[k];
        step_begin(c, w);
        for (int i = 0; i < c->ndim; ++i) w->xnew[i] = draw_unit(c, w);
        step_end(c, w);
        dom_scale(c, w->xnew);
        memcpy(w->xold, w->xnew, sizeof(double) * (size_t)c->ndim);
    }
    return MCI_OK;
}

mci_status orc_set_mrt2step(orc_ctx *c, int idx, double step)
{
    for (int64_t k = 0; k < c->W; ++k) {
        if (idx < 0) for (int i = 0; i < c->ntypes; ++i) c->w[k].steps[i] = step;
        else if (idx < c->ntypes) c->w[k].steps[idx] = step;
    }
    return MCI_OK;
}

mci_status orc_set_mrt2steps(orc_ctx *c, const double *steps, int n)
{
    if (n != c->ntypes) return fail(MCI_E_INVALID_ARGUMENT, "need one step per step-size type");
    for (int64_t k = 0; k < c->W; ++k) memcpy(c->w[k].steps, steps, sizeof(double) * (size_t)n);
    return MCI_OK;
}

mci_status orc_get_mrt2step(orc_ctx *c, int idx, double *step)
{
    *step = (idx >= 0 && idx < c->ntypes) ? c->w[0].steps[idx] : 0.;
    return MCI_OK;
}

mci_status orc_get_walker_mrt2steps(orc_ctx *c, double *steps, int64_t n)
{
    if (n < c->W * c->ntypes) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    for (int64_t k = 0; k < c->W; ++k) memcpy(steps + k * c->ntypes, c->w[k].steps, sizeof(double) * (size_t)c->ntypes);
    return MCI_OK;
}

mci_status orc_get_n_step_sizes(orc_ctx *c, int *n) { *n = c->ntypes; return MCI_OK; }

mci_status orc_set_trial_move(orc_ctx *c, int srrd, int veclen, int ntypes, const int *te)
{
    return set_move(c, srrd, veclen, ntypes, te, 0.05);
}

mci_status orc_set_trial_move_type(orc_ctx *c, int mt)
{
    if (mt == MCI_MOVE_ALL) return set_move(c, MCI_SRRD_UNIFORM, 0, 1, NULL, 0.05);
    if (mt == MCI_MOVE_VEC) return set_move(c, MCI_SRRD_UNIFORM, 1, 1, NULL, 0.05);
    if (mt == MCI_MOVE_MULTISTEP) return orc_set_trial_move_multistep(c, c->ndim, MCI_SRRD_UNIFORM, 1, 1, NULL); /* MultiStepMove(ndim) */
    return fail(MCI_E_DOMAIN, "[createMoveDefault] Unhandled MoveType enumerator.");
}

mci_status orc_get_change_rate(orc_ctx *c, double *r)
{
    *r = (c->movetype == MCI_MOVE_ALL) ? 1. : 1. / c->nvecs;
    if (c->ms_on) *r = fmin(1., *r * c->ms_nsteps); /* MultiStepMove.hpp:67 */
    return MCI_OK;
}

/* ------------------------------------------------------------------------- */
/* functor registration                                                       */
/* ------------------------------------------------------------------------- */
static mci_status make_pdf(int kind, int ndim, const double *params, int nparams, pdf_desc *out)
{
    pdf_desc p;
    memset(&p, 0, sizeof p);
    p.kind = kind;
    p.ndim = ndim;
    p.nparams = nparams;
    if (nparams > 0) {
        p.params = malloc(sizeof(double) * (size_t)nparams);
        memcpy(p.params, params, sizeof(double) * (size_t)nparams);
    }
    switch (kind) {
    case MCI_PDF_GAUSS_ISO: p.nproto = 1; break;
    case MCI_PDF_GAUSS_DIAG: case MCI_PDF_EXP_DIAG:
        if (nparams != 0 && nparams != ndim) { free(p.params); return fail(MCI_E_INVALID_ARGUMENT, "params must be empty or have ndim entries"); }
        p.nproto = ndim;
        break;
    case MCI_PDF_GMM_ISO: {
        if (nparams < 1) { free(p.params); return fail(MCI_E_INVALID_ARGUMENT, "GMM needs K"); }
        int K = (int)params[0];
        if (K < 1 || nparams != 1 + 2 * K + K * ndim) { free(p.params); return fail(MCI_E_INVALID_ARGUMENT, "GMM params must be [K, w[K], s[K], mu[K][ndim]]"); }
        p.K = K;
        p.nproto = K + 1;
        p.gw = malloc(sizeof(double) * (size_t)K);
        p.gs = malloc(sizeof(double) * (size_t)K);
        p.gc = malloc(sizeof(double) * (size_t)K);
        p.gb = malloc(sizeof(double) * (size_t)K);
        p.gmu = malloc(sizeof(double) * (size_t)(K * ndim));
        for (int k = 0; k < K; ++k) {
            p.gw[k] = params[1 + k];
            p.gs[k] = params[1 + K + k];
            p.gc[k] = log(p.gw[k]) - (double)ndim * log(p.gs[k]);
            p.gb[k] = 0.5 / (p.gs[k] * p.gs[k]);
        }
        memcpy(p.gmu, params + 1 + 2 * K, sizeof(double) * (size_t)(K * ndim));
        break;
    }
    default:
        free(p.params);
        return fail(MCI_E_DOMAIN, "unknown sampling-function kind %d", kind);
    }
    *out = p;
    return MCI_OK;
}

mci_status orc_add_sampling_function(orc_ctx *c, int kind, int ndim, const double *params, int nparams)
{
    if (ndim != c->ndim)
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addSamplingFunction] Passed sampling function's number of inputs is not equal to MCI's number of walkers.");
    if (c->npdf >= MAX_PDF) return fail(MCI_E_INVALID_ARGUMENT, "too many sampling functions");
    mci_status st = make_pdf(kind, ndim, params, nparams, &c->pdf[c->npdf]);
    if (st != MCI_OK) return st;
    c->npdf++;
    rebuild_protos(c);
    return MCI_OK;
}

/* MultiStepMove(ndim, nsteps) (MultiStepMove.hpp:40-46) with setTrialMove(the SRRD move built as
 * mci_set_trial_move does) (MultiStepMove.cpp:51-58); no own sampling functions yet */
mci_status orc_set_trial_move_multistep(orc_ctx *c, int nsteps, int srrd, int veclen, int ntypes, const int *te)
{
    if (nsteps < 0) return fail(MCI_E_INVALID_ARGUMENT, "the number of steps must not be negative");
    mci_status st = set_move(c, srrd, veclen, ntypes, te, 0.05);
    if (st != MCI_OK) return st;
    c->ms_on = 1;
    c->ms_nsteps = nsteps;
    rebuild_ms_protos(c);
    return MCI_OK;
}

/* MultiStepMove::addSamplingFunction (MultiStepMove.cpp:69-75) */
mci_status orc_add_multistep_sampling_function(orc_ctx *c, int kind, int ndim, const double *params, int nparams)
{
    if (!c->ms_on) return fail(MCI_E_RUNTIME, "the trial move is not a MultiStepMove");
    if (ndim != c->ndim)
        return fail(MCI_E_INVALID_ARGUMENT, "[MultiStepMove::addSamplingFunction] Passed sampling function's number of inputs is not equal to number of walkers.");
    if (c->ms_npdf >= MAX_PDF) return fail(MCI_E_INVALID_ARGUMENT, "too many sampling functions");
    mci_status st = make_pdf(kind, ndim, params, nparams, &c->ms_pdf[c->ms_npdf]);
    if (st != MCI_OK) return st;
    c->ms_npdf++;
    rebuild_ms_protos(c);
    return MCI_OK;
}

/* MultiStepMove::clearSamplingFunctions (MultiStepMove.cpp:64-67) */
mci_status orc_clear_multistep_sampling_functions(orc_ctx *c)
{
    if (!c->ms_on) return fail(MCI_E_RUNTIME, "the trial move is not a MultiStepMove");
    while (c->ms_npdf > 0) pdf_free(&c->ms_pdf[--c->ms_npdf]);
    rebuild_ms_protos(c);
    return MCI_OK;
}

static void pdf_free(pdf_desc *p)
{
    free(p->params); free(p->gw); free(p->gs); free(p->gmu); free(p->gc); free(p->gb);
}

mci_status orc_pop_sampling_function(orc_ctx *c)
{
    if (c->npdf == 0) return fail(MCI_E_RUNTIME, "no sampling function to pop");
    pdf_free(&c->pdf[--c->npdf]);
    rebuild_protos(c);
    return MCI_OK;
}

mci_status orc_clear_sampling_functions(orc_ctx *c)
{
    while (c->npdf > 0) pdf_free(&c->pdf[--c->npdf]);
    rebuild_protos(c);
    return MCI_OK;
}

mci_status orc_get_npdf(orc_ctx *c, int *n) { *n = c->npdf; return MCI_OK; }

/* MCI::addObservable MCIntegrator.cpp:455-469 */
mci_status orc_add_observable(orc_ctx *c, int kind, int ndim, const double *params, int nparams,
                              int blocksize, int nskip, int flag_equil, int estim)
{
    if (blocksize < 0) blocksize = 0;
    if (nskip < 1) nskip = 1;
    if (ndim != c->ndim)
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Passed observable function's number of inputs is not equal to MCI's number of walkers.");
    if (flag_equil && estim == MCI_ESTIM_NOOP)
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Requested automatic observable equilibration requires estimator with error calculation.");
    if (estim < MCI_ESTIM_NOOP || estim > MCI_ESTIM_MJBLOCKER) return fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator.");
    int nobs = obs_nobs(kind, ndim);
    int dep = -1;
    if (kind == MCI_OBS_DEP_SQUARE) { /* DependentObservableInterface::isObsDepValid (DependentObservableInterface.hpp:36-44) */
        dep = (nparams > 0) ? (int)params[0] : -1;
        if (nparams < 1 || params[0] != (double)dep || dep < 0 || dep >= c->nobs)
            return fail(MCI_E_INVALID_ARGUMENT, "a squared observable needs the index of an earlier observable");
        if (nskip % c->obs[dep].nskip != 0)
            return fail(MCI_E_INVALID_ARGUMENT, "the nskip of a dependent observable must be a multiple of its dependency's");
        nobs = c->obs[dep].nobs;
    }
    if ((kind == MCI_OBS_PDF_LOG || kind == MCI_OBS_PDF_PROTO) && nparams > 0 && (params[0] < 0 || params[0] >= MAX_PDF || params[0] != (int)params[0]))
        return fail(MCI_E_INVALID_ARGUMENT, "sampling-function index out of range");
    if (nobs < 0) return fail(MCI_E_DOMAIN, "unknown observable kind %d", kind);
    if (c->nobs >= MAX_OBS) return fail(MCI_E_INVALID_ARGUMENT, "too many observables");
    obs_desc o;
    memset(&o, 0, sizeof o);
    o.dep = dep;
    o.kind = kind; o.ndim = ndim; o.nparams = nparams; o.nobs = nobs;
    o.updateable = (kind == MCI_OBS_X2);
    if (nparams > 0) { o.params = malloc(sizeof(double) * (size_t)nparams); memcpy(o.params, params, sizeof(double) * (size_t)nparams); }
    o.blocksize = blocksize; o.nskip = nskip; o.flag_equil = flag_equil ? 1 : 0; o.estim = estim;
    c->obs[c->nobs++] = o;
    return MCI_OK;
}

mci_status orc_pop_observable(orc_ctx *c)
{
    if (c->nobs == 0) return fail(MCI_E_RUNTIME, "no observable to pop");
    free(c->obs[--c->nobs].params);
    return MCI_OK;
}

mci_status orc_clear_observables(orc_ctx *c)
{
    while (c->nobs > 0) free(c->obs[--c->nobs].params);
    return MCI_OK;
}

mci_status orc_get_nobs(orc_ctx *c, int *n) { *n = c->nobs; return MCI_OK; }

mci_status orc_get_nobsdim(orc_ctx *c, int *n)
{
    int s = 0;
    for (int i = 0; i < c->nobs; ++i) s += c->obs[i].nobs;
    *n = s;
    return MCI_OK;
}

/* ------------------------------------------------------------------------- */
/* Metropolis step                                                            */
/* ------------------------------------------------------------------------- */
/* SRRDAllMove::trialMove SRRDAllMove.hpp:67-80, SRRDVecMove::trialMove SRRDVecMove.hpp:75-96 */
static double srrd_move(const orc_ctx *c, walker *w, int outer)
{
    if (c->movetype == MCI_MOVE_ALL) {
        int xidx = 0;
        for (int t = 0; t < c->ntypes; ++t)
            while (xidx < c->typeEnds[t]) {
                w->xnew[xidx] = w->xnew[xidx] + w->steps[t] * srrd_draw(c, w);
                ++xidx;
            }
        w->nchanged = c->ndim;
        return 1.;
    }
    const int ib = outer ? orc_shared_index_bits(c) : -1;
    const int vidx = ib > 0 ? (int)(ph_word(c, w, 0) >> (32 - ib)) : ib == 0 ? 0 : draw_index(c, w, c->nvecs);
    const int xidx = vidx * c->veclen;
    int t = 0;
    while (t < c->ntypes) {
        if (xidx < c->typeEnds[t]) break;
        ++t;
    }
    for (int i = 0; i < c->veclen; ++i) {
        w->xnew[xidx + i] = w->xnew[xidx + i] + w->steps[t] * srrd_draw(c, w);
        w->changedIdx[i] = xidx + i;
    }
    w->nchanged = c->veclen;
    return 1.;
}

/* samplingFunction(protovalues) of each kind (the forms of oracle/ref_harness.cpp's classes) */
static double pdf_value(const pdf_desc *p, const double *pv)
{
    switch (p->kind) {
    case MCI_PDF_GAUSS_ISO: return exp(-pv[0]);
    case MCI_PDF_GAUSS_DIAG:
    case MCI_PDF_EXP_DIAG: {
        double s = 0.0;
        for (int i = 0; i < p->nproto; ++i) s = s + pv[i];
        return exp(-s);
    }
    case MCI_PDF_GMM_ISO: return exp(pv[p->K]);
    }
    return 0.0;
}

/* MultiStepMove::trialMove (MultiStepMove.cpp:6-47): nsteps Metropolis steps of the contained
 * move under the move's own sampling functions (none: every step accepted), from xold; xnew ends
 * at the last position, xold is restored, all coordinates count as changed, and the factor
 * oldPDF / newPDF of the own sampling functions is returned */
/* log q(y) of the independence mixture: log-sum-exp over components of lc_k - |(y - mu_k)/sigma_k|^2/2 */
static double lm_logq(const orc_ctx *c, const double *y)
{
    const int d = c->ndim, K = c->lm_K;
    const size_t KD = (size_t)K * d;
    const double *mu = c->lm, *isg = c->lm + KD, *lc = c->lm + 3 * KD;
    double e[8], mx = -HUGE_VAL;
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int j = 0; j < d; ++j) {
            const double z = (y[j] - mu[k * d + j]) * isg[k * d + j];
            s = s + z * z;
        }
        e[k] = lc[k] - 0.5 * s;
        if (e[k] > mx) mx = e[k];
    }
    double s = 0.0;
    for (int k = 0; k < K; ++k) s = s + exp(e[k] - mx);
    return mx + log(s);
}

/* the independence move of a mixed trial move: a component by cumulative weight, normals by the polar
 * method in pairs (nothing kept between steps), x' = mu_k + sigma_k z, factor q(x) / q(x') (the
 * reference's trial-move hook, TrialMoveInterface.hpp:63-69, MCIntegrator.cpp:343) */
static double lm_move(const orc_ctx *c, walker *w)
{
    const int d = c->ndim, K = c->lm_K;
    const size_t KD = (size_t)K * d;
    const double *mu = c->lm, *sg = c->lm + 2 * KD, *cw = c->lm + 3 * KD + K;
    const double uc = draw_unit(c, w);
    int k = K - 1;
    for (int kk = 0; kk < K - 1; ++kk)
        if (uc < cw[kk]) { k = kk; break; }
    for (int j = 0; j < d; j += 2) {
        double x, y, r2;
        do {
            x = 2.0 * draw_unit(c, w) - 1.0;
            y = 2.0 * draw_unit(c, w) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2 * log(r2) / r2);
        w->xnew[j] = mu[k * d + j] + sg[k * d + j] * (y * mult);
        if (j + 1 < d) w->xnew[j + 1] = mu[k * d + j + 1] + sg[k * d + j + 1] * (x * mult);
    }
    w->nchanged = d;
    return exp(lm_logq(c, w->xold) - lm_logq(c, w->xnew));
}

mci_status orc_set_learned_move(orc_ctx *c, int K, const double *weights, const double *means, const double *sigmas, double pmix)
{
    if (K < 1 || K > 8 || !weights || !means || !sigmas) return fail(MCI_E_INVALID_ARGUMENT, "need 1 to 8 mixture components");
    if (!(pmix >= 0. && pmix <= 1.)) return fail(MCI_E_INVALID_ARGUMENT, "mixing probability must be in [0, 1]");
    const int d = c->ndim;
    double wsum = 0.;
    for (int k = 0; k < K; ++k) {
        if (!(weights[k] > 0.)) return fail(MCI_E_INVALID_ARGUMENT, "mixture weights must be positive");
        wsum += weights[k];
    }
    for (int e = 0; e < K * d; ++e)
        if (!(sigmas[e] > 0.) || !isfinite(means[e])) return fail(MCI_E_INVALID_ARGUMENT, "mixture widths must be positive, means finite");
    const size_t KD = (size_t)K * d;
    free(c->lm);
    c->lm = malloc(sizeof(double) * (3 * KD + 2 * (size_t)K));
    double cum = 0.;
    for (int k = 0; k < K; ++k) {
        const double pi = weights[k] / wsum;
        double ls = 0.;
        for (int j = 0; j < d; ++j) {
            c->lm[(size_t)k * d + j] = means[(size_t)k * d + j];
            c->lm[KD + (size_t)k * d + j] = 1. / sigmas[(size_t)k * d + j];
            c->lm[2 * KD + (size_t)k * d + j] = sigmas[(size_t)k * d + j];
            ls = ls + log(1. / sigmas[(size_t)k * d + j]);
        }
        c->lm[3 * KD + k] = log(pi) + ls - 0.5 * (double)d * log(2. * M_PI);
        cum = cum + pi;
        c->lm[3 * KD + K + k] = (k == K - 1) ? 1. : cum;
    }
    c->lm_K = K;
    c->lm_pmix = pmix;
    return MCI_OK;
}

mci_status orc_clear_learned_move(orc_ctx *c)
{
    c->lm_K = 0;
    c->lm_pmix = 0.;
    return MCI_OK;
}

static double trial_move(const orc_ctx *c, walker *w)
{
    /* a mixed move: first the uniform that picks the independence move with probability lm_pmix */
    if (c->lm_K > 0 && c->lm_pmix > 0. && draw_unit(c, w) < c->lm_pmix) return lm_move(c, w);
    if (!c->ms_on) return srrd_move(c, w, 1);
    const int nd = c->ndim;
    memcpy(w->ms_x0, w->xold, sizeof(double) * (size_t)nd);
    int off = 0;
    for (int p = 0; p < c->ms_npdf; ++p) { /* initializeProtoValues(xold) */
        pdf_proto(&c->ms_pdf[p], w->xold, w->ms_pnew + off);
        memcpy(w->ms_pold + off, w->ms_pnew + off, sizeof(double) * (size_t)c->ms_pdf[p].nproto);
        off += c->ms_pdf[p].nproto;
    }
    double oldPDF = 1.;
    off = 0;
    for (int p = 0; p < c->ms_npdf; ++p) {
        oldPDF *= pdf_value(&c->ms_pdf[p], w->ms_pold + off);
        off += c->ms_pdf[p].nproto;
    }
    for (int s = 0; s < c->ms_nsteps; ++s) {
        const double moveAcc = srrd_move(c, w, 0);
        double pdfAcc = 1.;
        off = 0;
        for (int p = 0; p < c->ms_npdf; ++p) {
            pdfAcc *= pdf_compute_acceptance(&c->ms_pdf[p], w, w->ms_pold + off, w->ms_pnew + off);
            off += c->ms_pdf[p].nproto;
        }
        const int acc = (draw_unit(c, w) <= pdfAcc * moveAcc);
        if (acc) {
            memcpy(w->ms_pold, w->ms_pnew, sizeof(double) * (size_t)c->ms_nproto);
            memcpy(w->xold, w->xnew, sizeof(double) * (size_t)nd);
        }
        else {
            memcpy(w->ms_pnew, w->ms_pold, sizeof(double) * (size_t)c->ms_nproto);
            memcpy(w->xnew, w->xold, sizeof(double) * (size_t)nd);
        }
    }
    double newPDF = 1.;
    off = 0;
    for (int p = 0; p < c->ms_npdf; ++p) {
        newPDF *= pdf_value(&c->ms_pdf[p], w->ms_pold + off);
        off += c->ms_pdf[p].nproto;
    }
    memcpy(w->xold, w->ms_x0, sizeof(double) * (size_t)nd);
    w->nchanged = nd;
    w->accepted = 0;
    return oldPDF / newPDF;
}

/* MCI::initializeSampling MCIntegrator.cpp:252-271 (+ WalkerState::initialize WalkerState.hpp:41-49,
 * ProtoFunctionInterface::initializeProtoValues ProtoFunctionInterface.cpp:40-44) */
static void init_sampling(orc_ctx *c, walker *w)
{
    w->acc = 0;
    w->rej = 0;
    memcpy(w->xnew, w->xold, sizeof(double) * (size_t)c->ndim);
    w->nchanged = c->ndim;
    for (int i = 0; i < c->ndim; ++i) w->changedIdx[i] = i;
    w->accepted = 1;
    int off = 0;
    for (int p = 0; p < c->npdf; ++p) {
        pdf_proto(&c->pdf[p], w->xold, w->pnew + off);
        memcpy(w->pold + off, w->pnew + off, sizeof(double) * (size_t)c->pdf[p].nproto);
        off += c->pdf[p].nproto;
    }
}

/* MCI::doStepMRT2 MCIntegrator.cpp:326-360 (+ SamplingFunctionContainer::computeAcceptance
 * SamplingFunctionContainer.cpp:41-48) */
static void do_step_mrt2(orc_ctx *c, walker *w)
{
    step_begin(c, w);
    const double moveAcc = trial_move(c, w);
    if (w->nchanged < c->ndim) dom_apply_wlk(c, w);
    else dom_apply_all(c, w->xnew);
    double pdfAcc = 1.;
    int off = 0;
    for (int p = 0; p < c->npdf; ++p) {
        pdfAcc = pdfAcc * pdf_compute_acceptance(&c->pdf[p], w, w->pold + off, w->pnew + off);
        off += c->pdf[p].nproto;
    }
    const double u = draw_accept(c, w);
    w->accepted = (u <= pdfAcc * moveAcc);
    if (w->accepted) ++w->acc;
    else ++w->rej;
    step_end(c, w);
    size_t nb = sizeof(double) * (size_t)c->ndim, pb = sizeof(double) * (size_t)c->nproto_total;
    if (w->accepted) { memcpy(w->pold, w->pnew, pb); memcpy(w->xold, w->xnew, nb); }
    else { memcpy(w->pnew, w->pold, pb); memcpy(w->xnew, w->xold, nb); }
}

/* MCI::doStepRandom MCIntegrator.cpp:362-376 (no sampling function) */
static void do_step_random(orc_ctx *c, walker *w)
{
    step_begin(c, w);
    for (int i = 0; i < c->ndim; ++i) w->xnew[i] = draw_unit(c, w);
    step_end(c, w);
    dom_scale(c, w->xnew);
    w->nchanged = c->ndim;
    w->accepted = 1;
    ++w->acc;
    memcpy(w->xold, w->xnew, sizeof(double) * (size_t)c->ndim);
}

/* MCI::sample(npoints) MCIntegrator.cpp:273-288 */
static void sample_plain(orc_ctx *c, walker *w, int64_t n)
{
    init_sampling(c, w);
    for (int64_t i = 0; i < n; ++i) {
        if (c->npdf > 0) do_step_mrt2(c, w);
        else do_step_random(c, w);
    }
}

/* MCI::sample(npoints, container, flagMC) MCIntegrator.cpp:290-321 (PDF-observation bookkeeping is
 * out of scope: no catalogue observable depends on the PDF) */
static mci_status sample_obs(orc_ctx *c, walker *w, int64_t n, accu *accs, const obs_desc *obs, int nobs, int record)
{
    init_sampling(c, w);
    for (int i = 0; i < nobs; ++i) accu_reset(&accs[i]);
    for (int64_t s = 0; s < n; ++s) {
        if (c->npdf > 0) do_step_mrt2(c, w);
        else do_step_random(c, w);
        const obs_env env = { c, w, accs };
        for (int i = 0; i < nobs; ++i) accu_accumulate(&accs[i], &obs[i], w, &env);
        if (record && c->tracing && c->trace_n < c->trace_cap) {
            memcpy(c->trace_x + c->trace_n * c->ndim, w->xold, sizeof(double) * (size_t)c->ndim);
            c->trace_acc[c->trace_n] = (uint8_t)w->accepted;
            ++c->trace_n;
        }
    }
    for (int i = 0; i < nobs; ++i) TRY(accu_finalize(&accs[i]));
    return MCI_OK;
}

/* ------------------------------------------------------------------------- */
/* FP32 precision form (MCI_PREC_FP32).  The reference computes in FP64 only; this is the chain
 * of MCI::sample / doStepMRT2 (MCIntegrator.cpp:290-360) with SRRDVecMove's single-coordinate
 * uniform move (SRRDVecMove.hpp:75-96), one Gauss-form sampling function (p_i = a_i x_i^2,
 * TestMCIFunctions.hpp:151-190) and the X2 / X2_SUM / EXP_X2 observables into Block or Simple
 * accumulators (BlockAccumulator.cpp:20-41, SimpleAccumulator.cpp:14-28), carried out in FP32 as
 * the library's FP32 kernel does it:
 *  - per step one Philox block (t, walker): word 0 the index (Lemire, extra words on rejection),
 *    word 1 the proposal on the symmetric 24-bit grid, the high 24 bits of word 2 the uniform
 *    u on the 2^-24 grid; x' = fmaf(step, r, x); accept iff u <= bexp(-a_v (x' - x)(x' + x));
 *  - block sums by dwell time, as sums of positive terms: a value (x_i^2, |x|^2 or
 *    exp(-|x|^2/c)) enters its block sum once, when it is replaced or its chunk ends, times the
 *    steps it was held; the sums are Kahan-compensated FP32 pairs (the per-coordinate
 *    compensations stored in BF16) and each product enters exactly (its rounding error, by fused
 *    multiply-add, goes into the compensation); |x|^2 is itself a Kahan pair updated by the
 *    accepted changes; chunks of at most 1024 steps end at every block end, where the sums flush
 *    to FP64 block means ((S - C) / B); every chunk end closes all dwells and recomputes |x|^2
 *    from the position.
 * Calibration and the automatic decorrelation stay on the FP64 chain.  With c->fp32_shadow the
 * accumulators also get the block means of the same FP32 trajectory summed step by step in FP64
 * (the reference's BlockAccumulator applied to the FP32 chain), which bounds the FP32 sums' error. */
/* ------------------------------------------------------------------------- */
static void kahan32(float *s, float *cc, float y)
{
    const float yy = y - *cc;
    const float tt = *s + yy;
    *cc = (tt - *s) - yy;
    *s = tt;
}

static void kahan32_if(float *s, float *cc, float y, int on)
{
    const float yy = y - *cc;
    const float tt = *s + yy;
    const float c2 = (tt - *s) - yy;
    if (on) { *s = tt; *cc = c2; }
}

static uint32_t f2u_(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f_(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Kahan step adding the exact product a b (its rounding error, exact by fused multiply-add,
 * enters the compensation first) */
static void kahan32_prod(float *s, float *cc, float a, float b, int on)
{
    const float p = a * b;
    const float pe = fmaf(a, b, -p);
    const float yy = p - (*cc - pe);
    const float tt = *s + yy;
    const float c2 = (tt - *s) - yy;
    if (on) { *s = tt; *cc = c2; }
}

/* v rounded to BF16 (nearest, ties to even), as an FP32 value: the stored form of the per-coordinate
 * Kahan compensations (a compensation is below half an ulp of its sum, so 8 bits of it keep the pair
 * exact to about 2^-33 of the sum) */
static float bf16r(float v)
{
    const uint32_t u = f2u_(v);
    return u2f_(((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16);
}

static float sym24f(uint32_t wd) { return (float)(2 * (int32_t)(wd >> 8) + 1 - 16777216) * 0x1p-24f; }

/* the configuration the FP32 chain supports (the library refuses the others alike) */
static mci_status fp32_supported(const orc_ctx *c, const obs_desc *obs, int nobs)
{
    if (c->rng != MCI_RNG_PHILOX || c->movetype != MCI_MOVE_VEC || c->veclen != 1 || c->ntypes != 1 || c->periodic ||
        c->npdf != 1 || c->pdf[0].kind != MCI_PDF_GAUSS_DIAG || c->pdf[0].nparams < c->ndim || c->ms_on || c->lm_K > 0 || c->srrd != MCI_SRRD_UNIFORM)
        return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision supports Philox streams with single-coordinate uniform moves on one diagonal "
                                            "Gaussian (explicit weights) in an unbounded domain");
    int seen = 0, B = -1;
    for (int i = 0; i < nobs; ++i) {
        const obs_desc *o = &obs[i];
        const int bs = o->blocksize == 0 ? -2 : o->blocksize;
        if (o->nskip != 1 || o->blocksize == 1 || (o->kind != MCI_OBS_X2 && o->kind != MCI_OBS_X2_SUM && o->kind != MCI_OBS_EXP_X2) ||
            (seen & (1 << o->kind)) || (B != -1 && B != bs))
            return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision supports X2, X2_SUM and EXP_X2 observables (one each, one block size, Block or Simple accumulators, nskip 1)");
        seen |= 1 << o->kind;
        B = bs;
    }
    return MCI_OK;
}

/* n FP32 steps of walker w; accs/obs (nobs of them) may be NULL/0 (no accumulation) */
static mci_status sample_fp32(orc_ctx *c, walker *w, int64_t n, accu *accs, const obs_desc *obs, int nobs)
{
    TRY(fp32_supported(c, obs, nobs));
    const int d = c->ndim;
    const int acc_on = (nobs > 0);
    int ix2 = -1, isum = -1, iexp = -1;
    int64_t B = n;
    double x2sum_div = 1.0, exp_div = (double)d;
    for (int i = 0; i < nobs; ++i) {
        if (obs[i].kind == MCI_OBS_X2) ix2 = i;
        if (obs[i].kind == MCI_OBS_X2_SUM) { isum = i; x2sum_div = obs_param(&obs[i], 0, 1.0); }
        if (obs[i].kind == MCI_OBS_EXP_X2) { iexp = i; exp_div = obs_param(&obs[i], 0, (double)d); }
        B = (obs[i].blocksize == 0) ? n : obs[i].blocksize;
    }
    if (acc_on && B > INT32_MAX)
        return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision accumulates at most 2^31 - 1 steps per block (Simple accumulator: per run)");
    const int Bi = (int)B;
    const int64_t nblk = acc_on ? n / B : 0;
    for (int i = 0; i < nobs; ++i) {
        accu *a = &accs[i];
        if (a->nstore != nblk) return fail(MCI_E_RUNTIME, "FP32 chain: accumulator rows %lld, blocks %lld", (long long)a->nstore, (long long)nblk);
        memset(a->data, 0, sizeof(double) * (size_t)(a->nstore * a->nobs));
        free(a->shadow);
        a->shadow = c->fp32_shadow ? calloc((size_t)(a->nstore * a->nobs), sizeof(double)) : NULL;
    }
    float *xs = malloc(sizeof(float) * (size_t)d), *as = malloc(sizeof(float) * (size_t)d);
    float *S = calloc((size_t)d, sizeof(float)), *Cm = calloc((size_t)d, sizeof(float));
    int *jl = calloc((size_t)d, sizeof(int));           /* chunk step of the coordinate's last change */
    double *shx = calloc((size_t)d + 2, sizeof(double)); /* shadow sums: x_i^2 [d], |x|^2, exp */
    for (int i = 0; i < d; ++i) { xs[i] = (float)w->xold[i]; as[i] = (float)pdf_a(&c->pdf[0], i); }
    const float step = (float)w->steps[0];
    const float inv_ediv = (float)(1.0 / exp_div);
    float r = 0.f, rc = 0.f;
    for (int i = 0; i < d; ++i) kahan32(&r, &rc, xs[i] * xs[i]);
    float Sr = 0.f, Cr = 0.f, Se = 0.f, Ce = 0.f, e = 0.f;
    int jr = 0, je = 0, j = 0;
    if (iexp >= 0) e = orc_bexp(-r * inv_ediv);
    const uint32_t range = (uint32_t)c->nvecs;
    /* the index shares the step's word 0 with the 24-bit acceptance uniform when it takes at most 8 bits */
    const int ib0 = orc_shared_index_bits(c), ib = ib0 <= 8 ? ib0 : -1;
    int64_t b = 0, nacc = 0;
    const double norm = 1. / (double)Bi;
    for (int64_t s = 0; s < n;) {
        /* a chunk: at most 1024 steps (FAST32_REFRESH in mci_fast_plan.hpp), never past a block end */
        int64_t lim = n - s;
        if (acc_on && (int64_t)(Bi - j) < lim) lim = Bi - j;
        const int len = (int)(lim < 1024 ? lim : 1024);
        for (int q = 0; q < len; ++q) {
            /* the draws in the FP64 chain's order: main word 0 is shared (index bits, then the
             * acceptance uniform), the move's index (when not shared) and proposal follow */
            const uint32_t wd0 = ph_word(c, w, 0);
            int v;
            uint32_t ubits, wp;
            if (ib >= 0) {
                v = ib > 0 ? (int)(wd0 >> (32 - ib)) : 0;
                ubits = (wd0 << ib) >> 8;
                wp = ph_word(c, w, 1);
            }
            else {
                /* index: Lemire on main word 1, rejected draws from the extra-word blocks (draw_index) */
                uint64_t prod = (uint64_t)ph_word(c, w, 1) * range;
                uint32_t low = (uint32_t)prod;
                if (low < range) {
                    const uint32_t thr = (0u - range) % range;
                    w->kx = 0;
                    while (low < thr) {
                        prod = (uint64_t)ph_extra(c, w) * range;
                        low = (uint32_t)prod;
                    }
                }
                v = (int)(prod >> 32);
                ubits = wd0 >> 8;
                wp = ph_word(c, w, 2);
            }
            const float xo = xs[v];
            const float rr = sym24f(wp);
            const float u = (float)ubits * 0x1p-24f;
            ++w->t;
            const float xn = fmaf(step, rr, xo);
            const float dx2 = (xn - xo) * (xn + xo);
            const float ex = as[v] * dx2;
            const int acc = (u <= orc_bexp(-ex));
            nacc += acc;
            if (acc) xs[v] = xn;
            if (acc_on) {
                if (ix2 >= 0) { /* the old x_v^2 was held since chunk step jl */
                    kahan32_prod(&S[v], &Cm[v], xo * xo, (float)(q - jl[v]), acc);
                    Cm[v] = bf16r(Cm[v]);
                    if (acc) jl[v] = q;
                }
                if (isum >= 0) {
                    kahan32_prod(&Sr, &Cr, r, (float)(q - jr), acc);
                    if (acc) jr = q;
                }
                kahan32_if(&r, &rc, dx2, acc);
                if (iexp >= 0 && acc) { kahan32_prod(&Se, &Ce, e, (float)(q - je), 1); je = q; e = orc_bexp(-r * inv_ediv); }
                if (c->fp32_shadow) { /* FP64, step by step, over the same positions */
                    double qq = 0.0;
                    for (int i = 0; i < d; ++i) { const double xi = (double)xs[i]; shx[i] = shx[i] + xi * xi; qq = qq + xi * xi; }
                    shx[d] = shx[d] + qq;
                    shx[d + 1] = shx[d + 1] + exp(-(qq / exp_div));
                }
            }
        }
        s += len;
        if (acc_on) {
            /* the chunk ends: every value held counts until here; a block end flushes */
            j += len;
            const int bend = (j == Bi);
            float rx = 0.f, rxc = 0.f;
            for (int i = 0; i < d; ++i) {
                const float q2 = xs[i] * xs[i];
                if (ix2 >= 0) {
                    kahan32_prod(&S[i], &Cm[i], q2, (float)(len - jl[i]), 1);
                    if (bend) {
                        accs[ix2].data[b * d + i] = ((double)S[i] - (double)Cm[i]) * norm;
                        S[i] = Cm[i] = 0.f;
                    }
                    else Cm[i] = bf16r(Cm[i]);
                    jl[i] = 0;
                }
                kahan32(&rx, &rxc, q2);
            }
            if (isum >= 0) kahan32_prod(&Sr, &Cr, r, (float)(len - jr), 1);
            if (iexp >= 0) kahan32_prod(&Se, &Ce, e, (float)(len - je), 1);
            if (bend) {
                if (isum >= 0) accs[isum].data[b] = (((double)Sr - (double)Cr) * norm) / x2sum_div;
                if (iexp >= 0) accs[iexp].data[b] = ((double)Se - (double)Ce) * norm;
                if (c->fp32_shadow) {
                    if (ix2 >= 0) for (int i = 0; i < d; ++i) accs[ix2].shadow[b * d + i] = shx[i] * norm;
                    if (isum >= 0) accs[isum].shadow[b] = (shx[d] * norm) / x2sum_div;
                    if (iexp >= 0) accs[iexp].shadow[b] = shx[d + 1] * norm;
                    memset(shx, 0, sizeof(double) * ((size_t)d + 2));
                }
                Sr = Cr = Se = Ce = 0.f;
                j = 0;
                ++b;
            }
            jr = je = 0;
            r = rx;
            rc = rxc;
            if (iexp >= 0) e = orc_bexp(-r * inv_ediv);
        }
    }
    for (int i = 0; i < d; ++i) { w->xold[i] = (double)xs[i]; w->xnew[i] = w->xold[i]; }
    w->acc = nacc;
    w->rej = n - nacc;
    w->accepted = 0;
    for (int i = 0; i < nobs; ++i) { accs[i].stepidx = accs[i].nsteps; accs[i].flag_final = 1; }
    free(xs); free(as); free(S); free(Cm); free(jl); free(shx);
    return MCI_OK;
}

static double walker_rate(const walker *w) /* MCI::getAcceptanceRate MCIntegrator.cpp:587-592 */
{
    return (w->acc > 0) ? (double)w->acc / ((double)w->acc + (double)w->rej) : 0.;
}

/* ------------------------------------------------------------------------- */
/* calibration: MCI::findMRT2Step MCIntegrator.cpp:87-168                    */
/* walkers [w0, w1) share one loop when nranks > 1 (USE_MPI semantics)        */
/* ------------------------------------------------------------------------- */
static void find_mrt2(orc_ctx *c, int64_t w0, int64_t w1, int64_t nranks)
{
    const int ndim = c->ndim;
    const int64_t MIN_STAT = (int64_t)fmax(100., sqrt(40000. * ndim) / (double)nranks);
    const int MIN_CONS = 5;
    const double TOL = 0.05;
    const double SMALLEST = (double)FLT_MIN;
    double *sizes = malloc(sizeof(double) * (size_t)ndim);
    int *sidx = malloc(sizeof(int) * (size_t)ndim);
    double *rates = malloc(sizeof(double) * (size_t)(w1 - w0));
    dom_sizes(c, sizes);
    for (int i = 0; i < ndim; ++i) {
        int t = 0;
        while (t < c->ntypes && !(i < c->typeEnds[t])) ++t;
        sidx[i] = t;
    }
    int cons = 0, counter = 0;
    const int N = c->nfind;
    while ((N < 0 && cons < MIN_CONS) || counter < N) {
        for (int64_t k = w0; k < w1; ++k) {
            sample_plain(c, &c->w[k], MIN_STAT);
            rates[k - w0] = walker_rate(&c->w[k]);
        }
        double rate = (nranks > 1) ? orc_det_sum(rates, w1 - w0) / (double)nranks : rates[0];
        if (fabs(rate - c->target) < TOL) ++cons;
        else cons = 0;
        const double fact = fmin(2., fmax(0.5, rate / c->target));
        for (int64_t k = w0; k < w1; ++k) {
            double *st = c->w[k].steps;
            for (int t = 0; t < c->ntypes; ++t) st[t] = st[t] * fact;
            for (int i = 0; i < ndim; ++i)
                if (st[sidx[i]] > 0.5 * sizes[i]) st[sidx[i]] = 0.5 * sizes[i];
            for (int t = 0; t < c->ntypes; ++t)
                if (st[t] < SMALLEST) st[t] = SMALLEST;
        }
        ++counter;
        if (N < 0 && counter >= abs(N)) break;
    }
    free(rates); free(sidx); free(sizes);
}

/* MPIReduceAvgErr MCIntegrator.cpp:21-34 with the deterministic sum */
static void reduce_avg_err(int nd, int64_t R, const double *wa, const double *we, double *avg, double *err)
{
    double *tmp = malloc(sizeof(double) * (size_t)R);
    for (int j = 0; j < nd; ++j) {
        for (int64_t r = 0; r < R; ++r) tmp[r] = wa[r * nd + j];
        avg[j] = orc_det_sum(tmp, R) / (double)R;
        for (int64_t r = 0; r < R; ++r) tmp[r] = we[r * nd + j] * we[r * nd + j];
        err[j] = sqrt(orc_det_sum(tmp, R)) / (double)R;
    }
    free(tmp);
}

/* MCI::initialDecorrelation MCIntegrator.cpp:171-247 */
static mci_status decorrelate(orc_ctx *c, int64_t w0, int64_t w1, int64_t nranks)
{
    if (c->ndecorr > 0) {
        for (int64_t k = w0; k < w1; ++k) {
            if (c->prec == MCI_PREC_FP32) TRY(sample_fp32(c, &c->w[k], c->ndecorr, NULL, NULL, 0));
            else sample_plain(c, &c->w[k], c->ndecorr);
        }
        return MCI_OK;
    }
    if (c->ndecorr == 0) return MCI_OK;
    /* equilibration container: clones of flag_equil observables, Full accumulator, Correlated estimator */
    obs_desc eo[MAX_OBS];
    int eq_index[MAX_OBS];
    int neo = 0, nd = 0;
    for (int i = 0; i < c->nobs; ++i) {
        eq_index[i] = -1;
        if (c->obs[i].flag_equil) {
            eq_index[i] = neo;
            eo[neo] = c->obs[i];
            eo[neo].blocksize = 1; eo[neo].nskip = 1; eo[neo].estim = MCI_ESTIM_CORRELATED;
            if (eo[neo].kind == MCI_OBS_DEP_SQUARE) { /* its dependency's index in this container */
                eo[neo].dep = eq_index[c->obs[i].dep];
                if (eo[neo].dep < 0)
                    return fail(MCI_E_INVALID_ARGUMENT, "an equilibration observable depends on observable %d, which is not one", c->obs[i].dep);
            }
            nd += eo[neo].nobs;
            ++neo;
        }
    }
    const int64_t MIN_NMC = (int64_t)fmax(100., sqrt(40000. * c->ndim) / (double)nranks);
    const int64_t R = w1 - w0;
    accu *acc = calloc((size_t)(R * (neo > 0 ? neo : 1)), sizeof(accu));
    for (int64_t k = 0; k < R; ++k)
        for (int i = 0; i < neo; ++i) {
            accu_make(&acc[k * neo + i], &eo[i], c->ndim);
            TRY(accu_allocate(&acc[k * neo + i], MIN_NMC));
        }
    size_t nb = (size_t)(nd > 0 ? nd : 1);
    double *wa = calloc(nb * (size_t)R, sizeof(double)), *we = calloc(nb * (size_t)R, sizeof(double));
    double *oa = calloc(nb, sizeof(double)), *oe = calloc(nb, sizeof(double));
    double *na = calloc(nb, sizeof(double)), *ne = calloc(nb, sizeof(double));
    mci_status st = MCI_OK;

#define EQ_SAMPLE_ESTIMATE(outa, oute) \
    for (int64_t k = 0; k < R && st == MCI_OK; ++k) { \
        st = sample_obs(c, &c->w[w0 + k], MIN_NMC, &acc[k * neo], eo, neo, 0); \
        int off = 0; \
        for (int i = 0; i < neo && st == MCI_OK; ++i) { \
            st = estimate_any(MCI_ESTIM_CORRELATED, acc[k * neo + i].nstore, acc[k * neo + i].nobs, acc[k * neo + i].data, \
                              wa + k * nd + off, we + k * nd + off); \
            off += eo[i].nobs; \
        } \
    } \
    if (st == MCI_OK) { \
        if (nranks > 1) reduce_avg_err(nd, R, wa, we, outa, oute); \
        else { memcpy(outa, wa, sizeof(double) * (size_t)nd); memcpy(oute, we, sizeof(double) * (size_t)nd); } \
    }

    EQ_SAMPLE_ESTIMATE(oa, oe)
    int flag_loop = 1;
    int64_t count = 0;
    while (flag_loop && st == MCI_OK) {
        flag_loop = 0;
        for (int64_t k = 0; k < R; ++k) st = sample_obs(c, &c->w[w0 + k], MIN_NMC, &acc[k * neo], eo, neo, 0);
        if (st != MCI_OK) break;
        count += MIN_NMC;
        if (count >= llabs(c->ndecorr)) { ++c->nwarn; break; }
        /* estimate the samples just taken (the accumulators hold them) */
        for (int64_t k = 0; k < R && st == MCI_OK; ++k) {
            int off = 0;
            for (int i = 0; i < neo && st == MCI_OK; ++i) {
                st = estimate_any(MCI_ESTIM_CORRELATED, acc[k * neo + i].nstore, acc[k * neo + i].nobs, acc[k * neo + i].data,
                                  wa + k * nd + off, we + k * nd + off);
                off += eo[i].nobs;
            }
        }
        if (st != MCI_OK) break;
        if (nranks > 1) reduce_avg_err(nd, R, wa, we, na, ne);
        else { memcpy(na, wa, sizeof(double) * (size_t)nd); memcpy(ne, we, sizeof(double) * (size_t)nd); }
        for (int i = 0; i < nd; ++i)
            if (fabs(oa[i] - na[i]) > 2 * sqrt(oe[i] * oe[i] + ne[i] * ne[i])) { flag_loop = 1; break; }
        memcpy(oa, na, sizeof(double) * (size_t)nd);
        memcpy(oe, ne, sizeof(double) * (size_t)nd);
    }
#undef EQ_SAMPLE_ESTIMATE
    for (int64_t i = 0; i < R * neo; ++i) accu_free(&acc[i]);
    free(acc); free(wa); free(we); free(oa); free(oe); free(na); free(ne);
    return st;
}

mci_status orc_trace_enable(orc_ctx *c, double *xbuf, uint8_t *accbuf, int64_t cap)
{
    c->trace_x = xbuf;
    c->trace_acc = accbuf;
    c->trace_cap = cap;
    c->trace_n = 0;
    c->tracing = (xbuf != NULL && accbuf != NULL && cap > 0);
    return MCI_OK;
}

/* MCI::integrate MCIntegrator.cpp:43-82 + MPIMCI::integrate MPIMCI.cpp:68-93 */
mci_status orc_integrate(orc_ctx *c, int64_t nmc, double *average, double *error, int do_find, int do_decorr)
{
    int finite = dom_volume(c) != 0.;
    if (c->npdf == 0 && !finite)
        return fail(MCI_E_DOMAIN, "[MCI::integrate] Integrating over an infinite domain requires a sampling function.");
    for (int i = 0; i < c->nobs; ++i) { /* what the dependent observables depend on exists */
        const obs_desc *o = &c->obs[i];
        if (o->kind != MCI_OBS_PDF_LOG && o->kind != MCI_OBS_PDF_PROTO) continue;
        const int j = (int)obs_param(o, 0, 0.0);
        if (j >= c->npdf) return fail(MCI_E_INVALID_ARGUMENT, "an observable depends on sampling function %d, but there are %d", j, c->npdf);
        if (o->kind == MCI_OBS_PDF_PROTO && c->pdf[j].nproto != c->ndim)
            return fail(MCI_E_INVALID_ARGUMENT, "the proto-value observable needs a sampling function with one proto value per dimension");
    }
    c->nwarn = 0;
    const int64_t W = c->W;
    const int shared = (c->coupling == MCI_COUPLING_SHARED);
    if (c->npdf > 0) {
        if (shared) {
            if (do_find && c->ntypes > 0) find_mrt2(c, 0, W, W);
            if (do_decorr) TRY(decorrelate(c, 0, W, W));
        } else {
            for (int64_t k = 0; k < W; ++k) {
                if (do_find && c->ntypes > 0) find_mrt2(c, k, k + 1, 1);
                if (do_decorr) TRY(decorrelate(c, k, k + 1, 1));
            }
        }
    }
    if (nmc <= 0) return MCI_OK;
    int nd;
    orc_get_nobsdim(c, &nd);
    free_last_accus(c);
    c->acc_last = calloc((size_t)(W * (c->nobs > 0 ? c->nobs : 1)), sizeof(accu));
    c->acc_last_nobs = c->nobs;
    for (int64_t k = 0; k < W; ++k)
        for (int i = 0; i < c->nobs; ++i) {
            accu_make(&c->acc_last[k * c->nobs + i], &c->obs[i], c->ndim);
            TRY(accu_allocate(&c->acc_last[k * c->nobs + i], nmc));
        }
    free(c->wavg); free(c->werr);
    c->wavg = calloc((size_t)W * (size_t)(nd > 0 ? nd : 1), sizeof(double));
    c->werr = calloc((size_t)W * (size_t)(nd > 0 ? nd : 1), sizeof(double));
    c->nobsdim_last = nd;
    c->trace_n = 0;
    for (int64_t k = 0; k < W; ++k) {
        if (c->prec == MCI_PREC_FP32) TRY(sample_fp32(c, &c->w[k], nmc, &c->acc_last[k * c->nobs], c->obs, c->nobs));
        else TRY(sample_obs(c, &c->w[k], nmc, &c->acc_last[k * c->nobs], c->obs, c->nobs, k == 0));
    }
    /* ObservableContainer::estimate ObservableContainer.cpp:104-111 */
    for (int64_t k = 0; k < W; ++k) {
        int off = 0;
        for (int i = 0; i < c->nobs; ++i) {
            accu *a = &c->acc_last[k * c->nobs + i];
            TRY(estimate_any(c->obs[i].estim, a->nstore, a->nobs, a->data, c->wavg + k * nd + off, c->werr + k * nd + off));
            off += a->nobs;
        }
        if (c->npdf == 0) { /* volume scaling MCIntegrator.cpp:71-77 */
            const double vol = dom_volume(c);
            for (int j = 0; j < nd; ++j) { c->wavg[k * nd + j] = c->wavg[k * nd + j] * vol; c->werr[k * nd + j] = c->werr[k * nd + j] * vol; }
        }
    }
    if (W == 1) {
        memcpy(average, c->wavg, sizeof(double) * (size_t)nd);
        memcpy(error, c->werr, sizeof(double) * (size_t)nd);
    } else {
        reduce_avg_err(nd, W, c->wavg, c->werr, average, error);
    }
    return MCI_OK;
}

mci_status orc_get_walker_results(orc_ctx *c, double *a, double *e, int64_t n)
{
    if (n < c->W * c->nobsdim_last) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    memcpy(a, c->wavg, sizeof(double) * (size_t)(c->W * c->nobsdim_last));
    memcpy(e, c->werr, sizeof(double) * (size_t)(c->W * c->nobsdim_last));
    return MCI_OK;
}

mci_status orc_get_acceptance(orc_ctx *c, int64_t *acc, int64_t *rej, double *rate)
{
    int64_t a = 0, r = 0;
    double *rates = malloc(sizeof(double) * (size_t)c->W);
    for (int64_t k = 0; k < c->W; ++k) { a += c->w[k].acc; r += c->w[k].rej; rates[k] = walker_rate(&c->w[k]); }
    *acc = a;
    *rej = r;
    *rate = (c->W == 1) ? rates[0] : orc_det_sum(rates, c->W) / (double)c->W;
    free(rates);
    return MCI_OK;
}

mci_status orc_get_walker_acceptance(orc_ctx *c, int64_t *acc, int64_t *rej, int64_t n)
{
    if (n < c->W) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    for (int64_t k = 0; k < c->W; ++k) { acc[k] = c->w[k].acc; rej[k] = c->w[k].rej; }
    return MCI_OK;
}

mci_status orc_get_warnings(orc_ctx *c, int64_t *n) { *n = c->nwarn; return MCI_OK; }

mci_status orc_get_shadow_data(orc_ctx *c, int obs, int64_t walker, double *data, int64_t n, int64_t *nstore)
{
    if (!c->acc_last || obs < 0 || obs >= c->acc_last_nobs || walker < 0 || walker >= c->W)
        return fail(MCI_E_INVALID_ARGUMENT, "no such accumulator");
    accu *a = &c->acc_last[walker * c->acc_last_nobs + obs];
    if (!a->shadow) return fail(MCI_E_RUNTIME, "no FP64 shadow: the last integrate did not run the FP32 chain with the shadow on");
    *nstore = a->nstore;
    int64_t m = a->nstore * a->nobs;
    if (n < m) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    memcpy(data, a->shadow, sizeof(double) * (size_t)m);
    return MCI_OK;
}

mci_status orc_get_accu_data(orc_ctx *c, int obs, int64_t walker, double *data, int64_t n, int64_t *nstore)
{
    if (!c->acc_last || obs < 0 || obs >= c->acc_last_nobs || walker < 0 || walker >= c->W)
        return fail(MCI_E_INVALID_ARGUMENT, "no such accumulator");
    accu *a = &c->acc_last[walker * c->acc_last_nobs + obs];
    *nstore = a->nstore;
    int64_t m = a->nstore * a->nobs;
    if (n < m) return fail(MCI_E_INVALID_ARGUMENT, "output too small");
    memcpy(data, a->data, sizeof(double) * (size_t)m);
    return MCI_OK;
}
