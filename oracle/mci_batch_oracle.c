/*
 * mci_batch_oracle.c — CPU oracle of the batched Gaussian-mixture integrals (TEST
 * INFRASTRUCTURE ONLY; used by tests/ as the checker, never by the product).
 *
 * The batched path has no single reference counterpart: per integral it is what one
 * reference MCI object computes with a mixture sampling function, the |x|^2 observable,
 * single-coordinate uniform moves (include/mci/SRRDVecMove.hpp:75-96) and step calibration
 * (src/MCIntegrator.cpp:87-168, 64 pooled ranks as under USE_MPI, :107,131-138), combined over
 * walkers (src/MPIMCI.cpp:85-92, or the standard error of the walker means).  This file states
 * that algorithm in plain C, step for step, in FP32 where the device computes in FP32 and FP64
 * where it reduces in FP64, so a correct device implementation reproduces it bit for bit:
 *  - FP32 exp/log are the routines written out below (Cody–Waite + Taylor; atanh series),
 *    pinned against libm's double exp/log by tests/test_oracle_batch.py (<= 1 ulp);
 *  - fmaf, rintf, sqrtf and FP32 division are IEEE correctly rounded here and on the GPU;
 *  - the per-walker MJBlocker is orc_estimate (pinned by the reference's golden vectors);
 *  - sums over walkers use orc_det_sum.
 * Build with -ffp-contract=off (oracle/Makefile).
 */
#include "mci_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define BK 4 /* components per integral, at most */
#define BW 64
#define CALR 64 /* calibration lanes (pooled ranks) per integral */

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float orc_bexp(float x)
{
    if (!(x >= -87.0f)) return 0.0f;
    if (x > 88.0f) return INFINITY;
    const float n = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-n, 0x1.62e43p-1f, x);
    r = fmaf(-n, -0x1.05c61p-29f, r);
    float p = 0x1.a01a02p-13f;
    p = fmaf(p, r, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0x1p-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return p * u2f((uint32_t)((int)n + 127) << 23);
}

float orc_blog(float x)
{
    const uint32_t b = f2u(x);
    int e = (int)(b >> 23) - 127;
    float m = u2f((b & 0x7fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float z = s * s;
    float p = 0x1.c71c72p-4f;
    p = fmaf(p, z, 0x1.24924ap-3f);
    p = fmaf(p, z, 0x1.99999ap-3f);
    p = fmaf(p, z, 0x1.555556p-2f);
    const float ls = s + s;
    const float lm = fmaf(ls * z, p, ls);
    const float fe = (float)e;
    return fmaf(fe, 0x1.62e43p-1f, fmaf(fe, -0x1.05c61p-29f, lm));
}

static float sym24(uint32_t w) { return (float)(2 * (int32_t)(w >> 8) + 1 - 16777216) * 0x1p-24f; }

static void kahanf(float *s, float *c, float y)
{
    const float yy = y - *c;
    const float tt = *s + yy;
    *c = (tt - *s) - yy;
    *s = tt;
}

typedef struct {
    int nint;
    const int *d, *K;
    float *lw, *h, *sig, *cw, *mu;
    int64_t *moff;
    double *step0;
    uint32_t key[2];
} bparams;

typedef struct {
    int d, K;
    uint32_t thr;
    float lw[BK], h[BK];
    const float *mu;
    float r[BK], m, S;
    double q;
    float step;
    uint32_t id;
    uint64_t t;
    float x[MCI_BATCH_MAX_DIM];
} bwalker;

static void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, const uint32_t key[2], uint32_t o[4])
{
    const uint32_t ctr[4] = {c0, c1, c2, c3};
    orc_philox4x32_10(ctr, key, o);
}

static void bw_load(bwalker *L, const bparams *P, int i, uint32_t id)
{
    L->d = P->d[i];
    L->K = P->K[i];
    L->thr = (uint32_t)(0u - (uint32_t)L->d) % (uint32_t)L->d;
    for (int k = 0; k < BK; ++k) {
        L->lw[k] = P->lw[i * BK + k];
        L->h[k] = P->h[i * BK + k];
        L->r[k] = 0.f;
    }
    L->mu = P->mu + P->moff[i];
    L->id = id;
    L->t = 0;
    L->m = 0.f;
    L->S = 1.f;
    L->q = 0.;
}

/* exact draw: component by cumulative weight, coordinates by Marsaglia's polar method */
static void bw_draw(bwalker *L, const bparams *P, int i)
{
    uint32_t o[4];
    philox(0x40000000u, L->id, 0u, 0u, P->key, o);
    const float uc = (float)(o[0] >> 8) * 0x1p-24f;
    int kc = 0;
    for (int k = 0; k < BK - 1; ++k) kc += (k < L->K - 1 && uc >= P->cw[i * BK + k]) ? 1 : 0;
    const float sig = P->sig[i * BK + kc];
    const float *muk = L->mu + (int64_t)kc * L->d;
    uint32_t call = 1;
    int j = 0;
    while (j < L->d) {
        philox(0x40000000u + call, L->id, 0u, 0u, P->key, o);
        ++call;
        for (int a = 0; a < 2; ++a) {
            if (j < L->d) {
                const float v1 = sym24(o[2 * a]);
                const float v2 = sym24(o[2 * a + 1]);
                const float s = fmaf(v1, v1, v2 * v2);
                if (s < 1.0f) {
                    const float f = sqrtf(-2.0f * orc_blog(s) / s);
                    L->x[j] = fmaf(sig, v1 * f, muk[j]);
                    ++j;
                    if (j < L->d) {
                        L->x[j] = fmaf(sig, v2 * f, muk[j]);
                        ++j;
                    }
                }
            }
        }
    }
}

static void bw_refresh(bwalker *L)
{
    double q = 0.;
    for (int j = 0; j < L->d; ++j) q = fma((double)L->x[j], (double)L->x[j], q);
    L->q = q;
    float e[BK];
    float m = -INFINITY;
    for (int k = 0; k < L->K; ++k) {
        const float *mk = L->mu + (int64_t)k * L->d;
        float rk = 0.f;
        for (int j = 0; j < L->d; ++j) {
            const float dd = L->x[j] - mk[j];
            rk = fmaf(dd, dd, rk);
        }
        L->r[k] = rk;
        e[k] = fmaf(-L->h[k], rk, L->lw[k]);
        m = fmaxf(m, e[k]);
    }
    float S = 0.f;
    for (int k = 0; k < L->K; ++k) S = S + orc_bexp(e[k] - m);
    L->m = m;
    L->S = S;
}

/* one Metropolis step: SRRDVecMove (veclen 1) proposal, u <= p(x')/p(x) as u S < exp(m'-m) S' */
static int bw_step(bwalker *L, const bparams *P)
{
    uint32_t o[4];
    const uint32_t tl = (uint32_t)L->t, th = (uint32_t)(L->t >> 32);
    philox(0u, L->id, tl, th, P->key, o);
    uint64_t prod = (uint64_t)o[0] * (uint32_t)L->d;
    uint32_t low = (uint32_t)prod, kx = 0;
    while (low < L->thr) { /* Lemire rejection (uniform_int_dist.h) on extra words */
        const uint32_t k = 0x80000000u + (kx++);
        uint32_t e[4];
        philox(k >> 2, L->id, tl, th, P->key, e);
        prod = (uint64_t)e[k & 3u] * (uint32_t)L->d;
        low = (uint32_t)prod;
    }
    const int j = (int)(prod >> 32);
    const float xo = L->x[j];
    const float xn = fmaf(L->step, sym24(o[1]), xo);
    const float ua = (float)(o[2] >> 8) * 0x1p-24f;
    const float dx = xn - xo, sx = xn + xo;
    float rn[BK], en[BK];
    float mn = -INFINITY;
    for (int k = 0; k < L->K; ++k) {
        const float mk = L->mu[(int64_t)k * L->d + j];
        rn[k] = fmaf(dx, sx - (mk + mk), L->r[k]);
        en[k] = fmaf(-L->h[k], rn[k], L->lw[k]);
        mn = fmaxf(mn, en[k]);
    }
    float Sn = 0.f;
    for (int k = 0; k < L->K; ++k) Sn = Sn + orc_bexp(en[k] - mn);
    const float E = orc_bexp(mn - L->m);
    const int acc = ua * L->S < E * Sn;
    if (acc) {
        L->x[j] = xn;
        for (int k = 0; k < L->K; ++k) L->r[k] = rn[k];
        L->m = mn;
        L->S = Sn;
        const double xn64 = (double)xn, xo64 = (double)xo; /* |x|^2 update in FP64 */
        L->q = L->q + fma(xn64, xn64, -(xo64 * xo64));
    }
    ++L->t;
    return acc;
}

/* FP32 parameters of each integral from the double inputs */
static void derive(int nint, const int *d, const int *K, const double *w, const double *s, const double *mu, bparams *P)
{
    int64_t nc = 0, nm = 0;
    for (int i = 0; i < nint; ++i) { nc += K[i]; nm += (int64_t)K[i] * d[i]; }
    P->lw = calloc((size_t)nint * BK, sizeof(float));
    P->h = calloc((size_t)nint * BK, sizeof(float));
    P->sig = calloc((size_t)nint * BK, sizeof(float));
    P->cw = calloc((size_t)nint * BK, sizeof(float));
    P->mu = malloc(sizeof(float) * (size_t)(nm > 0 ? nm : 1));
    P->moff = malloc(sizeof(int64_t) * (size_t)nint);
    P->step0 = malloc(sizeof(double) * (size_t)nint);
    for (int64_t q = 0; q < nm; ++q) P->mu[q] = (float)mu[q];
    int64_t co = 0, mo = 0;
    for (int i = 0; i < nint; ++i) {
        P->moff[i] = mo;
        double wsum = 0.;
        for (int k = 0; k < K[i]; ++k) wsum = wsum + w[co + k];
        double cum = 0., st = 0.;
        for (int k = 0; k < BK; ++k) P->cw[i * BK + k] = 1.f;
        for (int k = 0; k < K[i]; ++k) {
            const double wn = w[co + k] / wsum, sk = s[co + k];
            P->lw[i * BK + k] = (float)(log(wn) - 0.5 * (double)d[i] * log(2.0 * M_PI * sk * sk));
            P->h[i * BK + k] = (float)(1.0 / (2.0 * sk * sk));
            P->sig[i * BK + k] = (float)sk;
            cum = cum + wn;
            P->cw[i * BK + k] = (k == K[i] - 1) ? 1.0f : (float)cum;
            st = st + wn * sk;
        }
        P->step0[i] = 2.93 * st; /* half acceptance on a Gaussian of width st */
        co += K[i];
        mo += (int64_t)K[i] * d[i];
    }
}

/* findMRT2Step over CALR pooled lanes (MCIntegrator.cpp:87-168 with R = CALR ranks) */
static void calibrate(const bparams *P, int i, int nfind, double target, double *step_out, double *rate_out)
{
    double step = P->step0[i], rate = 0.;
    const int N = nfind;
    if (N != 0) {
        bwalker *L = malloc(sizeof(bwalker) * CALR);
        for (int l = 0; l < CALR; ++l) {
            bw_load(&L[l], P, i, 0x80000000u | (uint32_t)(i * CALR + l));
            bw_draw(&L[l], P, i);
            bw_refresh(&L[l]);
        }
        const int64_t min_stat = (int64_t)fmax(100., sqrt(40000. * P->d[i]) / (double)CALR);
        int cons = 0, counter = 0;
        double v[CALR];
        while ((N < 0 && cons < 5) || counter < N) {
            for (int l = 0; l < CALR; ++l) {
                L[l].step = (float)step;
                int64_t acc = 0;
                for (int64_t s = 0; s < min_stat; ++s) acc += bw_step(&L[l], P);
                v[l] = (acc > 0) ? (double)acc / ((double)acc + (double)(min_stat - acc)) : 0.;
            }
            rate = orc_det_sum(v, CALR) / (double)CALR;
            if (fabs(rate - target) < 0.05) ++cons;
            else cons = 0;
            const double fact = fmin(2., fmax(0.5, rate / target));
            step = step * fact;
            if (step > (double)FLT_MAX) step = (double)FLT_MAX;
            if (step < (double)FLT_MIN) step = (double)FLT_MIN;
            ++counter;
            if (N < 0 && counter >= -N) break;
        }
        free(L);
    }
    *step_out = step;
    *rate_out = rate;
}

static const int64_t *g_T;
static const int *g_d, *g_K;
static int cmp_order(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    if (g_T[x] != g_T[y]) return g_T[x] > g_T[y] ? -1 : 1;
    if (g_d[x] != g_d[y]) return g_d[x] > g_d[y] ? -1 : 1;
    if (g_K[x] != g_K[y]) return g_K[x] > g_K[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

mci_status orc_batch_walker_order(int nint, const int *ndim, const int *ncomp, const int64_t *nmc, const int64_t *T,
                                  int64_t *walkers, int64_t *woff, int64_t *nwalkers)
{
    int *ord = malloc(sizeof(int) * (size_t)nint);
    for (int i = 0; i < nint; ++i) ord[i] = i;
    g_T = T; g_d = ndim; g_K = ncomp;
    qsort(ord, (size_t)nint, sizeof(int), cmp_order); /* total order: same as a stable sort */
    int64_t off = 0;
    for (int q = 0; q < nint; ++q) {
        const int i = ord[q];
        walkers[i] = (nmc[i] + T[i] - 1) / T[i];
        woff[i] = off;
        off += walkers[i];
    }
    *nwalkers = off;
    free(ord);
    return MCI_OK;
}

mci_status orc_batch_run(int nint, const int *ndim, const int64_t *nmc, const int *ncomp, const double *weights,
                         const double *widths, const double *means, uint64_t seed, int nblocks, int nfind, double target,
                         int64_t ndecorr, int mode, const int64_t *T, double *avg, double *err, double *steps,
                         double *calib_rate, double *accrate, double *wavg, double *werr, double *bmeans,
                         double *bmeans_fp64)
{
    bparams P;
    memset(&P, 0, sizeof P);
    P.nint = nint;
    P.d = ndim;
    P.K = ncomp;
    P.key[0] = (uint32_t)seed;
    P.key[1] = (uint32_t)(seed >> 32);
    derive(nint, ndim, ncomp, weights, widths, means, &P);
    int64_t *R = malloc(sizeof(int64_t) * (size_t)nint), *woff = malloc(sizeof(int64_t) * (size_t)nint), W = 0;
    orc_batch_walker_order(nint, ndim, ncomp, nmc, T, R, woff, &W);
    double *step = malloc(sizeof(double) * (size_t)nint);
    for (int i = 0; i < nint; ++i) {
        double r;
        calibrate(&P, i, nfind, target, &step[i], &r);
        if (steps) steps[i] = step[i];
        if (calib_rate) calib_rate[i] = r;
    }
    double *bm = malloc(sizeof(double) * (size_t)(nblocks * W));
    double *wa = malloc(sizeof(double) * (size_t)W), *we = malloc(sizeof(double) * (size_t)W);
    int32_t *acc = malloc(sizeof(int32_t) * (size_t)W);
    bwalker *L = malloc(sizeof(bwalker));
    double *col = malloc(sizeof(double) * (size_t)nblocks);
    for (int i = 0; i < nint; ++i) {
        const int64_t B = T[i] / nblocks;
        for (int64_t r = 0; r < R[i]; ++r) {
            const int64_t g = woff[i] + r;
            bw_load(L, &P, i, (uint32_t)g);
            L->step = (float)step[i];
            bw_draw(L, &P, i);
            bw_refresh(L);
            for (int64_t s = 0; s < ndecorr; ++s) bw_step(L, &P);
            float Sb = 0.f, Cb = 0.f;
            double S64 = 0.;
            int64_t jb = 0, b = 0;
            int32_t na = 0;
            for (int64_t s = 0; s < T[i]; ++s) {
                na += bw_step(L, &P);
                kahanf(&Sb, &Cb, (float)L->q);
                if (bmeans_fp64) { /* the same samples summed in FP64 from the FP32 coordinates */
                    double q64 = 0.;
                    for (int j = 0; j < L->d; ++j) q64 += (double)L->x[j] * (double)L->x[j];
                    S64 += q64;
                }
                if (++jb == B) {
                    bm[b * W + g] = ((double)Sb - (double)Cb) * (1. / (double)B);
                    if (bmeans_fp64) bmeans_fp64[b * W + g] = S64 / (double)B;
                    Sb = 0.f; Cb = 0.f; S64 = 0.; jb = 0; ++b;
                    bw_refresh(L);
                }
            }
            acc[g] = na;
            for (int b2 = 0; b2 < nblocks; ++b2) col[b2] = bm[b2 * W + g];
            orc_estimate(MCI_ESTIM_MJBLOCKER, nblocks, 1, col, &wa[g], &we[g]);
        }
    }
    double *tmp = malloc(sizeof(double) * (size_t)W);
    for (int i = 0; i < nint; ++i) {
        const int64_t o = woff[i], n = R[i];
        const double a = orc_det_sum(wa + o, n) / (double)n;
        double e;
        if (mode == MCI_BATCH_ERR_SPREAD && n >= 2) {
            for (int64_t r = 0; r < n; ++r) { const double dv = wa[o + r] - a; tmp[r] = dv * dv; }
            e = sqrt(orc_det_sum(tmp, n) / ((double)n * (double)(n - 1)));
        } else {
            for (int64_t r = 0; r < n; ++r) tmp[r] = we[o + r] * we[o + r];
            e = sqrt(orc_det_sum(tmp, n)) / (double)n;
        }
        int64_t na = 0;
        for (int64_t r = 0; r < n; ++r) na += acc[o + r];
        avg[i] = a;
        err[i] = e;
        if (accrate) accrate[i] = (double)na / ((double)n * (double)T[i]);
    }
    if (wavg) memcpy(wavg, wa, sizeof(double) * (size_t)W);
    if (werr) memcpy(werr, we, sizeof(double) * (size_t)W);
    if (bmeans) memcpy(bmeans, bm, sizeof(double) * (size_t)(nblocks * W));
    free(tmp); free(col); free(L); free(acc); free(we); free(wa); free(bm); free(step); free(woff); free(R);
    free(P.lw); free(P.h); free(P.sig); free(P.cw); free(P.mu); free(P.moff); free(P.step0);
    return MCI_OK;
}
