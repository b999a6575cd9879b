// mci_api.cpp — C-ABI of the MI355X Metropolis integrator (include/mci_amd.h) and the host
// orchestration of MCI::integrate (src/MCIntegrator.cpp:43-82): calibration, decorrelation,
// sampling, estimation and the cross-walker combine run as HIP kernels on the context's
// stream; the host only steers the loops whose exit depends on a statistic.
#include "mci_kernels.hpp"
#include "mci_fast.hpp"
#include "mci_fast_plan.hpp"
#include "mci_host.hpp"
#include "mci_rtc.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

using namespace mcid;

using mcih::check_estimator;
using mcih::DBuf;
using mcih::det_sum;
using mcih::fail;
using mcih::g_err;

namespace
{
#define CHECKCTX(c)                                                                                      \
    do {                                                                                                 \
        if (!(c)) { return fail(MCI_E_INVALID_ARGUMENT, "null context"); }                               \
    } while (0)

uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// seed of walker w when only MCI::setSeed(seed) was given: walker 0 keeps it (so one walker
// reproduces one reference MCI seeded the same way), the others get splitmix64 draws
uint64_t walker_seed(uint64_t seed, int64_t w)
{
    if (w == 0) { return seed; }
    uint64_t s = (seed ^ 0x6A09E667F3BCC909ULL) + (uint64_t)w * 0x9E3779B97F4A7C15ULL;
    return splitmix64(s);
}

int obs_nobs(int kind, int ndim)
{
    switch (kind) {
    case MCI_OBS_X2: case MCI_OBS_X: case MCI_OBS_PDF_PROTO: return ndim;
    case MCI_OBS_PDF_LOG: return 1;
    case MCI_OBS_X2_SUM: case MCI_OBS_X_SUM: case MCI_OBS_CONST: case MCI_OBS_EXP_X2: case MCI_OBS_INDICATOR:
    case MCI_OBS_GAUSS_X2: return 1;
    }
    return -1;
}

bool is_pow2(int64_t n) { return n > 0 && (((uint64_t)n & ((uint64_t)n - 1)) == 0); }

struct PdfHost
{
    int kind = 0, nproto = 0, K = 0;
    std::vector<double> params;
    DBuf<double> dev; // a[ndim] | gc[K], gb[K], gmu[K*ndim] | a device functor's parameters
    std::string fname, fsrc; // MCI_PDF_DEVICE: the functor's name and source
};

struct ObsHost
{
    int kind = 0, nobs = 0, blocksize = 1, nskip = 1, equil = 0, estim = 0;
    std::vector<double> params;
    DBuf<double> dev;        // MCI_OBS_DEVICE: its parameters on the device
    std::string fname, fsrc; // MCI_OBS_DEVICE: the functor's name and source
};

const char *kVersion = "mcintegratorplusplus_amd 0.1 (gfx950)";
} // namespace

struct mci_ctx
{
    int ndim = 0;
    int64_t W = 0;
    int rng = MCI_RNG_MT19937_64, prec = MCI_PREC_FP64, coupling = MCI_COUPLING_INDEPENDENT;
    int kernel_path = MCI_KERNELS_AUTO;
    uint64_t seed = 5489u;
    std::vector<uint64_t> wseeds;
    bool reseed = true;
    double target = 0.5;
    int nfind = -50;
    int64_t ndecorr = -10000;
    bool periodic = false;
    std::vector<double> lb, ub;
    int movetype = MCI_MOVE_ALL, srrd = MCI_SRRD_UNIFORM, veclen = 0, nvecs = 0, ntypes = 1;
    std::vector<int> typeEnds;
    int ms_on = 0, ms_nsteps = 0;     // MultiStepMove around the move above
    int lm_K = 0;                     // the independence half of a mixed move: mixture components (0: off)
    double lm_pmix = 0.;
    DBuf<double> lm;                  // mu [K][ndim], 1 / sigma [K][ndim], sigma [K][ndim], lc [K], cw [K]
    std::string obsfile, wlkfile;     // file output of the main sampling (empty: off)
    mci_callback_fn cb = nullptr;     // MCI::setCallback: replayed from walker 0's record
    void *cb_user = nullptr;
    int64_t cb_cap = 0;
    int64_t obsfreq = 0, wlkfreq = 0;
    std::vector<PdfHost> ms_pdfs;     // its own sampling functions
    std::vector<PdfHost> pdfs;
    std::vector<ObsHost> obs;
    RtcKernels rtc;                   // the step kernels built with the device functors (none: null)
    std::string rtc_text;             // their program text
    std::string mv_name, mv_src;      // MCI_MOVE_DEVICE: the caller's trial move
    int mv_nproto = 0;
    double mv_rate = 1.;              // its getChangeRate
    DBuf<double> mv_par, mv_pold, mv_pnew;

    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};

    DBuf<double> x, xn, pold, pnew, steps, lbd, ubd, obsval, dimsizes, wavg, werr, chunks_a, chunks_e, res, rates, fact;
    DBuf<double> oa, oe;  // decorrelation: previous estimates (independent walkers)
    DBuf<CalibCtrl> ctrl;
    int calib_hint = 0;    // iterations the last shared calibration took: the next one's first batch
    bool walker_results_host = false; // h_wavg/h_werr hold the device's per-walker results
    DBuf<uint64_t> mt, ctr, seeds;
    DBuf<int32_t> mtp;
    // the walkers' mt19937_64 next indices may differ (per-walker loops, generic-kernel launches,
    // a restored state): realigned before the next specialised launch of all walkers
    mutable bool mt_diverged = false;
    DBuf<double> nsv;     // [2][W] saved normals of the move's distribution
    DBuf<double> ms_pold, ms_pnew, ms_x0; // MultiStepMove: proto values of its own sampling functions, start position
    DBuf<double> rec_obs, rec_x;          // file output: recorded observable values and positions
    DBuf<double> cb_buf;                  // callback record [cb_cap][ndim + 2]
    DBuf<int64_t> cb_n;
    DBuf<int32_t> nsf;    // [W] their availability bits
    DBuf<int64_t> acc, rej;
    DBuf<uint8_t> active;
    DBuf<int32_t> cal_cons, cal_count, cal_any; // independent calibration on the specialised kernels
    // the walkers still calibrating, compacted: their sampling columns and the list of who they are
    DBuf<double> cmp_x, cmp_steps;
    DBuf<int64_t> cmp_acc, cmp_rej;
    DBuf<uint64_t> cmp_mt, cmp_ctr;
    DBuf<int32_t> cmp_mtp, cmp_list, cmp_count;
    int calib_indep_hint = 0; // iterations the last such calibration took: the next one's first batch
    std::vector<DBuf<double>> accdata, scratch;
    std::vector<double> h_wavg, h_werr;
    int nobsdim_last = 0;
    std::vector<int64_t> nstore_last; // rows of each observable's data after the last integrate
    double timings[5] = {0, 0, 0, 0, 0};
    FastState fast;
};

namespace
{
int nproto_total(const mci_ctx *c)
{
    int n = 0;
    for (auto &p : c->pdfs) { n += p.nproto; }
    return n;
}

int nobsdim(const mci_ctx *c)
{
    int n = 0;
    for (auto &o : c->obs) { n += o.nobs; }
    return n;
}

mci_status set_current(mci_ctx *c)
{
    HIPCHK(hipSetDevice(c->device));
    return MCI_OK;
}

mci_status fill_rows(mci_ctx *c, double *dev, int row0, int nrows, double v)
{
    std::vector<double> h((size_t)c->W * nrows, v);
    HIPCHK(hipMemcpyAsync(dev + (size_t)row0 * c->W, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status ensure_protos(mci_ctx *c)
{
    const size_t n = (size_t)std::max(nproto_total(c), 1) * c->W;
    HIPCHK(c->pold.ensure(n));
    HIPCHK(c->pnew.ensure(n));
    return MCI_OK;
}

mci_status ensure_rng(mci_ctx *c)
{
    if (!c->reseed) { return MCI_OK; }
    if (c->rng == MCI_RNG_MT19937_64) {
        std::vector<uint64_t> s(c->W);
        for (int64_t w = 0; w < c->W; ++w) { s[w] = c->wseeds.empty() ? walker_seed(c->seed, w) : c->wseeds[w]; }
        HIPCHK(c->seeds.ensure(c->W));
        HIPCHK(c->mt.ensure((size_t)MT_N * c->W));
        HIPCHK(c->mtp.ensure(c->W));
        HIPCHK(hipMemcpyAsync(c->seeds.p, s.data(), s.size() * sizeof(uint64_t), hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_seed_mt(c->mt.p, c->mtp.p, c->seeds.p, c->W, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->mt_diverged = false;
    }
    else {
        HIPCHK(c->ctr.ensure(c->W));
        HIPCHK(hipMemsetAsync(c->ctr.p, 0, c->W * sizeof(uint64_t), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->reseed = false;
    return MCI_OK;
}

// Chain descriptor of the context for the given observable list (main or equilibration)
Chain make_chain(mci_ctx *c, const std::vector<ObsDev> *obsd)
{
    Chain ch;
    std::memset(&ch, 0, sizeof ch);
    ch.ndim = c->ndim;
    ch.W = c->W;
    ch.rng = c->rng;
    ch.key0 = (uint32_t)c->seed;
    ch.key1 = (uint32_t)(c->seed >> 32);
    ch.mt = c->mt.p;
    ch.mtp = c->mtp.p;
    ch.ctr = c->ctr.p;
    ch.xold = c->x.p;
    ch.xnew = c->xn.p;
    ch.pold = c->pold.p;
    ch.pnew = c->pnew.p;
    ch.nproto = nproto_total(c);
    ch.steps = c->steps.p;
    ch.ntypes = c->ntypes;
    for (int t = 0; t < c->ntypes && t < MAX_TYPES; ++t) { ch.typeEnds[t] = c->typeEnds[t]; }
    ch.movetype = c->movetype;
    ch.veclen = c->veclen;
    ch.nvecs = c->nvecs;
    ch.ibits = shared_index_bits(c->rng == MCI_RNG_PHILOX, c->movetype == MCI_MOVE_VEC, c->ms_on != 0, c->nvecs);
    ch.srrd = c->srrd;
    ch.nsv = c->nsv.p;
    ch.nsf = c->nsf.p;
    if (c->cb) {
        ch.cb_buf = c->cb_buf.p;
        ch.cb_n = c->cb_n.p;
        ch.cb_cap = c->cb_cap;
    }
    ch.ms_on = c->ms_on;
    ch.ms_nsteps = c->ms_nsteps;
    ch.lm_on = c->lm_K > 0 && c->lm_pmix > 0. ? 1 : 0;
    ch.lm_K = c->lm_K;
    ch.lm_pmix = c->lm_pmix;
    if (c->lm_K > 0) {
        const size_t KD = (size_t)c->lm_K * c->ndim;
        ch.lm_mu = c->lm.p;
        ch.lm_isig = c->lm.p + KD;
        ch.lm_sig = c->lm.p + 2 * KD;
        ch.lm_lc = c->lm.p + 3 * KD;
        ch.lm_cw = c->lm.p + 3 * KD + c->lm_K;
    }
    ch.ms_npdf = (int)c->ms_pdfs.size();
    ch.ms_pold = c->ms_pold.p;
    ch.mv_par = c->mv_par.p;
    ch.mv_nproto = c->movetype == MCI_MOVE_DEVICE ? c->mv_nproto : 0;
    ch.mv_pold = c->mv_pold.p;
    ch.mv_pnew = c->mv_pnew.p;
    ch.ms_pnew = c->ms_pnew.p;
    ch.ms_x0 = c->ms_x0.p;
    {
        int moff = 0;
        for (size_t i = 0; i < c->ms_pdfs.size(); ++i) {
            const PdfHost &p = c->ms_pdfs[i];
            PdfDev &d = ch.ms_pdf[i];
            d.kind = p.kind;
            d.nproto = p.nproto;
            d.off = moff;
            d.K = p.K;
            d.uid = 0;
            moff += p.nproto;
            if (p.kind == MCI_PDF_GMM_ISO) {
                d.a = nullptr;
                d.gc = p.dev.p;
                d.gb = p.dev.p + p.K;
                d.gmu = p.dev.p + 2 * p.K;
            }
            else { d.a = p.params.empty() ? nullptr : p.dev.p; }
        }
        ch.ms_nproto = moff;
    }
    ch.periodic = c->periodic ? 1 : 0;
    ch.lb = c->lbd.p;
    ch.ub = c->ubd.p;
    ch.npdf = (int)c->pdfs.size();
    int off = 0, uid = 0;
    for (size_t i = 0; i < c->pdfs.size(); ++i) {
        const PdfHost &p = c->pdfs[i];
        PdfDev &d = ch.pdf[i];
        d.kind = p.kind;
        d.nproto = p.nproto;
        d.off = off;
        d.K = p.K;
        d.uid = p.kind == MCI_PDF_DEVICE ? uid++ : 0;
        off += p.nproto;
        if (p.kind == MCI_PDF_GMM_ISO) {
            d.a = nullptr;
            d.gc = p.dev.p;
            d.gb = p.dev.p + p.K;
            d.gmu = p.dev.p + 2 * p.K;
        }
        else { d.a = p.params.empty() ? nullptr : p.dev.p; }
    }
    ch.obsval = c->obsval.p;
    ch.acc = c->acc.p;
    ch.rej = c->rej.p;
    if (obsd) {
        ch.nobs = (int)obsd->size();
        for (size_t i = 0; i < obsd->size(); ++i) { ch.obs[i] = (*obsd)[i]; }
    }
    return ch;
}

// a device observable's index among the context's device observables (the module's dispatch)
int obs_uid(const mci_ctx *c, size_t i)
{
    int u = 0;
    for (size_t j = 0; j < i; ++j) { u += c->obs[j].kind == MCI_OBS_DEVICE ? 1 : 0; }
    return c->obs[i].kind == MCI_OBS_DEVICE ? u : 0;
}

// the step kernels for the context's device functors: built (or taken from the process's cache)
// when the set of functors changed; none: the library's own kernels
mci_status ensure_rtc(mci_ctx *c)
{
    std::vector<RtcFunctor> fp, fo;
    for (const PdfHost &p : c->pdfs) {
        if (p.kind == MCI_PDF_DEVICE) { fp.push_back({p.fname, p.fsrc}); }
    }
    for (const ObsHost &o : c->obs) {
        if (o.kind == MCI_OBS_DEVICE) { fo.push_back({o.fname, o.fsrc}); }
    }
    const bool dmove = c->movetype == MCI_MOVE_DEVICE;
    if (fp.empty() && fo.empty() && !dmove) {
        c->rtc = RtcKernels{};
        c->rtc_text.clear();
        return MCI_OK;
    }
    const RtcFunctor mv{c->mv_name, c->mv_src};
    std::string text, msg;
    if (!rtc_program(fp, fo, dmove ? &mv : nullptr, text, msg)) { return fail(MCI_E_INVALID_ARGUMENT, "%s", msg.c_str()); }
    if (text == c->rtc_text && c->rtc.sample) { return MCI_OK; }
    RtcKernels k;
    std::string log;
    const int st = rtc_get(text, k, log);
    if (st != MCI_OK) { return fail(st, "building the device functors failed:\n%s", log.c_str()); }
    c->rtc = k;
    c->rtc_text = text;
    return MCI_OK;
}

hipError_t sample_launch(const mci_ctx *c, const Chain &ch, int64_t nsteps, int acc, const uint8_t *active, hipStream_t s,
                         const CalibCtrl *skip = nullptr)
{
    c->mt_diverged = true;
    if (c->rtc.sample) { return rtc_launch_sample(c->rtc, ch, nsteps, acc, active, s, skip); }
    return launch_sample(ch, nsteps, acc, active, s, skip);
}

hipError_t calibrate_indep_launch(const mci_ctx *c, const Chain &ch, int nfind, double target, int64_t min_stat, const double *dimSizes,
                                  hipStream_t s)
{
    c->mt_diverged = true;
    if (c->rtc.calibrate) { return rtc_launch_calibrate_indep(c->rtc, ch, nfind, target, min_stat, dimSizes, s); }
    return launch_calibrate_indep(ch, nfind, target, min_stat, dimSizes, s);
}

// Before a specialised launch in which every walker steps: if the walkers' generator next indices
// may differ, rotate each state to next index 0 (the same generator), so that every state access
// of a wave is one row instead of a gather over 64 rows.
mci_status mt_align_if_needed(mci_ctx *c)
{
    if (c->rng != MCI_RNG_MT19937_64 || !c->mt_diverged) { return MCI_OK; }
    HIPCHK(launch_mt_align(c->mt.p, c->mtp.p, c->W, c->stream));
    c->mt_diverged = false;
    return MCI_OK;
}

double obs_p0(const ObsHost &o, int ndim)
{
    switch (o.kind) {
    case MCI_OBS_X2_SUM: return o.params.empty() ? 1.0 : o.params[0];
    case MCI_OBS_CONST: return o.params.empty() ? 1.0 : o.params[0];
    case MCI_OBS_EXP_X2: return o.params.empty() ? (double)ndim : o.params[0];
    case MCI_OBS_INDICATOR: return o.params.empty() ? 0.0 : o.params[0];
    case MCI_OBS_GAUSS_X2: return o.params.empty() ? 1. / sqrt(M_PI * M_PI * M_PI) / 3. : o.params[0];
    case MCI_OBS_PDF_LOG: case MCI_OBS_PDF_PROTO: case MCI_OBS_DEP_SQUARE: return o.params.empty() ? 0.0 : o.params[0];
    }
    return 0.0;
}

int64_t nstore_of(int blocksize, int nskip, int64_t nmc)
{
    const int64_t naccu = 1 + (nmc - 1) / nskip;
    if (blocksize == 0) { return 1; }
    if (blocksize == 1) { return naccu; }
    return naccu / blocksize;
}

bool needs_scratch(int estim, int64_t n) { return estim == MCI_ESTIM_MJBLOCKER || (estim == MCI_ESTIM_CORRELATED && is_pow2(n)); }

float elapsed(mci_ctx *c, int a, int b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) { return 0.f; }
    return ms;
}

// the specialised kernels apply to this chain (observables are checked separately)
bool fast_chain(const mci_ctx *c)
{
    return c->kernel_path == MCI_KERNELS_AUTO && c->srrd == MCI_SRRD_UNIFORM && !c->ms_on && c->lm_K == 0 && !c->cb && fast_eligible(c->prec, c->rng, c->movetype, c->veclen, c->ntypes, c->periodic, (int)c->pdfs.size(),
                         c->pdfs.empty() ? 0 : c->pdfs[0].kind, c->nvecs) &&
           (c->movetype == MCI_MOVE_ALL || !c->pdfs[0].params.empty()) && fast_obs_eligible(nullptr, 0, c->ndim, c->prec, c->movetype, c->rng);
}

// MCI::findMRT2Step for walkers that share the rank-averaged rate (MCIntegrator.cpp:87-168, USE_MPI).
// The loop's bookkeeping runs on the device (k_calib_control); the host enqueues iterations in
// batches and reads the loop flag once per batch; iterations after the exit are no-ops.
mci_status find_shared(mci_ctx *c)
{
    const int N = c->nfind;
    if (N == 0) { return MCI_OK; }
    Chain ch = make_chain(c, nullptr);
    const int64_t min_stat = (int64_t)std::fmax(100., std::sqrt(40000. * c->ndim) / (double)c->W);
    const int64_t nch = (c->W + 255) / 256;
    HIPCHK(c->rates.ensure(c->W));
    HIPCHK(c->chunks_a.ensure(nch));
    HIPCHK(c->ctrl.ensure(1));
    CalibCtrl init;
    std::memset(&init, 0, sizeof init);
    HIPCHK(hipMemcpyAsync(c->ctrl.p, &init, sizeof init, hipMemcpyHostToDevice, c->stream));
    const bool fast = fast_chain(c);
    if (fast) { TRY(mt_align_if_needed(c)); }
    const int maxit = std::abs(N);
    // the first batch is as long as the last calibration of this context took (iterations after
    // the exit are no-ops, so the batch length changes the time, never the result)
    int batch = (c->calib_hint > 0) ? c->calib_hint : 4;
    for (int it = 0; it < maxit;) {
        for (int b = 0; b < batch && it < maxit; ++b, ++it) {
            // the calibration runs the FP64 chain in both precisions (step sizes stay FP64 quantities)
            if (fast) {
                // two launches per iteration: the sampling launch first applies the previous
                // iteration's factor; rates, sums and control are one launch
                TRY(fast_sample(c->fast, ch, MCI_PREC_FP64, min_stat, nullptr, 0, c->stream, c->ctrl.p, c->dimsizes.p));
                HIPCHK(launch_calib_round(c->acc.p, c->rej.p, c->W, c->chunks_a.p, c->target, N, c->ctrl.p, c->stream));
            }
            else {
                HIPCHK(sample_launch(c, ch, min_stat, 0, nullptr, c->stream, c->ctrl.p));
                HIPCHK(launch_walker_rates(c->acc.p, c->rej.p, c->rates.p, c->W, c->stream));
                HIPCHK(launch_chunk_sum(c->rates.p, c->W, 1, 0, c->chunks_a.p, c->stream));
                HIPCHK(launch_calib_control(c->chunks_a.p, c->W, c->target, N, c->ctrl.p, c->stream));
                HIPCHK(launch_scale_steps(ch, c->ctrl.p, c->dimsizes.p, c->stream));
            }
        }
        CalibCtrl h;
        HIPCHK(hipMemcpyAsync(&h, c->ctrl.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (h.done) {
            c->calib_hint = std::min(std::max(h.counter, 1), 64);
            break;
        }
        batch = 4;
    }
    // the fast path's last factor is still pending when no sampling launch followed its control
    if (fast) { HIPCHK(launch_scale_steps(ch, c->ctrl.p, c->dimsizes.p, c->stream)); }
    return MCI_OK;
}

// The compacted copies of a chain's sampling columns (positions, step size, generator state,
// acceptance counters), for the walkers still in a per-walker loop.
mci_status dense_chain(mci_ctx *c, const Chain &ch, Chain &dense)
{
    const int64_t W = c->W;
    const bool mt = c->rng == MCI_RNG_MT19937_64;
    HIPCHK(c->cmp_x.ensure((size_t)c->ndim * W));
    HIPCHK(c->cmp_steps.ensure(W));
    HIPCHK(c->cmp_acc.ensure(W));
    HIPCHK(c->cmp_rej.ensure(W));
    if (mt) {
        HIPCHK(c->cmp_mt.ensure((size_t)MT_N * W));
        HIPCHK(c->cmp_mtp.ensure(W));
    }
    else { HIPCHK(c->cmp_ctr.ensure(W)); }
    HIPCHK(c->cmp_list.ensure(W));
    HIPCHK(c->cmp_count.ensure(1));
    dense = ch;
    dense.xold = c->cmp_x.p;
    dense.steps = c->cmp_steps.p;
    dense.acc = c->cmp_acc.p;
    dense.rej = c->cmp_rej.p;
    dense.mt = mt ? c->cmp_mt.p : nullptr;
    dense.mtp = mt ? c->cmp_mtp.p : nullptr;
    dense.ctr = mt ? nullptr : c->cmp_ctr.p;
    return MCI_OK;
}

// One specialised sampling launch of the walkers whose flag (c->active) is set, compacted: the
// flagged walkers listed (c->cmp_list, c->cmp_count), their columns gathered into the dense chain,
// the launch over the list (a wave with one flagged walker would otherwise run in full), the columns
// scattered back.  Full traces (obs) are written to the dense columns.
mci_status compacted_sample(mci_ctx *c, const Chain &ch, const Chain &dense, int prec, int64_t nsteps, const ObsDev *obs, int nobs)
{
    HIPCHK(hipMemsetAsync(c->cmp_count.p, 0, sizeof(int32_t), c->stream));
    HIPCHK(launch_compact_active(c->active.p, c->W, c->cmp_list.p, c->cmp_count.p, c->stream));
    HIPCHK(launch_walker_columns(ch, dense, c->cmp_list.p, c->cmp_count.p, 1, c->stream));
    TRY(fast_sample(c->fast, dense, prec, nsteps, obs, nobs, c->stream, nullptr, nullptr, nullptr, c->cmp_list.p, c->cmp_count.p));
    HIPCHK(launch_walker_columns(ch, dense, c->cmp_list.p, c->cmp_count.p, 0, c->stream));
    c->mt_diverged = true;
    return MCI_OK;
}

// MCI::findMRT2Step for independent walkers (MCIntegrator.cpp:87-168, one loop per walker) on the
// specialised kernels: an iteration is a sampling launch of the walkers still in their loop and
// k_calib_indep_round's bookkeeping; the host issues batches of iterations and reads, after each
// batch, whether any walker is left (iterations with none left are no-ops).  The walkers still in
// their loop are compacted into dense columns for the sampling launch (most have left after a few
// iterations, and a wave with one of them would otherwise run in full), then moved back.  The same
// walker states, counters and step sizes, bit for bit, as k_calibrate_indep.
mci_status find_indep_fast(mci_ctx *c, int64_t min_stat)
{
    const int N = c->nfind;
    if (N == 0) { return MCI_OK; }
    const int64_t W = c->W;
    const int maxit = std::abs(N);
    HIPCHK(c->active.ensure(W));
    HIPCHK(c->cal_cons.ensure(W));
    HIPCHK(c->cal_count.ensure(W));
    HIPCHK(c->cal_any.ensure(maxit));
    HIPCHK(hipMemsetAsync(c->active.p, 1, W, c->stream));
    HIPCHK(hipMemsetAsync(c->cal_cons.p, 0, sizeof(int32_t) * W, c->stream));
    HIPCHK(hipMemsetAsync(c->cal_count.p, 0, sizeof(int32_t) * W, c->stream));
    HIPCHK(hipMemsetAsync(c->cal_any.p, 0, sizeof(int32_t) * maxit, c->stream));
    TRY(mt_align_if_needed(c));
    Chain ch = make_chain(c, nullptr), dense;
    TRY(dense_chain(c, ch, dense));
    int batch = (c->calib_indep_hint > 0) ? c->calib_indep_hint : 4;
    for (int it = 0; it < maxit;) {
        for (int b = 0; b < batch && it < maxit; ++b, ++it) {
            // the calibration runs the FP64 chain in both precisions (step sizes stay FP64 quantities)
            TRY(compacted_sample(c, ch, dense, MCI_PREC_FP64, min_stat, nullptr, 0));
            HIPCHK(launch_calib_indep_round(ch, c->active.p, c->cal_cons.p, c->cal_count.p, N, c->target, c->dimsizes.p,
                                            c->cal_any.p + it, c->stream));
        }
        int32_t left = 0;
        HIPCHK(hipMemcpyAsync(&left, c->cal_any.p + (it - 1), sizeof left, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (!left) {
            c->calib_indep_hint = std::min(std::max(it, 1), 64);
            break;
        }
        batch = 4;
    }
    return MCI_OK;
}

// what the dependent observables depend on exists (checked when integrate starts: sampling
// functions may be added after the observables)
mci_status check_dependencies(const mci_ctx *c)
{
    for (const ObsHost &o : c->obs) {
        if (o.kind != MCI_OBS_PDF_LOG && o.kind != MCI_OBS_PDF_PROTO) { continue; }
        const int j = (int)obs_p0(o, c->ndim);
        if (j >= (int)c->pdfs.size()) {
            return fail(MCI_E_INVALID_ARGUMENT, "an observable depends on sampling function %d, but there are %d", j, (int)c->pdfs.size());
        }
        if (o.kind == MCI_OBS_PDF_PROTO && c->pdfs[j].nproto != c->ndim) {
            return fail(MCI_E_INVALID_ARGUMENT, "the proto-value observable needs a sampling function with one proto value per dimension");
        }
    }
    return MCI_OK;
}

// MCI::initialDecorrelation, automatic branch (MCIntegrator.cpp:173-243): equilibration
// observables with Full accumulators and the Correlated estimator; walkers either stop
// individually (independent) or together on the rank-combined estimates (shared)
mci_status decorrelate_auto(mci_ctx *c)
{
    const bool shared = (c->coupling == MCI_COUPLING_SHARED) && c->W > 1;
    const int64_t nranks = (c->coupling == MCI_COUPLING_SHARED) ? c->W : 1;
    const int64_t min_nmc = (int64_t)std::fmax(100., std::sqrt(40000. * c->ndim) / (double)nranks);
    const int64_t W = c->W;
    std::vector<ObsDev> od;
    std::vector<DBuf<double>> data, scr;
    std::vector<int> eq_index(c->obs.size(), -1); // an observable's index in the equilibration container
    int nd = 0;
    for (size_t i = 0; i < c->obs.size(); ++i) {
        const ObsHost &o = c->obs[i];
        if (!o.equil) { continue; }
        eq_index[i] = (int)od.size();
        ObsDev d;
        std::memset(&d, 0, sizeof d);
        d.kind = o.kind;
        d.nobs = o.nobs;
        d.voff = nd;
        d.nskip = 1;
        d.acc = ACC_FULL;
        d.blocksize = 1;
        d.p0 = obs_p0(o, c->ndim);
        d.uid = obs_uid(c, i);
        d.par = o.dev.p;
        if (o.kind == MCI_OBS_DEP_SQUARE) {
            const int k = eq_index[(size_t)d.p0];
            if (k < 0) {
                return fail(MCI_E_INVALID_ARGUMENT, "an equilibration observable depends on observable %d, which is not one", (int)d.p0);
            }
            d.p0 = k;
        }
        d.nstore = min_nmc;
        data.emplace_back();
        HIPCHK(data.back().ensure((size_t)min_nmc * o.nobs * W));
        d.data = data.back().p;
        scr.emplace_back();
        if (is_pow2(min_nmc)) { HIPCHK(scr.back().ensure((size_t)mj_scratch_rows(min_nmc) * o.nobs * W)); }
        od.push_back(d);
        nd += o.nobs;
    }
    HIPCHK(c->obsval.ensure((size_t)std::max(nd, std::max(nobsdim(c), 1)) * W));
    const size_t ndw = (size_t)std::max(nd, 1) * W;
    HIPCHK(c->wavg.ensure(ndw));
    HIPCHK(c->werr.ensure(ndw));
    HIPCHK(c->oa.ensure(ndw));
    HIPCHK(c->oe.ensure(ndw));
    HIPCHK(c->active.ensure(W));
    const int64_t nch = (W + 255) / 256;
    HIPCHK(c->chunks_a.ensure((size_t)nch * std::max(nd, 1)));
    HIPCHK(c->chunks_e.ensure((size_t)nch * std::max(nd, 1)));
    HIPCHK(c->res.ensure(2 * (size_t)std::max(nd, 1)));
    HIPCHK(hipMemsetAsync(c->active.p, 1, W, c->stream));
    Chain ch = make_chain(c, &od);
    // the equilibration traces on the specialised kernels where they apply (Full accumulators of the
    // X2 family; the independent rounds skip the finished walkers as the generic kernel does) --
    // the same traces, bit for bit, as the generic kernel's
    const bool fast_dec = fast_chain(c) && fast_obs_eligible(od.data(), (int)od.size(), c->ndim, c->prec, c->movetype, c->rng);
    // the independent rounds after the first step the walkers still decorrelating, compacted (their
    // traces in the dense columns, which the estimators map back to the walkers)
    Chain dense;
    if (fast_dec) {
        TRY(mt_align_if_needed(c));
        if (!shared) { TRY(dense_chain(c, ch, dense)); }
    }
    bool compacted = false;
    auto dec_launch = [&](const uint8_t *act) -> mci_status {
        compacted = fast_dec && act != nullptr;
        if (compacted) { return compacted_sample(c, ch, dense, c->prec, min_nmc, od.data(), (int)od.size()); }
        if (act) { c->mt_diverged = true; }
        if (fast_dec) { return fast_sample(c->fast, ch, c->prec, min_nmc, od.data(), (int)od.size(), c->stream); }
        HIPCHK(sample_launch(c, ch, min_nmc, 1, act, c->stream));
        return MCI_OK;
    };

    auto sample_estimate = [&](const uint8_t *act) -> mci_status {
        TRY(dec_launch(act));
        for (size_t i = 0; i < od.size(); ++i) {
            EstimJob j;
            j.data = od[i].data;
            j.scratch = scr[i].p;
            j.n = min_nmc;
            j.W = W;
            j.nobs = od[i].nobs;
            j.off = od[i].voff;
            j.estim = MCI_ESTIM_CORRELATED;
            j.scale = 1.0;
            j.active = act;
            j.avg = c->wavg.p;
            j.err = c->werr.p;
            if (compacted) {
                j.wmap = c->cmp_list.p;
                j.wcount = c->cmp_count.p;
            }
            HIPCHK(launch_estimate(j, c->stream));
        }
        return MCI_OK;
    };
    auto combine = [&](std::vector<double> &ha, std::vector<double> &he) -> mci_status {
        HIPCHK(launch_chunk_sum(c->wavg.p, W, nd, 0, c->chunks_a.p, c->stream));
        HIPCHK(launch_chunk_sum(c->werr.p, W, nd, 1, c->chunks_e.p, c->stream));
        HIPCHK(launch_combine(c->chunks_a.p, c->chunks_e.p, W, nd, c->res.p, c->res.p + nd, c->stream));
        ha.resize(nd);
        he.resize(nd);
        HIPCHK(hipMemcpyAsync(ha.data(), c->res.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(he.data(), c->res.p + nd, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return MCI_OK;
    };

    std::vector<double> oa, oe, na, ne;
    TRY(sample_estimate(nullptr));
    if (shared) { TRY(combine(oa, oe)); }
    else {
        HIPCHK(hipMemcpyAsync(c->oa.p, c->wavg.p, ndw * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->oe.p, c->werr.p, ndw * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    int64_t count = 0;
    const int64_t cap = std::llabs(c->ndecorr);
    std::vector<uint8_t> hact(W);
    while (true) {
        const uint8_t *act = shared ? nullptr : c->active.p;
        TRY(dec_launch(act));
        count += min_nmc;
        if (count >= cap) {
            std::printf("Warning [MCI::initialDecorrelation]: Max number of MC steps reached without equilibration.\n");
            std::fflush(stdout);
            break;
        }
        for (size_t i = 0; i < od.size(); ++i) {
            EstimJob j;
            j.data = od[i].data;
            j.scratch = scr[i].p;
            j.n = min_nmc;
            j.W = W;
            j.nobs = od[i].nobs;
            j.off = od[i].voff;
            j.estim = MCI_ESTIM_CORRELATED;
            j.scale = 1.0;
            j.active = act;
            j.avg = c->wavg.p;
            j.err = c->werr.p;
            if (compacted) {
                j.wmap = c->cmp_list.p;
                j.wcount = c->cmp_count.p;
            }
            HIPCHK(launch_estimate(j, c->stream));
        }
        if (shared) {
            TRY(combine(na, ne));
            bool again = false;
            for (int i = 0; i < nd; ++i) {
                if (std::fabs(oa[i] - na[i]) > 2 * std::sqrt(oe[i] * oe[i] + ne[i] * ne[i])) { again = true; break; }
            }
            oa = na;
            oe = ne;
            if (!again) { break; }
        }
        else {
            HIPCHK(launch_decorr_compare(c->oa.p, c->oe.p, c->wavg.p, c->werr.p, nd, W, c->active.p, c->stream));
            HIPCHK(hipMemcpyAsync(hact.data(), c->active.p, W, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            int64_t nact = 0;
            for (int64_t w = 0; w < W; ++w) { nact += hact[w]; }
            if (nact == 0) { break; }
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}
} // namespace

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

const char *mci_last_error(void) { return g_err.c_str(); }

const char *mci_version(void) { return kVersion; }

mci_status mci_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { n = 0; }
    *count = n;
    return MCI_OK;
}

mci_status mci_create(int ndim, int64_t nwalkers, mci_ctx **out)
{
    if (!out) { return fail(MCI_E_INVALID_ARGUMENT, "null output pointer"); }
    *out = nullptr;
    if (ndim < 1) { return fail(MCI_E_INVALID_ARGUMENT, "[ProtoFunctionInterface] Number of dimensions must be at least 1."); }
    if (nwalkers < 1) { return fail(MCI_E_INVALID_ARGUMENT, "number of walkers must be at least 1"); }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { return fail(MCI_E_DEVICE, "no HIP device available"); }
    mci_ctx *c = new mci_ctx();
    c->ndim = ndim;
    c->W = nwalkers;
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) { HIPCHK(hipEventCreate(&e)); }
    c->lb.assign(ndim, 0.);
    c->ub.assign(ndim, 0.);
    HIPCHK(c->x.ensure((size_t)ndim * nwalkers));
    HIPCHK(c->xn.ensure((size_t)ndim * nwalkers));
    HIPCHK(c->acc.ensure(nwalkers));
    HIPCHK(c->rej.ensure(nwalkers));
    HIPCHK(c->lbd.ensure(ndim));
    HIPCHK(c->ubd.ensure(ndim));
    HIPCHK(c->dimsizes.ensure(ndim));
    HIPCHK(hipMemsetAsync(c->x.p, 0, sizeof(double) * ndim * nwalkers, c->stream)); // WalkerState.hpp:30 x = 0
    HIPCHK(hipMemsetAsync(c->acc.p, 0, sizeof(int64_t) * nwalkers, c->stream));
    HIPCHK(hipMemsetAsync(c->rej.p, 0, sizeof(int64_t) * nwalkers, c->stream));
    std::vector<double> sz(ndim, (double)FLT_MAX + (double)FLT_MAX);
    HIPCHK(hipMemcpyAsync(c->dimsizes.p, sz.data(), sizeof(double) * ndim, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    TRY(ensure_protos(c));
    mci_status st = mci_set_trial_move(c, MCI_SRRD_UNIFORM, 0, 1, nullptr); // MCI ctor default (MoveType::All)
    if (st != MCI_OK) { mci_destroy(c); return st; }
    *out = c;
    return MCI_OK;
}

mci_status mci_destroy(mci_ctx *c)
{
    if (!c) { return MCI_OK; }
    (void)hipSetDevice(c->device);
    if (c->stream) { (void)hipStreamSynchronize(c->stream); }
    for (auto &e : c->ev) { if (e) { (void)hipEventDestroy(e); } }
    fast_release(c->fast);
    hipStream_t s = c->stream;
    delete c;
    if (s) { (void)hipStreamDestroy(s); }
    return MCI_OK;
}

mci_status mci_set_seed(mci_ctx *c, uint64_t seed)
{
    CHECKCTX(c);
    c->seed = seed;
    c->wseeds.clear();
    c->reseed = true;
    return MCI_OK;
}

mci_status mci_set_walker_seeds(mci_ctx *c, const uint64_t *seeds, int64_t n)
{
    CHECKCTX(c);
    if (n != c->W || !seeds) { return fail(MCI_E_INVALID_ARGUMENT, "need one seed per walker"); }
    c->wseeds.assign(seeds, seeds + n);
    c->reseed = true;
    return MCI_OK;
}

mci_status mci_set_rng(mci_ctx *c, int k)
{
    CHECKCTX(c);
    if (k != MCI_RNG_MT19937_64 && k != MCI_RNG_PHILOX) { return fail(MCI_E_DOMAIN, "unknown random-number mode %d", k); }
    c->rng = k;
    c->reseed = true;
    return MCI_OK;
}

mci_status mci_set_precision(mci_ctx *c, int p)
{
    CHECKCTX(c);
    if (p != MCI_PREC_FP64 && p != MCI_PREC_FP32) { return fail(MCI_E_DOMAIN, "unknown precision %d", p); }
    c->prec = p;
    return MCI_OK;
}

mci_status mci_set_coupling(mci_ctx *c, int m)
{
    CHECKCTX(c);
    if (m != MCI_COUPLING_INDEPENDENT && m != MCI_COUPLING_SHARED) { return fail(MCI_E_DOMAIN, "unknown coupling %d", m); }
    c->coupling = m;
    return MCI_OK;
}

mci_status mci_set_kernel_path(mci_ctx *c, int path)
{
    CHECKCTX(c);
    if (path != MCI_KERNELS_AUTO && path != MCI_KERNELS_GENERIC) { return fail(MCI_E_DOMAIN, "unknown kernel path %d", path); }
    c->kernel_path = path;
    return MCI_OK;
}

mci_status mci_set_target_acceptance_rate(mci_ctx *c, double r) { CHECKCTX(c); c->target = r; return MCI_OK; }
mci_status mci_set_nfind_mrt2_iterations(mci_ctx *c, int n) { CHECKCTX(c); c->nfind = n; return MCI_OK; }
mci_status mci_set_ndecorrelation_steps(mci_ctx *c, int64_t n) { CHECKCTX(c); c->ndecorr = n; return MCI_OK; }

mci_status mci_get_settings(mci_ctx *c, double *t, int *nf, int64_t *nd)
{
    CHECKCTX(c);
    if (t) { *t = c->target; }
    if (nf) { *nf = c->nfind; }
    if (nd) { *nd = c->ndecorr; }
    return MCI_OK;
}

mci_status mci_set_x(mci_ctx *c, const double *x, int64_t n)
{
    CHECKCTX(c);
    if (!x || (n != c->ndim && n != c->ndim * c->W)) { return fail(MCI_E_INVALID_ARGUMENT, "x must have ndim or nwalkers*ndim entries"); }
    TRY(set_current(c));
    std::vector<double> h((size_t)c->ndim * c->W);
    for (int64_t w = 0; w < c->W; ++w) {
        for (int i = 0; i < c->ndim; ++i) { h[(size_t)i * c->W + w] = (n == c->ndim) ? x[i] : x[w * c->ndim + i]; }
    }
    HIPCHK(hipMemcpyAsync(c->x.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->periodic) { HIPCHK(launch_apply_domain(make_chain(c, nullptr), c->stream)); }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_get_x(mci_ctx *c, double *x, int64_t n)
{
    CHECKCTX(c);
    if (!x || n < c->ndim * c->W) { return fail(MCI_E_INVALID_ARGUMENT, "output too small"); }
    TRY(set_current(c));
    std::vector<double> h((size_t)c->ndim * c->W);
    HIPCHK(hipMemcpyAsync(h.data(), c->x.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t w = 0; w < c->W; ++w) {
        for (int i = 0; i < c->ndim; ++i) { x[w * c->ndim + i] = h[(size_t)i * c->W + w]; }
    }
    return MCI_OK;
}

mci_status mci_center_x(mci_ctx *c) // DomainInterface::getCenter (DomainInterface.hpp:39-43)
{
    CHECKCTX(c);
    std::vector<double> ctr(c->ndim);
    for (int i = 0; i < c->ndim; ++i) {
        ctr[i] = c->periodic ? c->lb[i] + 0.5 * (c->ub[i] - c->lb[i])
                             : -(double)FLT_MAX + 0.5 * ((double)FLT_MAX + (double)FLT_MAX);
    }
    return mci_set_x(c, ctr.data(), c->ndim);
}

mci_status mci_move_x(mci_ctx *c)
{
    CHECKCTX(c);
    TRY(set_current(c));
    TRY(ensure_protos(c));
    TRY(ensure_rng(c));
    c->mt_diverged = true;
    HIPCHK(launch_move_x(make_chain(c, nullptr), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_new_random_x(mci_ctx *c)
{
    CHECKCTX(c);
    TRY(set_current(c));
    TRY(ensure_rng(c));
    c->mt_diverged = true;
    HIPCHK(launch_new_random_x(make_chain(c, nullptr), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_set_mrt2step(mci_ctx *c, int idx, double step)
{
    CHECKCTX(c);
    TRY(set_current(c));
    if (idx < 0) { return fill_rows(c, c->steps.p, 0, c->ntypes, step); }
    if (idx < c->ntypes) { return fill_rows(c, c->steps.p, idx, 1, step); }
    std::printf("[MCI::setMRT2Step] Warning: Tried to set non-existing MRT2step index.\n");
    return MCI_OK;
}

mci_status mci_set_mrt2steps(mci_ctx *c, const double *steps, int n)
{
    CHECKCTX(c);
    if (!steps || n != c->ntypes) { return fail(MCI_E_INVALID_ARGUMENT, "need one step per step-size type"); }
    for (int t = 0; t < n; ++t) { TRY(mci_set_mrt2step(c, t, steps[t])); }
    return MCI_OK;
}

mci_status mci_get_mrt2step(mci_ctx *c, int idx, double *step)
{
    CHECKCTX(c);
    if (!step) { return fail(MCI_E_INVALID_ARGUMENT, "null output"); }
    if (idx < 0 || idx >= c->ntypes) { *step = 0.; return MCI_OK; }
    TRY(set_current(c));
    HIPCHK(hipMemcpyAsync(step, c->steps.p + (size_t)idx * c->W, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_get_walker_mrt2steps(mci_ctx *c, double *steps, int64_t n)
{
    CHECKCTX(c);
    if (!steps || n < c->W * c->ntypes) { return fail(MCI_E_INVALID_ARGUMENT, "output too small"); }
    TRY(set_current(c));
    std::vector<double> h((size_t)c->ntypes * c->W);
    HIPCHK(hipMemcpyAsync(h.data(), c->steps.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t w = 0; w < c->W; ++w) {
        for (int t = 0; t < c->ntypes; ++t) { steps[w * c->ntypes + t] = h[(size_t)t * c->W + w]; }
    }
    return MCI_OK;
}

mci_status mci_get_n_step_sizes(mci_ctx *c, int *n) { CHECKCTX(c); *n = c->ntypes; return MCI_OK; }

mci_status mci_reset_domain(mci_ctx *c)
{
    CHECKCTX(c);
    TRY(set_current(c));
    c->periodic = false;
    std::vector<double> sz(c->ndim, (double)FLT_MAX + (double)FLT_MAX);
    HIPCHK(hipMemcpyAsync(c->dimsizes.p, sz.data(), sizeof(double) * c->ndim, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_set_irange(mci_ctx *c, const double *lb, const double *ub)
{
    CHECKCTX(c);
    if (!lb || !ub) { return fail(MCI_E_INVALID_ARGUMENT, "null bounds"); }
    for (int i = 0; i < c->ndim; ++i) {
        if (ub[i] <= lb[i]) {
            return fail(MCI_E_INVALID_ARGUMENT,
                        "[OrthoPeriodicDomain::checkBounds] All upper bounds must be truly greater than their corresponding lower bounds.");
        }
    }
    TRY(set_current(c));
    c->lb.assign(lb, lb + c->ndim);
    c->ub.assign(ub, ub + c->ndim);
    c->periodic = true;
    std::vector<double> sz(c->ndim);
    for (int i = 0; i < c->ndim; ++i) { sz[i] = ub[i] - lb[i]; }
    HIPCHK(hipMemcpyAsync(c->lbd.p, lb, sizeof(double) * c->ndim, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ubd.p, ub, sizeof(double) * c->ndim, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->dimsizes.p, sz.data(), sizeof(double) * c->ndim, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_apply_domain(make_chain(c, nullptr), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_domain_is_finite(mci_ctx *c, int *f) { CHECKCTX(c); *f = c->periodic ? 1 : 0; return MCI_OK; }

mci_status mci_set_trial_move(mci_ctx *c, int srrd, int veclen, int ntypes, const int *te)
{
    CHECKCTX(c);
    if (srrd < MCI_SRRD_UNIFORM || srrd > MCI_SRRD_FISHER) { return fail(MCI_E_DOMAIN, "[createSRRDAllMove] Unhandled SRRDType enumerator."); }
    if (ntypes < 1) { ntypes = 1; }
    if (ntypes > MAX_TYPES) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d step-size types", MAX_TYPES); }
    if (ntypes > 1 && !te) { return fail(MCI_E_INVALID_ARGUMENT, "[checkTrialMoveSanity] ntypes>1 requires passed typeEnds array."); }
    if (veclen > 0 && c->ndim % veclen != 0) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::setTrialMove] MCI's number of walkers must be a multiple of passed veclen.");
    }
    if (veclen > 0 && ntypes > 1) {
        for (int i = 0; i < ntypes; ++i) {
            if (te[i] % veclen != 0) { return fail(MCI_E_INVALID_ARGUMENT, "[SRRDVecMove] All type end indices must be multiples of vector length."); }
        }
    }
    if (ntypes > 1) {
        for (int i = 0; i < ntypes; ++i) {
            if (te[i] < 1 || te[i] > c->ndim || (i > 0 && te[i] <= te[i - 1])) { return fail(MCI_E_INVALID_ARGUMENT, "typeEnds must increase within [1, ndim]"); }
        }
        if (te[ntypes - 1] != c->ndim) { return fail(MCI_E_INVALID_ARGUMENT, "the last type end must equal ndim"); }
    }
    TRY(set_current(c));
    c->typeEnds.assign(ntypes, c->ndim);
    if (ntypes > 1) { c->typeEnds.assign(te, te + ntypes); }
    c->ntypes = ntypes;
    c->srrd = srrd;
    c->movetype = (veclen > 0) ? MCI_MOVE_VEC : MCI_MOVE_ALL;
    c->veclen = veclen > 0 ? veclen : 0;
    c->nvecs = veclen > 0 ? c->ndim / veclen : 0;
    HIPCHK(c->steps.ensure((size_t)ntypes * c->W));
    // a new move has fresh distribution objects: no saved normals
    HIPCHK(c->nsv.ensure((size_t)2 * c->W));
    HIPCHK(c->nsf.ensure((size_t)c->W));
    HIPCHK(hipMemsetAsync(c->nsf.p, 0, sizeof(int32_t) * (size_t)c->W, c->stream));
    c->ms_on = 0;
    c->ms_nsteps = 0;
    c->ms_pdfs.clear();
    c->lm_K = 0;  // a new trial move replaces a mixed one too
    c->lm_pmix = 0.;
    return fill_rows(c, c->steps.p, 0, ntypes, 0.05); // DEFAULT_MRT2STEP (Factories.hpp:145)
}

mci_status mci_set_trial_move_type(mci_ctx *c, int t)
{
    CHECKCTX(c);
    if (t == MCI_MOVE_ALL) { return mci_set_trial_move(c, MCI_SRRD_UNIFORM, 0, 1, nullptr); }
    if (t == MCI_MOVE_VEC) { return mci_set_trial_move(c, MCI_SRRD_UNIFORM, 1, 1, nullptr); }
    if (t == MCI_MOVE_MULTISTEP) { return mci_set_trial_move_multistep(c, c->ndim, MCI_SRRD_UNIFORM, 1, 1, nullptr); } // MultiStepMove(ndim)
    return fail(MCI_E_DOMAIN, "[createMoveDefault] Unhandled MoveType enumerator.");
}

mci_status mci_get_change_rate(mci_ctx *c, double *r)
{
    CHECKCTX(c);
    *r = (c->movetype == MCI_MOVE_ALL) ? 1. : (c->movetype == MCI_MOVE_DEVICE ? c->mv_rate : 1. / c->nvecs);
    if (c->ms_on) { *r = std::fmin(1., *r * c->ms_nsteps); } // MultiStepMove.hpp:67
    return MCI_OK;
}

} // extern "C"

namespace
{
// a sampling function of the catalogue: proto count and device parameter blob
mci_status build_pdf(int kind, int ndim, const double *params, int nparams, PdfHost &p)
{
    if (nparams < 0 || (nparams > 0 && !params)) { return fail(MCI_E_INVALID_ARGUMENT, "bad parameter array"); }
    p.kind = kind;
    p.params.assign(params, params + nparams);
    std::vector<double> blob;
    switch (kind) {
    case MCI_PDF_GAUSS_ISO:
        p.nproto = 1;
        break;
    case MCI_PDF_GAUSS_DIAG:
    case MCI_PDF_EXP_DIAG:
        if (nparams != 0 && nparams != ndim) { return fail(MCI_E_INVALID_ARGUMENT, "params must be empty or have ndim entries"); }
        p.nproto = ndim;
        blob = p.params;
        break;
    case MCI_PDF_GMM_ISO: {
        if (nparams < 1) { return fail(MCI_E_INVALID_ARGUMENT, "GMM needs K"); }
        const int K = (int)params[0];
        if (K < 1 || nparams != 1 + 2 * K + K * ndim) { return fail(MCI_E_INVALID_ARGUMENT, "GMM params must be [K, w[K], s[K], mu[K][ndim]]"); }
        p.K = K;
        p.nproto = K + 1;
        blob.resize(2 * K + (size_t)K * ndim);
        for (int k = 0; k < K; ++k) {
            const double w = params[1 + k], s = params[1 + K + k];
            if (!(w > 0.) || !(s > 0.)) { return fail(MCI_E_INVALID_ARGUMENT, "GMM weights and widths must be positive"); }
            blob[k] = std::log(w) - (double)ndim * std::log(s);
            blob[K + k] = 0.5 / (s * s);
        }
        std::copy(params + 1 + 2 * K, params + nparams, blob.begin() + 2 * K);
        break;
    }
    default:
        return fail(MCI_E_DOMAIN, "unknown sampling-function kind %d", kind);
    }
    if (!blob.empty()) {
        HIPCHK(p.dev.ensure(blob.size()));
        HIPCHK(hipMemcpy(p.dev.p, blob.data(), blob.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return MCI_OK;
}

// the files of MCI::storeObservables / storeWalkerPositions (MCIntegrator.cpp:509-516, 533-542),
// one per walker (the file name itself for one walker, name.<walker> for several), written with the
// same stream formatting from the values the sampling kernel recorded
mci_status write_files(mci_ctx *c, int64_t nmc, int nd, bool rec_obs, bool rec_x)
{
    const int64_t W = c->W;
    HIPCHK(hipStreamSynchronize(c->stream));
    auto name = [&](const std::string &p, int64_t w) { return W == 1 ? p : p + "." + std::to_string(w); };
    if (rec_obs) {
        const int64_t no = (nmc - 1) / c->obsfreq + 1, nc = std::max(nd, 1);
        std::vector<double> h((size_t)no * nc * W);
        HIPCHK(hipMemcpy(h.data(), c->rec_obs.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t w = 0; w < W; ++w) {
            std::ofstream f(name(c->obsfile, w));
            for (int64_t r = 0; r < no; ++r) {
                f << r * c->obsfreq;
                for (int j = 0; j < nd; ++j) { f << " " << h[((size_t)r * nc + j) * W + w]; }
                f << " ";
                f << std::endl;
            }
        }
    }
    if (rec_x) {
        const int64_t nx = (nmc - 1) / c->wlkfreq + 1;
        std::vector<double> h((size_t)nx * c->ndim * W);
        HIPCHK(hipMemcpy(h.data(), c->rec_x.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t w = 0; w < W; ++w) {
            std::ofstream f(name(c->wlkfile, w));
            for (int64_t r = 0; r < nx; ++r) {
                f << r * c->wlkfreq;
                for (int j = 0; j < c->ndim; ++j) { f << "   " << h[((size_t)r * c->ndim + j) * W + w]; }
                f << std::endl;
            }
        }
    }
    return MCI_OK;
}

mci_status ensure_ms_protos(mci_ctx *c)
{
    int n = 0;
    for (const auto &p : c->ms_pdfs) { n += p.nproto; }
    const size_t sz = (size_t)std::max(n, 1) * c->W;
    HIPCHK(c->ms_pold.ensure(sz));
    HIPCHK(c->ms_pnew.ensure(sz));
    return MCI_OK;
}
} // namespace

extern "C" {

mci_status mci_add_sampling_function(mci_ctx *c, int kind, int ndim, const double *params, int nparams)
{
    CHECKCTX(c);
    if (ndim != c->ndim) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addSamplingFunction] Passed sampling function's number of inputs is not equal to MCI's number of walkers.");
    }
    if ((int)c->pdfs.size() >= MAX_PDF) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d sampling functions", MAX_PDF); }
    TRY(set_current(c));
    PdfHost p;
    TRY(build_pdf(kind, ndim, params, nparams, p));
    c->pdfs.push_back(std::move(p));
    return ensure_protos(c);
}

// MultiStepMove(ndim, nsteps) (MultiStepMove.hpp:40-46) whose contained move is the SRRD move
// that mci_set_trial_move builds from the same arguments (MultiStepMove::setTrialMove,
// MultiStepMove.cpp:51-58); it has no sampling function of its own yet
mci_status mci_set_trial_move_multistep(mci_ctx *c, int nsteps, int srrd, int veclen, int ntypes, const int *te)
{
    CHECKCTX(c);
    if (nsteps < 0) { return fail(MCI_E_INVALID_ARGUMENT, "the number of steps must not be negative"); }
    TRY(mci_set_trial_move(c, srrd, veclen, ntypes, te));
    HIPCHK(c->ms_x0.ensure((size_t)c->ndim * c->W));
    c->ms_on = 1;
    c->ms_nsteps = nsteps;
    return ensure_ms_protos(c);
}

// MCI::storeObservablesOnFile / clearObservableFile / storeWalkerPositionsOnFile / clearWalkerFile
// (MCIntegrator.cpp:495-531)
mci_status mci_store_observables_on_file(mci_ctx *c, const char *path, int freq)
{
    CHECKCTX(c);
    if (!path || freq < 1) { return fail(MCI_E_INVALID_ARGUMENT, "need a file path and a frequency of at least 1"); }
    c->obsfile = path;
    c->obsfreq = freq;
    return MCI_OK;
}

mci_status mci_clear_observable_file(mci_ctx *c)
{
    CHECKCTX(c);
    c->obsfile.clear();
    c->obsfreq = 0;
    return MCI_OK;
}

mci_status mci_store_walker_positions_on_file(mci_ctx *c, const char *path, int freq)
{
    CHECKCTX(c);
    if (!path || freq < 1) { return fail(MCI_E_INVALID_ARGUMENT, "need a file path and a frequency of at least 1"); }
    c->wlkfile = path;
    c->wlkfreq = freq;
    return MCI_OK;
}

mci_status mci_clear_walker_file(mci_ctx *c)
{
    CHECKCTX(c);
    c->wlkfile.clear();
    c->wlkfreq = 0;
    return MCI_OK;
}

// MultiStepMove::addSamplingFunction (MultiStepMove.cpp:69-75)
mci_status mci_add_multistep_sampling_function(mci_ctx *c, int kind, int ndim, const double *params, int nparams)
{
    CHECKCTX(c);
    if (!c->ms_on) { return fail(MCI_E_RUNTIME, "the trial move is not a MultiStepMove"); }
    if (ndim != c->ndim) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MultiStepMove::addSamplingFunction] Passed sampling function's number of inputs is not equal to number of walkers.");
    }
    if ((int)c->ms_pdfs.size() >= MAX_PDF) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d sampling functions", MAX_PDF); }
    TRY(set_current(c));
    PdfHost p;
    TRY(build_pdf(kind, ndim, params, nparams, p));
    c->ms_pdfs.push_back(std::move(p));
    return ensure_ms_protos(c);
}

// A mixed trial move: with probability pmix an independence draw from a diagonal Gaussian mixture
// (the learned proposal's), else the move set before. No reference counterpart; it enters through the
// reference's trial-move hook, whose factor multiplies the acceptance (MCIntegrator.cpp:343).
mci_status mci_set_learned_move(mci_ctx *c, int K, const double *weights, const double *means, const double *sigmas, double pmix)
{
    CHECKCTX(c);
    if (K < 1 || K > 8 || !weights || !means || !sigmas) { return fail(MCI_E_INVALID_ARGUMENT, "need 1 to 8 mixture components"); }
    if (!(pmix >= 0. && pmix <= 1.)) { return fail(MCI_E_INVALID_ARGUMENT, "mixing probability must be in [0, 1]"); }
    const int d = c->ndim;
    double wsum = 0.;
    for (int k = 0; k < K; ++k) {
        if (!(weights[k] > 0.)) { return fail(MCI_E_INVALID_ARGUMENT, "mixture weights must be positive"); }
        wsum += weights[k];
    }
    for (int e = 0; e < K * d; ++e) {
        if (!(sigmas[e] > 0.) || !std::isfinite(means[e])) { return fail(MCI_E_INVALID_ARGUMENT, "mixture widths must be positive, means finite"); }
    }
    const size_t KD = (size_t)K * d;
    std::vector<double> h(3 * KD + 2 * (size_t)K);
    double cum = 0.;
    for (int k = 0; k < K; ++k) {
        const double pi = weights[k] / wsum;
        double ls = 0.;
        for (int j = 0; j < d; ++j) {
            h[(size_t)k * d + j] = means[(size_t)k * d + j];
            h[KD + (size_t)k * d + j] = 1. / sigmas[(size_t)k * d + j];
            h[2 * KD + (size_t)k * d + j] = sigmas[(size_t)k * d + j];
            ls = ls + log(1. / sigmas[(size_t)k * d + j]);
        }
        h[3 * KD + k] = log(pi) + ls - 0.5 * (double)d * log(2. * M_PI);
        cum = cum + pi;
        h[3 * KD + K + k] = (k == K - 1) ? 1. : cum;
    }
    TRY(set_current(c));
    HIPCHK(c->lm.ensure(h.size()));
    HIPCHK(hipMemcpy(c->lm.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    c->lm_K = K;
    c->lm_pmix = pmix;
    return MCI_OK;
}

mci_status mci_clear_learned_move(mci_ctx *c)
{
    CHECKCTX(c);
    c->lm_K = 0;
    c->lm_pmix = 0.;
    return MCI_OK;
}

// MultiStepMove::clearSamplingFunctions (MultiStepMove.cpp:64-67)
mci_status mci_clear_multistep_sampling_functions(mci_ctx *c)
{
    CHECKCTX(c);
    if (!c->ms_on) { return fail(MCI_E_RUNTIME, "the trial move is not a MultiStepMove"); }
    c->ms_pdfs.clear();
    return MCI_OK;
}

mci_status mci_pop_sampling_function(mci_ctx *c)
{
    CHECKCTX(c);
    if (c->pdfs.empty()) { return fail(MCI_E_RUNTIME, "no sampling function to pop"); }
    c->pdfs.pop_back();
    return MCI_OK;
}

mci_status mci_clear_sampling_functions(mci_ctx *c) { CHECKCTX(c); c->pdfs.clear(); return MCI_OK; }
mci_status mci_get_npdf(mci_ctx *c, int *n) { CHECKCTX(c); *n = (int)c->pdfs.size(); return MCI_OK; }

mci_status mci_add_observable(mci_ctx *c, int kind, int ndim, const double *params, int nparams, int blocksize, int nskip,
                              int flag_equil, int estim)
{
    CHECKCTX(c);
    blocksize = std::max(0, blocksize);
    nskip = std::max(1, nskip);
    if (ndim != c->ndim) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Passed observable function's number of inputs is not equal to MCI's number of walkers.");
    }
    if (flag_equil && estim == MCI_ESTIM_NOOP) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Requested automatic observable equilibration requires estimator with error calculation.");
    }
    if (estim < MCI_ESTIM_NOOP || estim > MCI_ESTIM_MJBLOCKER) { return fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator."); }
    if (nparams < 0 || (nparams > 0 && !params)) { return fail(MCI_E_INVALID_ARGUMENT, "bad parameter array"); }
    int nobs = obs_nobs(kind, ndim);
    if (kind == MCI_OBS_DEP_SQUARE) {
        // DependentObservableInterface::isObsDepValid (DependentObservableInterface.hpp:36-44): an earlier
        // observable whose nskip divides this one's
        const int k = nparams > 0 ? (int)params[0] : -1;
        if (nparams < 1 || params[0] != (double)k || k < 0 || k >= (int)c->obs.size()) {
            return fail(MCI_E_INVALID_ARGUMENT, "a squared observable needs the index of an earlier observable");
        }
        if (nskip % c->obs[k].nskip != 0) {
            return fail(MCI_E_INVALID_ARGUMENT, "the nskip of a dependent observable must be a multiple of its dependency's");
        }
        nobs = c->obs[k].nobs;
    }
    if ((kind == MCI_OBS_PDF_LOG || kind == MCI_OBS_PDF_PROTO) && nparams > 0 && (params[0] < 0 || params[0] >= MAX_PDF || params[0] != (int)params[0])) {
        return fail(MCI_E_INVALID_ARGUMENT, "sampling-function index out of range");
    }
    if (nobs < 0) { return fail(MCI_E_DOMAIN, "unknown observable kind %d", kind); }
    if ((int)c->obs.size() >= MAX_OBS) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d observables", MAX_OBS); }
    if (kind == MCI_OBS_INDICATOR && nparams > 0 && (params[0] < 0 || params[0] >= ndim)) {
        return fail(MCI_E_INVALID_ARGUMENT, "indicator index out of range");
    }
    ObsHost o;
    o.kind = kind;
    o.nobs = nobs;
    o.blocksize = blocksize;
    o.nskip = nskip;
    o.equil = flag_equil ? 1 : 0;
    o.estim = estim;
    o.params.assign(params, params + nparams);
    c->obs.push_back(std::move(o));
    return MCI_OK;
}

mci_status mci_pop_observable(mci_ctx *c)
{
    CHECKCTX(c);
    if (c->obs.empty()) { return fail(MCI_E_RUNTIME, "no observable to pop"); }
    c->obs.pop_back();
    return MCI_OK;
}

mci_status mci_clear_observables(mci_ctx *c) { CHECKCTX(c); c->obs.clear(); return MCI_OK; }

mci_status mci_add_sampling_function_device(mci_ctx *c, const char *name, const char *source, int ndim, int nproto, const double *params,
                                            int nparams)
{
    CHECKCTX(c);
    if (ndim != c->ndim) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addSamplingFunction] Passed sampling function's number of inputs is not equal to MCI's number of walkers.");
    }
    if (!name || !source || !rtc_valid_name(name)) { return fail(MCI_E_INVALID_ARGUMENT, "a device functor needs a source and a C identifier as its name"); }
    if (nproto < 1) { return fail(MCI_E_INVALID_ARGUMENT, "a sampling function needs at least one proto value"); }
    if (nparams < 0 || (nparams > 0 && !params)) { return fail(MCI_E_INVALID_ARGUMENT, "bad parameter array"); }
    if ((int)c->pdfs.size() >= MAX_PDF) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d sampling functions", MAX_PDF); }
    TRY(set_current(c));
    PdfHost p;
    p.kind = MCI_PDF_DEVICE;
    p.nproto = nproto;
    p.params.assign(params, params + nparams);
    p.fname = name;
    p.fsrc = source;
    if (nparams > 0) {
        HIPCHK(p.dev.ensure((size_t)nparams));
        HIPCHK(hipMemcpy(p.dev.p, params, sizeof(double) * (size_t)nparams, hipMemcpyHostToDevice));
    }
    c->pdfs.push_back(std::move(p));
    return ensure_protos(c);
}

mci_status mci_add_observable_device(mci_ctx *c, const char *name, const char *source, int ndim, int nobs, const double *params, int nparams,
                                     int blocksize, int nskip, int flag_equil, int estim)
{
    CHECKCTX(c);
    blocksize = std::max(0, blocksize);
    nskip = std::max(1, nskip);
    if (ndim != c->ndim) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Passed observable function's number of inputs is not equal to MCI's number of walkers.");
    }
    if (flag_equil && estim == MCI_ESTIM_NOOP) {
        return fail(MCI_E_INVALID_ARGUMENT, "[MCI::addObservable] Requested automatic observable equilibration requires estimator with error calculation.");
    }
    if (estim < MCI_ESTIM_NOOP || estim > MCI_ESTIM_MJBLOCKER) { return fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator."); }
    if (!name || !source || !rtc_valid_name(name)) { return fail(MCI_E_INVALID_ARGUMENT, "a device functor needs a source and a C identifier as its name"); }
    if (nobs < 1) { return fail(MCI_E_INVALID_ARGUMENT, "an observable needs at least one value"); }
    if (nparams < 0 || (nparams > 0 && !params)) { return fail(MCI_E_INVALID_ARGUMENT, "bad parameter array"); }
    if ((int)c->obs.size() >= MAX_OBS) { return fail(MCI_E_INVALID_ARGUMENT, "at most %d observables", MAX_OBS); }
    TRY(set_current(c));
    ObsHost o;
    o.kind = MCI_OBS_DEVICE;
    o.nobs = nobs;
    o.blocksize = blocksize;
    o.nskip = nskip;
    o.equil = flag_equil ? 1 : 0;
    o.estim = estim;
    o.params.assign(params, params + nparams);
    o.fname = name;
    o.fsrc = source;
    if (nparams > 0) {
        HIPCHK(o.dev.ensure((size_t)nparams));
        HIPCHK(hipMemcpy(o.dev.p, params, sizeof(double) * (size_t)nparams, hipMemcpyHostToDevice));
    }
    c->obs.push_back(std::move(o));
    return MCI_OK;
}

mci_status mci_set_trial_move_device(mci_ctx *c, const char *name, const char *source, int nproto, int nsteps, const int *te, double change_rate,
                                     const double *steps, const double *params, int nparams)
{
    CHECKCTX(c);
    if (!name || !source || !rtc_valid_name(name)) { return fail(MCI_E_INVALID_ARGUMENT, "a device functor needs a source and a C identifier as its name"); }
    if (nproto < 0) { return fail(MCI_E_INVALID_ARGUMENT, "nproto must not be negative"); }
    if (nsteps < 1 || nsteps > MAX_TYPES) { return fail(MCI_E_INVALID_ARGUMENT, "a trial move needs 1 to %d step sizes", MAX_TYPES); }
    if (nsteps > 1 && !te) { return fail(MCI_E_INVALID_ARGUMENT, "[checkTrialMoveSanity] ntypes>1 requires passed typeEnds array."); }
    if (nsteps > 1) {
        for (int i = 0; i < nsteps; ++i) {
            if (te[i] < 1 || te[i] > c->ndim || (i > 0 && te[i] <= te[i - 1])) { return fail(MCI_E_INVALID_ARGUMENT, "typeEnds must increase within [1, ndim]"); }
        }
        if (te[nsteps - 1] != c->ndim) { return fail(MCI_E_INVALID_ARGUMENT, "the last type end must equal ndim"); }
    }
    if (!(change_rate > 0. && change_rate <= 1.)) { return fail(MCI_E_INVALID_ARGUMENT, "the change rate must lie in (0, 1]"); }
    if (nparams < 0 || (nparams > 0 && !params)) { return fail(MCI_E_INVALID_ARGUMENT, "bad parameter array"); }
    TRY(set_current(c));
    c->typeEnds.assign(nsteps, c->ndim);
    if (nsteps > 1) { c->typeEnds.assign(te, te + nsteps); }
    c->ntypes = nsteps;
    c->srrd = MCI_SRRD_UNIFORM;
    c->movetype = MCI_MOVE_DEVICE;
    c->veclen = 0;
    c->nvecs = 0;
    c->ms_on = 0;
    c->ms_nsteps = 0;
    c->ms_pdfs.clear();
    c->lm_K = 0;
    c->lm_pmix = 0.;
    c->mv_name = name;
    c->mv_src = source;
    c->mv_nproto = nproto;
    c->mv_rate = change_rate;
    if (nparams > 0) {
        HIPCHK(c->mv_par.ensure((size_t)nparams));
        HIPCHK(hipMemcpy(c->mv_par.p, params, sizeof(double) * (size_t)nparams, hipMemcpyHostToDevice));
    }
    HIPCHK(c->mv_pold.ensure((size_t)std::max(nproto, 1) * c->W));
    HIPCHK(c->mv_pnew.ensure((size_t)std::max(nproto, 1) * c->W));
    HIPCHK(c->steps.ensure((size_t)nsteps * c->W));
    for (int t = 0; t < nsteps; ++t) { TRY(fill_rows(c, c->steps.p, t, 1, steps ? steps[t] : 0.05)); }
    return MCI_OK;
}

mci_status mci_build_device_functors(mci_ctx *c)
{
    CHECKCTX(c);
    TRY(set_current(c));
    return ensure_rtc(c);
}

mci_status mci_check_device_functor(const char *name, const char *source, int what, char *log, int64_t logcap)
{
    if (log && logcap > 0) { log[0] = '\0'; }
    if (!name || !source || !rtc_valid_name(name)) { return fail(MCI_E_INVALID_ARGUMENT, "a device functor needs a source and a C identifier as its name"); }
    if (what < 0 || what > 2) { return fail(MCI_E_DOMAIN, "what: 0 sampling function, 1 observable, 2 trial move"); }
    std::vector<RtcFunctor> fp, fo;
    const RtcFunctor mv{name, source};
    if (what == 0) { fp.push_back(mv); }
    if (what == 1) { fo.push_back(mv); }
    std::string text, msg, lg;
    if (!rtc_program(fp, fo, what == 2 ? &mv : nullptr, text, msg)) { return fail(MCI_E_INVALID_ARGUMENT, "%s", msg.c_str()); }
    const int st = rtc_check(text, "gfx950", lg);
    if (log && logcap > 0) {
        const size_t n = std::min<size_t>(lg.size(), (size_t)logcap - 1);
        std::memcpy(log, lg.data(), n);
        log[n] = '\0';
    }
    if (st != MCI_OK) { return fail(st, "the device functor does not compile:\n%s", lg.c_str()); }
    return MCI_OK;
}
mci_status mci_get_nobs(mci_ctx *c, int *n) { CHECKCTX(c); *n = (int)c->obs.size(); return MCI_OK; }
mci_status mci_get_nobsdim(mci_ctx *c, int *n) { CHECKCTX(c); *n = nobsdim(c); return MCI_OK; }

mci_status integrate_phases(mci_ctx *c, int64_t nmc, double *average, double *error, int do_find, int do_decorr);

// the callback calls of the last integrate, in order, from walker 0's record
mci_status replay_callback(mci_ctx *c)
{
    int64_t n = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(&n, c->cb_n.p, sizeof n, hipMemcpyDeviceToHost));
    const int64_t m = std::min(n, c->cb_cap), rw = c->ndim + 2;
    std::vector<double> h((size_t)std::max<int64_t>(m, 1) * rw);
    if (m > 0) { HIPCHK(hipMemcpy(h.data(), c->cb_buf.p, sizeof(double) * (size_t)(m * rw), hipMemcpyDeviceToHost)); }
    for (int64_t k = 0; k < m; ++k) {
        const double *r = &h[(size_t)(k * rw)];
        mci_callback_view v;
        v.call = k;
        v.ndim = c->ndim;
        v.x = r;
        v.acc = (int64_t)r[c->ndim];
        v.rej = (int64_t)r[c->ndim + 1];
        v.rate = (v.acc > 0) ? static_cast<double>(v.acc) / (static_cast<double>(v.acc) + v.rej) : 0.; // MCI::getAcceptanceRate (MCIntegrator.cpp:587-592)
        c->cb(c->cb_user, &v);
    }
    if (n > c->cb_cap) {
        return fail(MCI_E_RUNTIME, "the callback was due %lld times, its record holds %lld: the first %lld calls were made", (long long)n,
                    (long long)c->cb_cap, (long long)m);
    }
    return MCI_OK;
}

mci_status mci_integrate(mci_ctx *c, int64_t nmc, double *average, double *error, int do_find, int do_decorr)
{
    CHECKCTX(c);
    if (c->cb) {
        TRY(set_current(c));
        HIPCHK(hipMemsetAsync(c->cb_n.p, 0, sizeof(int64_t), c->stream));
    }
    TRY(integrate_phases(c, nmc, average, error, do_find, do_decorr));
    return c->cb ? replay_callback(c) : MCI_OK;
}

// MCI::setCallback / clearCallback (MCIntegrator.hpp:189-190)
mci_status mci_set_callback(mci_ctx *c, mci_callback_fn fn, void *user, int64_t capacity)
{
    CHECKCTX(c);
    if (!fn) { return fail(MCI_E_INVALID_ARGUMENT, "null callback (use mci_clear_callback)"); }
    if (capacity < 1 || (double)capacity * (c->ndim + 2) * 8.0 > 4e9) { return fail(MCI_E_INVALID_ARGUMENT, "callback record capacity must be in [1, 4 GB]"); }
    TRY(set_current(c));
    HIPCHK(c->cb_buf.ensure((size_t)capacity * (c->ndim + 2)));
    HIPCHK(c->cb_n.ensure(1));
    c->cb = fn;
    c->cb_user = user;
    c->cb_cap = capacity;
    return MCI_OK;
}

mci_status mci_clear_callback(mci_ctx *c)
{
    CHECKCTX(c);
    c->cb = nullptr;
    c->cb_user = nullptr;
    c->cb_cap = 0;
    return MCI_OK;
}

mci_status integrate_phases(mci_ctx *c, int64_t nmc, double *average, double *error, int do_find, int do_decorr)
{
    if (c->pdfs.empty() && !c->periodic) {
        return fail(MCI_E_DOMAIN, "[MCI::integrate] Integrating over an infinite domain requires a sampling function.");
    }
    const int nd = nobsdim(c);
    if (nmc > 0 && nd > 0 && (!average || !error)) { return fail(MCI_E_INVALID_ARGUMENT, "null output arrays"); }
    TRY(check_dependencies(c));
    // accumulator and estimator requirements (BlockAccumulator.cpp:8-13 and the estimators)
    if (nmc > 0) {
        for (auto &o : c->obs) {
            const int64_t naccu = 1 + (nmc - 1) / o.nskip;
            if (o.blocksize > 1) {
                if (naccu < o.blocksize) {
                    return fail(MCI_E_INVALID_ARGUMENT, "[BlockAccumulator::allocate] Requested number of accumulations is smaller than the requested block size.");
                }
                if (naccu % o.blocksize != 0) {
                    return fail(MCI_E_INVALID_ARGUMENT, "[BlockAccumulator::allocate] Requested number of accumulations is not a multiple of the requested block size.");
                }
            }
            TRY(check_estimator(o.estim, nstore_of(o.blocksize, o.nskip, nmc)));
        }
    }
    TRY(set_current(c));
    TRY(ensure_protos(c));
    TRY(ensure_rng(c));
    TRY(ensure_rtc(c));
    const int64_t W = c->W;
    HIPCHK(c->obsval.ensure((size_t)std::max(nd, 1) * W));
    for (double &t : c->timings) { t = 0.; }
    const bool fast = fast_chain(c);
    if (c->prec == MCI_PREC_FP32 && !fast) {
        return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision supports Philox streams with single-coordinate uniform moves on one diagonal "
                                            "Gaussian (explicit weights) in an unbounded domain");
    }
    std::vector<ObsDev> od;
    c->nstore_last.clear();
    bool fast_main = false;
    if (nmc > 0) {
        // ---- accumulators (MCIntegrator.cpp:54-57), allocated before any work so that a
        //      configuration the kernels cannot run fails before calibration ----
        if (c->accdata.size() < c->obs.size()) { c->accdata.resize(c->obs.size()); c->scratch.resize(c->obs.size()); }
        int voff = 0;
        for (size_t i = 0; i < c->obs.size(); ++i) {
            const ObsHost &o = c->obs[i];
            ObsDev d;
            std::memset(&d, 0, sizeof d);
            d.kind = o.kind;
            d.nobs = o.nobs;
            d.voff = voff;
            d.nskip = o.nskip;
            d.acc = (o.blocksize == 0) ? ACC_SIMPLE : (o.blocksize == 1 ? ACC_FULL : ACC_BLOCK);
            d.blocksize = std::max(o.blocksize, 1);
            d.p0 = obs_p0(o, c->ndim);
            d.uid = obs_uid(c, i);
            d.par = o.dev.p;
            d.nstore = nstore_of(o.blocksize, o.nskip, nmc);
            HIPCHK(c->accdata[i].ensure((size_t)d.nstore * o.nobs * W));
            d.data = c->accdata[i].p;
            if (needs_scratch(o.estim, d.nstore)) { HIPCHK(c->scratch[i].ensure((size_t)mj_scratch_rows(d.nstore) * o.nobs * W)); }
            od.push_back(d);
            voff += o.nobs;
        }
        fast_main = fast && fast_obs_eligible(od.data(), (int)od.size(), c->ndim, c->prec, c->movetype, c->rng);
        if (c->prec == MCI_PREC_FP32 && !fast_main) {
            return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision supports X2, X2_SUM and EXP_X2 observables (one each, one block size, Block or Simple accumulators, nskip 1)");
        }
        if (c->prec == MCI_PREC_FP32 && nmc > INT32_MAX) {
            for (const ObsDev &d : od) {
                if (d.acc == ACC_SIMPLE) { return fail(MCI_E_INVALID_ARGUMENT, "FP32 precision accumulates at most 2^31 - 1 steps per block (Simple accumulator: per run)"); }
            }
        }
    }
    // ---- file output of the main sampling (MCIntegrator.cpp:61-65, 495-542): the generic kernel
    //      records the values every freq steps, the host writes the files afterwards ----
    const bool rec_obs = nmc > 0 && !c->obsfile.empty(), rec_x = nmc > 0 && !c->wlkfile.empty();
    if (rec_obs || rec_x) {
        if (c->prec == MCI_PREC_FP32) { return fail(MCI_E_INVALID_ARGUMENT, "file output is written from the FP64 chain"); }
        fast_main = false;
        const int64_t no = rec_obs ? (nmc - 1) / c->obsfreq + 1 : 0, nx = rec_x ? (nmc - 1) / c->wlkfreq + 1 : 0;
        const double bytes = 8.0 * (double)W * ((double)no * std::max(nd, 1) + (double)nx * c->ndim);
        if (bytes > 16e9) {
            return fail(MCI_E_INVALID_ARGUMENT, "file output would record %.3g GB on the device (at most 16 GB): raise the frequency or use fewer walkers",
                        bytes / 1e9);
        }
        if (rec_obs) { HIPCHK(c->rec_obs.ensure((size_t)no * std::max(nd, 1) * W)); }
        if (rec_x) { HIPCHK(c->rec_x.ensure((size_t)nx * c->ndim * W)); }
    }
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    // ---- calibration (MCIntegrator.cpp:49-51) ----
    if (!c->pdfs.empty() && do_find && c->ntypes > 0) {
        if (c->coupling == MCI_COUPLING_SHARED && W > 1) { TRY(find_shared(c)); }
        else {
            const int64_t min_stat = (int64_t)std::fmax(100., std::sqrt(40000. * c->ndim));
            if (fast_chain(c)) { TRY(find_indep_fast(c, min_stat)); }
            else { HIPCHK(calibrate_indep_launch(c, make_chain(c, nullptr), c->nfind, c->target, min_stat, c->dimsizes.p, c->stream)); }
        }
    }
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    // ---- initial decorrelation (MCIntegrator.cpp:171-247) ----
    if (!c->pdfs.empty() && do_decorr) {
        if (c->ndecorr > 0) {
            if (fast) {
                TRY(mt_align_if_needed(c));
                TRY(fast_sample(c->fast, make_chain(c, nullptr), c->prec, c->ndecorr, nullptr, 0, c->stream));
            }
            else { HIPCHK(sample_launch(c, make_chain(c, nullptr), c->ndecorr, 0, nullptr, c->stream)); }
        }
        else if (c->ndecorr < 0) { TRY(decorrelate_auto(c)); }
    }
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (nmc <= 0) {
        HIPCHK(hipStreamSynchronize(c->stream));
        c->timings[0] = elapsed(c, 0, 1);
        c->timings[1] = elapsed(c, 1, 2);
        c->timings[4] = elapsed(c, 0, 2);
        return MCI_OK;
    }
    // ---- sample (MCIntegrator.cpp:58-62) ----
    for (auto &d : od) { c->nstore_last.push_back(d.nstore); }
    Chain ch = make_chain(c, &od);
    if (rec_obs) {
        ch.rec_obs = c->rec_obs.p;
        ch.rec_fobs = c->obsfreq;
        ch.rec_nobsdim = nd;
    }
    if (rec_x) {
        ch.rec_x = c->rec_x.p;
        ch.rec_fx = c->wlkfreq;
    }
    if (fast_main) {
        TRY(mt_align_if_needed(c));
        TRY(fast_sample(c->fast, ch, c->prec, nmc, od.data(), (int)od.size(), c->stream));
    }
    else { HIPCHK(sample_launch(c, ch, nmc, 1, nullptr, c->stream)); }
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    if (rec_obs || rec_x) { TRY(write_files(c, nmc, nd, rec_obs, rec_x)); }
    // ---- estimate (ObservableContainer.cpp:104-111), volume scaling (:71-77), combine ----
    double vol = 1.0;
    if (c->pdfs.empty()) {
        for (int i = 0; i < c->ndim; ++i) { vol = vol * (c->ub[i] - c->lb[i]); }
    }
    const size_t ndw = (size_t)std::max(nd, 1) * W;
    HIPCHK(c->wavg.ensure(ndw));
    HIPCHK(c->werr.ensure(ndw));
    for (size_t i = 0; i < od.size(); ++i) {
        EstimJob j;
        j.data = od[i].data;
        j.scratch = c->scratch[i].p;
        j.n = od[i].nstore;
        j.W = W;
        j.nobs = od[i].nobs;
        j.off = od[i].voff;
        j.estim = c->obs[i].estim;
        j.scale = vol;
        j.active = nullptr;
        j.avg = c->wavg.p;
        j.err = c->werr.p;
        HIPCHK(launch_estimate(j, c->stream));
    }
    c->walker_results_host = false;
    if (nd > 0) {
        if (W > 1) {
            const int64_t nch = (W + 255) / 256;
            HIPCHK(c->chunks_a.ensure((size_t)nch * nd));
            HIPCHK(c->chunks_e.ensure((size_t)nch * nd));
            HIPCHK(c->res.ensure(2 * (size_t)nd));
            HIPCHK(launch_chunk_sum(c->wavg.p, W, nd, 0, c->chunks_a.p, c->stream));
            HIPCHK(launch_chunk_sum(c->werr.p, W, nd, 1, c->chunks_e.p, c->stream));
            HIPCHK(launch_combine(c->chunks_a.p, c->chunks_e.p, W, nd, c->res.p, c->res.p + nd, c->stream));
            HIPCHK(hipMemcpyAsync(average, c->res.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(error, c->res.p + nd, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
        else {
            HIPCHK(hipMemcpyAsync(average, c->wavg.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(error, c->werr.p, nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(hipEventRecord(c->ev[4], c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nobsdim_last = nd;
    c->timings[0] = elapsed(c, 0, 1);
    c->timings[1] = elapsed(c, 1, 2);
    c->timings[2] = elapsed(c, 2, 3);
    c->timings[3] = elapsed(c, 3, 4);
    c->timings[4] = elapsed(c, 0, 4);
    return MCI_OK;
}

mci_status mci_get_walker_results(mci_ctx *c, double *a, double *e, int64_t n)
{
    CHECKCTX(c);
    const int nd = c->nobsdim_last;
    if (!a || !e || n < c->W * nd) { return fail(MCI_E_INVALID_ARGUMENT, "output too small"); }
    if (!c->walker_results_host && nd > 0) {
        TRY(set_current(c));
        c->h_wavg.assign((size_t)nd * c->W, 0.);
        c->h_werr.assign((size_t)nd * c->W, 0.);
        HIPCHK(hipMemcpyAsync(c->h_wavg.data(), c->wavg.p, (size_t)nd * c->W * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_werr.data(), c->werr.p, (size_t)nd * c->W * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->walker_results_host = true;
    }
    for (int64_t w = 0; w < c->W; ++w) {
        for (int i = 0; i < nd; ++i) {
            a[w * nd + i] = c->h_wavg[(size_t)i * c->W + w];
            e[w * nd + i] = c->h_werr[(size_t)i * c->W + w];
        }
    }
    return MCI_OK;
}

mci_status mci_get_observable_data(mci_ctx *c, int obs, double *data, int64_t n, int64_t *nstore)
{
    CHECKCTX(c);
    if (obs < 0 || obs >= (int)c->nstore_last.size() || obs >= (int)c->obs.size()) {
        return fail(MCI_E_INVALID_ARGUMENT, "no data for observable %d", obs);
    }
    const int64_t ns = c->nstore_last[obs];
    const int nobs = c->obs[obs].nobs;
    if (nstore) { *nstore = ns; }
    if (!data) { return MCI_OK; }
    if (n < ns * nobs * c->W) { return fail(MCI_E_INVALID_ARGUMENT, "output too small"); }
    TRY(set_current(c));
    std::vector<double> h((size_t)ns * nobs * c->W);
    HIPCHK(hipMemcpyAsync(h.data(), c->accdata[obs].p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t w = 0; w < c->W; ++w) {
        for (int64_t r = 0; r < ns * nobs; ++r) { data[(w * ns * nobs) + r] = h[(size_t)r * c->W + w]; }
    }
    return MCI_OK;
}

mci_status mci_get_acceptance(mci_ctx *c, int64_t *acc, int64_t *rej, double *rate)
{
    CHECKCTX(c);
    TRY(set_current(c));
    std::vector<int64_t> a(c->W), r(c->W);
    HIPCHK(hipMemcpyAsync(a.data(), c->acc.p, c->W * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(r.data(), c->rej.p, c->W * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int64_t sa = 0, sr = 0;
    std::vector<double> rates(c->W);
    for (int64_t w = 0; w < c->W; ++w) {
        sa += a[w];
        sr += r[w];
        rates[w] = (a[w] > 0) ? (double)a[w] / ((double)a[w] + (double)r[w]) : 0.;
    }
    if (acc) { *acc = sa; }
    if (rej) { *rej = sr; }
    if (rate) { *rate = (c->W == 1) ? rates[0] : det_sum(rates.data(), c->W) / (double)c->W; }
    return MCI_OK;
}

mci_status mci_get_walker_acceptance(mci_ctx *c, int64_t *acc, int64_t *rej, int64_t n)
{
    CHECKCTX(c);
    if (!acc || !rej || n < c->W) { return fail(MCI_E_INVALID_ARGUMENT, "output too small"); }
    TRY(set_current(c));
    HIPCHK(hipMemcpyAsync(acc, c->acc.p, c->W * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(rej, c->rej.p, c->W * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

// ---- checkpoint of the walkers: positions, step sizes, random streams, saved normals ----
namespace
{
struct StateHeader
{
    uint32_t magic, version;
    int32_t ndim, ntypes;
    int64_t W;
    int32_t rng, reserved;
    uint64_t seed;
};
constexpr uint32_t STATE_MAGIC = 0x5349434Du; // "MCIS"

int64_t state_bytes(const mci_ctx *c)
{
    const int64_t W = c->W;
    int64_t n = (int64_t)sizeof(StateHeader) + 8 * W * (c->ndim + c->ntypes) + 8 * 2 * W + 4 * W;
    n += (c->rng == MCI_RNG_MT19937_64) ? 8 * (int64_t)MT_N * W + 4 * W : 8 * W;
    return n;
}

// the sections of the state after the header, in order: device buffer, bytes
std::vector<std::pair<void *, size_t>> state_sections(mci_ctx *c)
{
    const size_t W = (size_t)c->W;
    std::vector<std::pair<void *, size_t>> v = {{c->x.p, 8 * W * c->ndim}, {c->steps.p, 8 * W * c->ntypes}};
    if (c->rng == MCI_RNG_MT19937_64) {
        v.push_back({c->mt.p, 8 * W * MT_N});
        v.push_back({c->mtp.p, 4 * W});
    }
    else { v.push_back({c->ctr.p, 8 * W}); }
    v.push_back({c->nsv.p, 8 * 2 * W});
    v.push_back({c->nsf.p, 4 * W});
    return v;
}
} // namespace

mci_status mci_get_state_size(mci_ctx *c, int64_t *nbytes)
{
    CHECKCTX(c);
    if (!nbytes) { return fail(MCI_E_INVALID_ARGUMENT, "null output"); }
    *nbytes = state_bytes(c);
    return MCI_OK;
}

mci_status mci_get_state(mci_ctx *c, void *buf, int64_t nbytes)
{
    CHECKCTX(c);
    if (!buf || nbytes < state_bytes(c)) { return fail(MCI_E_INVALID_ARGUMENT, "state buffer too small (%lld bytes needed)", (long long)state_bytes(c)); }
    TRY(set_current(c));
    TRY(ensure_rng(c));
    StateHeader h{STATE_MAGIC, 1u, c->ndim, c->ntypes, c->W, c->rng, 0, c->seed};
    char *p = static_cast<char *>(buf);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    for (auto &sec : state_sections(c)) {
        HIPCHK(hipMemcpyAsync(p, sec.first, sec.second, hipMemcpyDeviceToHost, c->stream));
        p += sec.second;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return MCI_OK;
}

mci_status mci_set_state(mci_ctx *c, const void *buf, int64_t nbytes)
{
    CHECKCTX(c);
    StateHeader h;
    if (!buf || nbytes < (int64_t)sizeof h) { return fail(MCI_E_INVALID_ARGUMENT, "not a walker state"); }
    std::memcpy(&h, buf, sizeof h);
    if (h.magic != STATE_MAGIC || h.version != 1u) { return fail(MCI_E_INVALID_ARGUMENT, "not a walker state of this library"); }
    if (h.ndim != c->ndim || h.W != c->W || h.ntypes != c->ntypes || h.rng != c->rng) {
        return fail(MCI_E_INVALID_ARGUMENT, "the state holds %d dimensions, %lld walkers, %d step sizes and random mode %d; the context %d, %lld, %d, %d",
                    h.ndim, (long long)h.W, h.ntypes, h.rng, c->ndim, (long long)c->W, c->ntypes, c->rng);
    }
    if (nbytes < state_bytes(c)) { return fail(MCI_E_INVALID_ARGUMENT, "truncated walker state"); }
    TRY(set_current(c));
    c->seed = h.seed;  // the Philox key
    if (c->rng == MCI_RNG_MT19937_64) {
        HIPCHK(c->mt.ensure((size_t)MT_N * c->W));
        HIPCHK(c->mtp.ensure(c->W));
    }
    else { HIPCHK(c->ctr.ensure(c->W)); }
    const char *p = static_cast<const char *>(buf) + sizeof h;
    for (auto &sec : state_sections(c)) {
        HIPCHK(hipMemcpyAsync(sec.first, p, sec.second, hipMemcpyHostToDevice, c->stream));
        p += sec.second;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mt_diverged = true;
    c->reseed = false;  // the streams continue from the state
    return MCI_OK;
}

mci_status mci_get_timings(mci_ctx *c, double *ms, int n)
{
    CHECKCTX(c);
    for (int i = 0; i < n && i < 5; ++i) { ms[i] = c->timings[i]; }
    return MCI_OK;
}

mci_status mci_estimate(int estim, int64_t n, int ndim, const double *x, double *average, double *error)
{
    if (!x || !average || !error || ndim < 1) { return fail(MCI_E_INVALID_ARGUMENT, "bad arguments"); }
    if (estim < MCI_ESTIM_NOOP || estim > MCI_ESTIM_MJBLOCKER) { return fail(MCI_E_DOMAIN, "[createEstimator] Unhandled estimator enumerator."); }
    if (n < 1) { return fail(MCI_E_INVALID_ARGUMENT, "n must be at least 1"); }
    TRY(check_estimator(estim, n));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { return fail(MCI_E_DEVICE, "no HIP device available"); }
    DBuf<double> d, s, a, e;
    HIPCHK(d.ensure((size_t)n * ndim));
    HIPCHK(s.ensure((size_t)mj_scratch_rows(n) * ndim));
    HIPCHK(a.ensure(ndim));
    HIPCHK(e.ensure(ndim));
    HIPCHK(hipMemcpy(d.p, x, sizeof(double) * n * ndim, hipMemcpyHostToDevice));
    EstimJob j;
    j.data = d.p;
    j.scratch = s.p;
    j.n = n;
    j.W = 1;
    j.nobs = ndim;
    j.off = 0;
    j.estim = estim;
    j.scale = 1.0;
    j.active = nullptr;
    j.avg = a.p;
    j.err = e.p;
    HIPCHK(launch_estimate(j, nullptr));
    HIPCHK(hipMemcpy(average, a.p, sizeof(double) * ndim, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(error, e.p, sizeof(double) * ndim, hipMemcpyDeviceToHost));
    return MCI_OK;
}

} // extern "C"
