// mci_batch_api.cpp — C-ABI of the batched-integral path (include/mci_amd.h, mci_batch_*):
// parameter preparation, the plan autotuner, and one submission = calibrate + persistent
// sampling + MJBlocker per walker + per-integral combine on the batch's stream, optionally
// captured once into a hipGraph and replayed.
#include "mci_batch.hpp"
#include "mci_host.hpp"
#include "mci_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

using mcih::DBuf;
using mcih::fail;

namespace
{
constexpr int MAXK = mcib::MAXK;
// launches are split by dimension: a wave holds 64 d floats of LDS, so small-d items can keep
// more waves resident per CU (the limits match 16, 8 and 5+ waves per CU of 160 KiB)
constexpr int NCLASS = 3;
constexpr int CLASS_DMAX[NCLASS] = {40, 80, MCI_BATCH_MAX_DIM};
int dim_class(int d)
{
    int c = 0;
    while (c < NCLASS - 1 && d > CLASS_DMAX[c]) { ++c; }
    return c;
}

// cost model of k_batch_sample in instructions issued per wave: per step a fixed part (Philox,
// index draw, proposal, test, Kahan add) and a part per component (two fused ops, exp); the
// block-end refresh adds about 3 (K + 1) d per block; a walker's exact start draw costs about
// 60 d (polar method) plus one refresh
struct CostModel
{
    double base = 110., per_comp = 36., refresh = 3., init_per_dim = 60.;
    double cycles_per_instr_alone = 5.6; // one wave on its SIMD: dependent issue
    double cycles_per_instr_busy = 1.25; // several waves per SIMD: issue-bound
    double clock_hz = 2.4e9;
};

int64_t pow2_mult_floor(int64_t x, int64_t unit)
{
    // largest unit * 2^k <= x (at least unit)
    int64_t v = unit;
    while (v * 2 <= x) { v *= 2; }
    return v;
}
} // namespace

struct mci_batch
{
    int nint = 0;
    std::vector<int> d, K;
    std::vector<int64_t> nmc, coff, moff;
    std::vector<double> w, s, mu;
    uint64_t seed = 1;
    int nblk = 16, rmin = 64, mode = MCI_BATCH_ERR_SPREAD, nfind = -50;
    double target = 0.5;
    int64_t ndecorr = 0;
    bool user_plan = false, planned = false;
    std::vector<int64_t> T, R, woff;
    std::vector<int> order;
    int64_t nwalkers = 0;
    int dmax = 1;
    std::vector<float> lw, h, sig, cw, muf;
    std::vector<int64_t> muoff;
    std::vector<double> step0;
    CostModel cost;

    bool dev_ready = false, params_up = false, plan_up = false;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    // graph form: the dimension classes' launches on side branches (fork/join events per phase)
    hipStream_t side[NCLASS - 1] = {};
    hipEvent_t fork_ev[2] = {}, join_ev[2][NCLASS - 1] = {};
    DBuf<int> d_nd, d_nc, d_wint;
    DBuf<float> d_lw, d_h, d_sig, d_cw, d_mu;
    DBuf<int64_t> d_moff, d_T, d_woff, d_R;
    DBuf<double> d_step0, d_step, d_rate, d_avg, d_err, d_accr, d_bm, d_wavg, d_werr;
    DBuf<int32_t> d_acc;
    DBuf<unsigned int> d_queue;
    DBuf<int> d_items, d_ilist;
    int item_off[NCLASS + 1] = {}, int_off[NCLASS + 1] = {};
    int item_dmax[NCLASS] = {}, int_dmax[NCLASS] = {}, nwaves[NCLASS] = {};

    bool graph_on = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double timings[4] = {0, 0, 0, 0};

    void drop_graph()
    {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    }
    ~mci_batch()
    {
        drop_graph();
        for (auto &e : ev) { if (e) { (void)hipEventDestroy(e); } }
        for (auto &e : fork_ev) { if (e) { (void)hipEventDestroy(e); } }
        for (auto &row : join_ev) {
            for (auto &e : row) { if (e) { (void)hipEventDestroy(e); } }
        }
        for (auto &q : side) { if (q) { (void)hipStreamDestroy(q); } }
        if (stream) { (void)hipStreamDestroy(stream); }
    }
};

namespace
{
#define CHECKB(b)                                                                                        \
    do {                                                                                                 \
        if (!(b)) { return fail(MCI_E_INVALID_ARGUMENT, "null batch"); }                                 \
    } while (0)

// FP32 parameters of each integral (normalised weights; the oracle derives them the same way)
void derive(mci_batch *b)
{
    b->lw.assign((size_t)b->nint * MAXK, 0.f);
    b->h.assign((size_t)b->nint * MAXK, 0.f);
    b->sig.assign((size_t)b->nint * MAXK, 0.f);
    b->cw.assign((size_t)b->nint * MAXK, 1.f);
    b->step0.assign(b->nint, 0.);
    b->muf.resize(b->mu.size());
    for (size_t q = 0; q < b->mu.size(); ++q) { b->muf[q] = (float)b->mu[q]; }
    b->muoff = b->moff;
    for (int i = 0; i < b->nint; ++i) {
        const int K = b->K[i];
        const double *w = b->w.data() + b->coff[i], *s = b->s.data() + b->coff[i];
        double wsum = 0.;
        for (int k = 0; k < K; ++k) { wsum = wsum + w[k]; }
        double cum = 0., st = 0.;
        for (int k = 0; k < K; ++k) {
            const double wn = w[k] / wsum;
            b->lw[i * MAXK + k] = (float)(log(wn) - 0.5 * (double)b->d[i] * log(2.0 * M_PI * s[k] * s[k]));
            b->h[i * MAXK + k] = (float)(1.0 / (2.0 * s[k] * s[k]));
            b->sig[i * MAXK + k] = (float)s[k];
            cum = cum + wn;
            b->cw[i * MAXK + k] = (k == K - 1) ? 1.0f : (float)cum;
            st = st + wn * s[k];
        }
        // a uniform single-coordinate move of half-width 2.93 s is accepted half the time on a
        // Gaussian of width s: start calibration there (with the weighted mean width)
        b->step0[i] = 2.93 * st;
    }
}

struct DeviceShape
{
    int ncu = 256;
    int waves_per_cu = 0;
};

DeviceShape device_shape(mci_batch *b)
{
    DeviceShape sh;
    if (b->dev_ready) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, b->device) == hipSuccess) { sh.ncu = p.multiProcessorCount; }
        sh.waves_per_cu = mcib::batch_sample_waves_per_cu(b->dmax);
    }
    if (sh.waves_per_cu <= 0) {
        // 160 KiB of LDS per CU, at most 8 waves per SIMD
        sh.waves_per_cu = std::max(1, std::min(32, (int)((160 * 1024) / mcib::batch_lds_bytes(b->dmax))));
    }
    return sh;
}

// walkers, offsets and walker order for steps-per-walker T (integral order)
void apply_plan(mci_batch *b, const std::vector<int64_t> &T)
{
    b->T = T;
    b->R.assign(b->nint, 0);
    for (int i = 0; i < b->nint; ++i) { b->R[i] = (b->nmc[i] + T[i] - 1) / T[i]; }
    b->order.resize(b->nint);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        if (b->T[x] != b->T[y]) { return b->T[x] > b->T[y]; }
        if (b->d[x] != b->d[y]) { return b->d[x] > b->d[y]; }
        return b->K[x] > b->K[y];
    });
    b->woff.assign(b->nint, 0);
    int64_t off = 0;
    for (int i : b->order) {
        b->woff[i] = off;
        off += b->R[i];
    }
    b->nwalkers = off;
    b->planned = true;
    b->plan_up = false;
    b->drop_graph();
}

// model time (s) of the sampling launch for a plan: the longer of the slowest item run alone
// and the total work spread over every SIMD
double model_seconds(const mci_batch *b, const std::vector<int64_t> &T, const DeviceShape &sh)
{
    const CostModel &c = b->cost;
    std::vector<int> ord(b->nint);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        if (T[x] != T[y]) { return T[x] > T[y]; }
        if (b->d[x] != b->d[y]) { return b->d[x] > b->d[y]; }
        return b->K[x] > b->K[y];
    });
    double total = 0., longest = 0.;
    int64_t lanes = 0;
    int64_t steps = 0;
    double instr = 0., init = 0.;
    auto close_item = [&]() {
        const double cyc = (double)steps * instr + init;
        total += cyc;
        longest = std::max(longest, cyc);
        lanes = 0;
        steps = 0;
        instr = 0.;
        init = 0.;
    };
    for (int i : ord) {
        const int64_t R = (b->nmc[i] + T[i] - 1) / T[i];
        const int64_t B = T[i] / b->nblk;
        const double ins = c.base + c.per_comp * b->K[i] + c.refresh * (b->K[i] + 1) * b->d[i] / (double)B;
        for (int64_t r = 0; r < R;) {
            const int64_t take = std::min<int64_t>(R - r, 64 - lanes);
            steps = std::max<int64_t>(steps, T[i] + b->ndecorr);
            instr = std::max(instr, ins);
            init = std::max(init, c.init_per_dim * b->d[i] + c.refresh * (b->K[i] + 1) * b->d[i]);
            lanes += take;
            r += take;
            if (lanes == 64) { close_item(); }
        }
    }
    if (lanes > 0) { close_item(); }
    const double nsimd = 4.0 * sh.ncu;
    const double alone = longest * c.cycles_per_instr_alone;
    const double busy = total * c.cycles_per_instr_busy / nsimd;
    return std::max(alone, busy) / c.clock_hz;
}

std::vector<int64_t> plan_for_cap(const mci_batch *b, int64_t cap)
{
    std::vector<int64_t> T(b->nint);
    for (int i = 0; i < b->nint; ++i) {
        const int64_t want = std::max<int64_t>(b->nmc[i] / b->rmin, b->nblk);
        T[i] = std::min(cap, pow2_mult_floor(want, b->nblk));
    }
    return T;
}

// candidate caps ranked by the model (best first)
std::vector<std::pair<double, int64_t>> ranked_caps(const mci_batch *b, const DeviceShape &sh)
{
    std::vector<std::pair<double, int64_t>> out;
    int64_t maxT = b->nblk;
    for (int i = 0; i < b->nint; ++i) { maxT = std::max(maxT, pow2_mult_floor(std::max<int64_t>(b->nmc[i], b->nblk), b->nblk)); }
    for (int64_t cap = b->nblk; cap <= maxT; cap *= 2) { out.push_back({model_seconds(b, plan_for_cap(b, cap), sh), cap}); }
    std::stable_sort(out.begin(), out.end(), [](const std::pair<double, int64_t> &x, const std::pair<double, int64_t> &y) {
        if (x.first != y.first) { return x.first < y.first; }
        return x.second > y.second;
    });
    return out;
}

mci_status ensure_device(mci_batch *b)
{
    if (b->dev_ready) { return MCI_OK; }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) { return fail(MCI_E_DEVICE, "no HIP device available"); }
    HIPCHK(hipGetDevice(&b->device));
    HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto &e : b->ev) { HIPCHK(hipEventCreate(&e)); }
    for (auto &q : b->side) { HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking)); }
    for (auto &e : b->fork_ev) { HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    for (auto &row : b->join_ev) {
        for (auto &e : row) { HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    }
    b->dev_ready = true;
    return MCI_OK;
}

// setup copies are synchronous: the host vectors are temporaries
template <class T>
mci_status upload(DBuf<T> &dst, const std::vector<T> &src, hipStream_t s)
{
    HIPCHK(dst.ensure(src.size()));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipMemcpy(dst.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return MCI_OK;
}

mci_status ensure_uploaded(mci_batch *b)
{
    TRY(ensure_device(b));
    HIPCHK(hipSetDevice(b->device));
    if (!b->params_up) {
        TRY(upload(b->d_nd, b->d, b->stream));
        TRY(upload(b->d_nc, b->K, b->stream));
        TRY(upload(b->d_lw, b->lw, b->stream));
        TRY(upload(b->d_h, b->h, b->stream));
        TRY(upload(b->d_sig, b->sig, b->stream));
        TRY(upload(b->d_cw, b->cw, b->stream));
        TRY(upload(b->d_mu, b->muf, b->stream));
        TRY(upload(b->d_moff, b->muoff, b->stream));
        TRY(upload(b->d_step0, b->step0, b->stream));
        HIPCHK(b->d_step.ensure(b->nint));
        HIPCHK(b->d_rate.ensure(b->nint));
        HIPCHK(b->d_avg.ensure(b->nint));
        HIPCHK(b->d_err.ensure(b->nint));
        HIPCHK(b->d_accr.ensure(b->nint));
        HIPCHK(b->d_queue.ensure(NCLASS));
        // integrals by dimension class for the calibration launches, each class sorted by
        // component count and dimension (the groups of one wave then mostly take one path)
        std::vector<int> il;
        for (int c = 0; c < NCLASS; ++c) {
            b->int_off[c] = (int)il.size();
            b->int_dmax[c] = 1;
            for (int i = 0; i < b->nint; ++i) {
                if (dim_class(b->d[i]) == c) {
                    il.push_back(i);
                    b->int_dmax[c] = std::max(b->int_dmax[c], b->d[i]);
                }
            }
            std::stable_sort(il.begin() + b->int_off[c], il.end(), [&](int x, int y) {
                if (b->K[x] != b->K[y]) { return b->K[x] < b->K[y]; }
                return b->d[x] < b->d[y];
            });
        }
        b->int_off[NCLASS] = (int)il.size();
        TRY(upload(b->d_ilist, il, b->stream));
        b->params_up = true;
    }
    if (!b->planned) {
        const DeviceShape sh = device_shape(b);
        apply_plan(b, plan_for_cap(b, ranked_caps(b, sh).front().second));
    }
    if (!b->plan_up) {
        std::vector<int> wint((size_t)b->nwalkers);
        for (int i = 0; i < b->nint; ++i) {
            for (int64_t r = 0; r < b->R[i]; ++r) { wint[(size_t)(b->woff[i] + r)] = i; }
        }
        TRY(upload(b->d_wint, wint, b->stream));
        TRY(upload(b->d_T, b->T, b->stream));
        TRY(upload(b->d_woff, b->woff, b->stream));
        TRY(upload(b->d_R, b->R, b->stream));
        HIPCHK(b->d_bm.ensure((size_t)b->nblk * b->nwalkers));
        HIPCHK(b->d_wavg.ensure(b->nwalkers));
        HIPCHK(b->d_werr.ensure(b->nwalkers));
        HIPCHK(b->d_acc.ensure(b->nwalkers));
        // items (64 consecutive walkers) by the dimension class of their largest d
        const int64_t nitems = (b->nwalkers + mcib::WAVE - 1) / mcib::WAVE;
        std::vector<int> idmax((size_t)nitems, 1);
        for (int64_t g = 0; g < b->nwalkers; ++g) {
            int &v = idmax[(size_t)(g / mcib::WAVE)];
            v = std::max(v, b->d[wint[(size_t)g]]);
        }
        std::vector<int> items;
        const DeviceShape sh = device_shape(b);
        for (int c = 0; c < NCLASS; ++c) {
            b->item_off[c] = (int)items.size();
            b->item_dmax[c] = 1;
            for (int64_t q = 0; q < nitems; ++q) {
                if (dim_class(idmax[(size_t)q]) == c) {
                    items.push_back((int)q);
                    b->item_dmax[c] = std::max(b->item_dmax[c], idmax[(size_t)q]);
                }
            }
            const int64_t n = (int64_t)items.size() - b->item_off[c];
            int wpc = b->dev_ready ? mcib::batch_sample_waves_per_cu(b->item_dmax[c]) : 0;
            if (wpc <= 0) { wpc = sh.waves_per_cu; }
            b->nwaves[c] = (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)sh.ncu * wpc));
        }
        b->item_off[NCLASS] = (int)items.size();
        TRY(upload(b->d_items, items, b->stream));
        b->plan_up = true;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return MCI_OK;
}

mcib::BatchParams params(const mci_batch *b)
{
    mcib::BatchParams P;
    P.nint = b->nint;
    P.ndim = b->d_nd.p;
    P.ncomp = b->d_nc.p;
    P.lw = b->d_lw.p;
    P.h = b->d_h.p;
    P.sig = b->d_sig.p;
    P.cw = b->d_cw.p;
    P.mu = b->d_mu.p;
    P.mu_off = b->d_moff.p;
    P.key0 = (uint32_t)b->seed;
    P.key1 = (uint32_t)(b->seed >> 32);
    return P;
}

// The phases of one submission.  Timed (events, eager): every launch on b->stream in order.
// Captured (graph): the dimension classes' calibration launches run as parallel branches, and
// so do their sampling launches, so that the classes' tails overlap.  A sampling item's class
// is that of the largest d among its walkers, so it may hold walkers of lower classes: all
// calibrations join before any sampling starts.  Every class writes its own integrals' steps
// and its own items' results, so the order changes no number.
// Joins the side streams of a fork back into b->stream if the enqueue returns between fork and
// join (a failed launch): during capture, side streams that never join leave the capture
// unterminated.  Errors of the join itself are ignored; the failure being returned is reported.
struct ForkGuard
{
    mci_batch *b = nullptr;
    int k = -1;
    ~ForkGuard()
    {
        if (k < 0) { return; }
        for (int c = 0; c < NCLASS - 1; ++c) {
            (void)hipEventRecord(b->join_ev[k][c], b->side[c]);
            (void)hipStreamWaitEvent(b->stream, b->join_ev[k][c], 0);
        }
    }
};

mci_status enqueue(mci_batch *b, bool events)
{
    const mcib::BatchParams P = params(b);
    const bool branches = !events;
    ForkGuard open;
    auto lane_stream = [&](int c) { return (branches && c > 0) ? b->side[c - 1] : b->stream; };
    auto fork = [&](int k) -> mci_status {
        if (!branches) { return MCI_OK; }
        HIPCHK(hipEventRecord(b->fork_ev[k], b->stream));
        open.b = b;
        open.k = k;
        for (auto &q : b->side) { HIPCHK(hipStreamWaitEvent(q, b->fork_ev[k], 0)); }
        return MCI_OK;
    };
    auto join = [&](int k) -> mci_status {
        if (!branches) { return MCI_OK; }
        open.k = -1;
        for (int c = 0; c < NCLASS - 1; ++c) {
            HIPCHK(hipEventRecord(b->join_ev[k][c], b->side[c]));
            HIPCHK(hipStreamWaitEvent(b->stream, b->join_ev[k][c], 0));
        }
        return MCI_OK;
    };
    if (events) { HIPCHK(hipEventRecord(b->ev[0], b->stream)); }
    HIPCHK(mcib::launch_batch_zero(b->d_queue.p, NCLASS, b->stream));
    TRY(fork(0));
    for (int c = 0; c < NCLASS; ++c) {
        const int n = b->int_off[c + 1] - b->int_off[c];
        if (n == 0) { continue; }
        mcib::CalibArgs ca{P, b->d_ilist.p + b->int_off[c], b->nfind, b->target, b->d_step0.p, b->d_step.p, b->d_rate.p};
        HIPCHK(mcib::launch_batch_calibrate(ca, n, b->int_dmax[c], lane_stream(c)));
    }
    TRY(join(0));
    if (events) { HIPCHK(hipEventRecord(b->ev[1], b->stream)); }
    TRY(fork(1));
    mcib::SampleArgs sa;
    sa.P = P;
    sa.step = b->d_step.p;
    sa.wint = b->d_wint.p;
    sa.wsteps = b->d_T.p;
    sa.nwalkers = b->nwalkers;
    sa.ndecorr = b->ndecorr;
    sa.nblocks = b->nblk;
    sa.bmeans = b->d_bm.p;
    sa.acc = b->d_acc.p;
    for (int c = 0; c < NCLASS; ++c) {
        const int n = b->item_off[c + 1] - b->item_off[c];
        if (n == 0) { continue; }
        sa.items = b->d_items.p + b->item_off[c];
        sa.nitems = n;
        sa.queue = b->d_queue.p + c;
        HIPCHK(mcib::launch_batch_sample(sa, b->item_dmax[c], b->nwaves[c], lane_stream(c)));
    }
    TRY(join(1));
    if (events) { HIPCHK(hipEventRecord(b->ev[2], b->stream)); }
    mcid::EstimJob j;
    j.data = b->d_bm.p;
    j.scratch = nullptr;
    j.n = b->nblk;
    j.W = b->nwalkers;
    j.nobs = 1;
    j.off = 0;
    j.estim = MCI_ESTIM_MJBLOCKER;
    j.scale = 1.0;
    j.active = nullptr;
    j.avg = b->d_wavg.p;
    j.err = b->d_werr.p;
    HIPCHK(mcid::launch_estimate(j, b->stream));
    mcib::CombineArgs cb{b->nint, b->d_woff.p, b->d_R.p, b->d_T.p, b->d_wavg.p, b->d_werr.p, b->d_acc.p, b->mode,
                         b->d_avg.p, b->d_err.p, b->d_accr.p};
    HIPCHK(mcib::launch_batch_combine(cb, b->stream));
    if (events) { HIPCHK(hipEventRecord(b->ev[3], b->stream)); }
    return MCI_OK;
}

float ev_ms(mci_batch *b, int x, int y)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev[x], b->ev[y]) != hipSuccess) { return 0.f; }
    return ms;
}

mci_status run_once(mci_batch *b)
{
    TRY(ensure_uploaded(b));
    if (b->graph_on) {
        if (!b->exec) {
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const mci_status st = enqueue(b, false);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (st != MCI_OK || e != hipSuccess) {
                if (g) { (void)hipGraphDestroy(g); }
                // a failed capture can leave the side streams unusable: replace them
                for (auto &q : b->side) {
                    (void)hipStreamDestroy(q);
                    q = nullptr;
                    HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
                }
                if (st != MCI_OK) { return st; }
            }
            HIPCHK(e);
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        HIPCHK(hipEventRecord(b->ev[0], b->stream));
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
        HIPCHK(hipEventRecord(b->ev[3], b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        b->timings[0] = b->timings[1] = b->timings[2] = 0.;
        b->timings[3] = ev_ms(b, 0, 3);
    }
    else {
        TRY(enqueue(b, true));
        HIPCHK(hipStreamSynchronize(b->stream));
        b->timings[0] = ev_ms(b, 0, 1);
        b->timings[1] = ev_ms(b, 1, 2);
        b->timings[2] = ev_ms(b, 2, 3);
        b->timings[3] = ev_ms(b, 0, 3);
    }
    return MCI_OK;
}

template <class T>
mci_status download(mci_batch *b, T *dst, const DBuf<T> &src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src.p, n * sizeof(T), hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return MCI_OK;
}

mci_status need_run(const mci_batch *b)
{
    if (!b->dev_ready || !b->plan_up) { return fail(MCI_E_RUNTIME, "the batch has not been run"); }
    return MCI_OK;
}
} // namespace

extern "C" {

mci_status mci_batch_create(int nint, const int *ndim, const int64_t *nmc, const int *ncomp, const double *weights,
                            const double *widths, const double *means, mci_batch **out)
{
    if (!out) { return fail(MCI_E_INVALID_ARGUMENT, "null output pointer"); }
    *out = nullptr;
    if (nint < 1) { return fail(MCI_E_INVALID_ARGUMENT, "number of integrals must be at least 1"); }
    if (!ndim || !nmc || !ncomp || !weights || !widths || !means) { return fail(MCI_E_INVALID_ARGUMENT, "null parameter array"); }
    mci_batch *b = new mci_batch();
    b->nint = nint;
    b->d.assign(ndim, ndim + nint);
    b->K.assign(ncomp, ncomp + nint);
    b->nmc.assign(nmc, nmc + nint);
    int64_t nc = 0, nm = 0;
    for (int i = 0; i < nint; ++i) {
        if (b->d[i] < 1 || b->d[i] > MCI_BATCH_MAX_DIM) {
            delete b;
            return fail(MCI_E_INVALID_ARGUMENT, "integral %d: ndim must be in 1..%d", i, MCI_BATCH_MAX_DIM);
        }
        if (b->K[i] < 1 || b->K[i] > MCI_BATCH_MAX_COMP) {
            delete b;
            return fail(MCI_E_INVALID_ARGUMENT, "integral %d: ncomp must be in 1..%d", i, MCI_BATCH_MAX_COMP);
        }
        if (b->nmc[i] < 1) {
            delete b;
            return fail(MCI_E_INVALID_ARGUMENT, "integral %d: nmc must be at least 1", i);
        }
        b->coff.push_back(nc);
        b->moff.push_back(nm);
        nc += b->K[i];
        nm += (int64_t)b->K[i] * b->d[i];
        b->dmax = std::max(b->dmax, b->d[i]);
    }
    b->w.assign(weights, weights + nc);
    b->s.assign(widths, widths + nc);
    b->mu.assign(means, means + nm);
    for (int i = 0; i < nint; ++i) {
        double ws = 0.;
        for (int k = 0; k < b->K[i]; ++k) {
            const double wk = b->w[b->coff[i] + k], sk = b->s[b->coff[i] + k];
            if (!(wk >= 0.) || !std::isfinite(wk) || !(sk > 0.) || !std::isfinite(sk)) {
                delete b;
                return fail(MCI_E_INVALID_ARGUMENT, "integral %d: weights must be finite and >= 0, widths finite and > 0", i);
            }
            ws += wk;
        }
        if (!(ws > 0.)) {
            delete b;
            return fail(MCI_E_INVALID_ARGUMENT, "integral %d: weights sum to zero", i);
        }
    }
    for (double v : b->mu) {
        if (!std::isfinite(v)) {
            delete b;
            return fail(MCI_E_INVALID_ARGUMENT, "means must be finite");
        }
    }
    derive(b);
    *out = b;
    return MCI_OK;
}

mci_status mci_batch_destroy(mci_batch *b)
{
    delete b;
    return MCI_OK;
}

mci_status mci_batch_set_seed(mci_batch *b, uint64_t seed)
{
    CHECKB(b);
    b->seed = seed;
    b->drop_graph();
    return MCI_OK;
}

mci_status mci_batch_set_blocks(mci_batch *b, int nblocks)
{
    CHECKB(b);
    if (nblocks < 4 || nblocks > 64 || (nblocks & (nblocks - 1)) != 0) {
        return fail(MCI_E_INVALID_ARGUMENT, "blocks per walker must be a power of two in 4..64");
    }
    b->nblk = nblocks;
    b->planned = b->user_plan = false;
    b->plan_up = false;
    b->drop_graph();
    return MCI_OK;
}

mci_status mci_batch_set_min_walkers(mci_batch *b, int rmin)
{
    CHECKB(b);
    if (rmin < 1 || rmin > 65536) { return fail(MCI_E_INVALID_ARGUMENT, "minimum walkers must be in 1..65536"); }
    b->rmin = rmin;
    if (!b->user_plan) { b->planned = false; b->plan_up = false; b->drop_graph(); }
    return MCI_OK;
}

mci_status mci_batch_set_error_mode(mci_batch *b, int mode)
{
    CHECKB(b);
    if (mode != MCI_BATCH_ERR_SPREAD && mode != MCI_BATCH_ERR_COMBINE) { return fail(MCI_E_DOMAIN, "unknown error mode %d", mode); }
    b->mode = mode;
    b->drop_graph();
    return MCI_OK;
}

mci_status mci_batch_set_calibration(mci_batch *b, int nfind, double target)
{
    CHECKB(b);
    if (!(target > 0. && target < 1.)) { return fail(MCI_E_INVALID_ARGUMENT, "target acceptance rate must be in (0, 1)"); }
    b->nfind = nfind;
    b->target = target;
    b->drop_graph();
    return MCI_OK;
}

mci_status mci_batch_set_ndecorrelation_steps(mci_batch *b, int64_t n)
{
    CHECKB(b);
    if (n < 0) { return fail(MCI_E_INVALID_ARGUMENT, "decorrelation steps must be >= 0"); }
    b->ndecorr = n;
    b->drop_graph();
    return MCI_OK;
}

mci_status mci_batch_set_plan(mci_batch *b, const int64_t *T)
{
    CHECKB(b);
    if (!T) {
        b->user_plan = false;
        b->planned = false;
        b->plan_up = false;
        b->drop_graph();
        return MCI_OK;
    }
    std::vector<int64_t> t(T, T + b->nint);
    for (int i = 0; i < b->nint; ++i) {
        const int64_t q = t[i] / b->nblk;
        if (t[i] < b->nblk || t[i] % b->nblk != 0 || (q & (q - 1)) != 0 || t[i] > (int64_t(1) << 30)) {
            return fail(MCI_E_INVALID_ARGUMENT, "integral %d: steps per walker must be blocks x a power of two", i);
        }
    }
    int64_t total = 0;
    for (int i = 0; i < b->nint; ++i) { total += (b->nmc[i] + t[i] - 1) / t[i]; }
    if (total >= (int64_t(1) << 31)) { return fail(MCI_E_INVALID_ARGUMENT, "too many walkers (%lld)", (long long)total); }
    apply_plan(b, t);
    b->user_plan = true;
    return MCI_OK;
}

mci_status mci_batch_autotune(mci_batch *b, int measure)
{
    CHECKB(b);
    if (measure > 0) { TRY(ensure_device(b)); }
    const DeviceShape sh = device_shape(b);
    const auto caps = ranked_caps(b, sh);
    apply_plan(b, plan_for_cap(b, caps.front().second));
    if (measure <= 0) { b->user_plan = false; return MCI_OK; }
    const bool graph = b->graph_on;
    b->graph_on = false;
    double best = 1e300;
    int64_t best_cap = caps.front().second;
    for (int c = 0; c < std::min<int>(measure, (int)caps.size()); ++c) {
        apply_plan(b, plan_for_cap(b, caps[c].second));
        mci_status st = run_once(b);
        if (st == MCI_OK) { st = run_once(b); } // the first run of a plan includes first-touch costs
        if (st != MCI_OK) { b->graph_on = graph; return st; }
        const double ms = b->timings[1] + b->timings[2];
        if (ms < best) { best = ms; best_cap = caps[c].second; }
    }
    b->graph_on = graph;
    apply_plan(b, plan_for_cap(b, best_cap));
    b->user_plan = false;
    return MCI_OK;
}

mci_status mci_batch_get_plan(mci_batch *b, int64_t *walkers, int64_t *T, int64_t *woff)
{
    CHECKB(b);
    if (!b->planned) { apply_plan(b, plan_for_cap(b, ranked_caps(b, device_shape(b)).front().second)); }
    for (int i = 0; i < b->nint; ++i) {
        if (walkers) { walkers[i] = b->R[i]; }
        if (T) { T[i] = b->T[i]; }
        if (woff) { woff[i] = b->woff[i]; }
    }
    return MCI_OK;
}

mci_status mci_batch_get_nwalkers(mci_batch *b, int64_t *n)
{
    CHECKB(b);
    if (!b->planned) { apply_plan(b, plan_for_cap(b, ranked_caps(b, device_shape(b)).front().second)); }
    *n = b->nwalkers;
    return MCI_OK;
}

mci_status mci_batch_set_graph(mci_batch *b, int on)
{
    CHECKB(b);
    b->graph_on = on != 0;
    if (!b->graph_on) { b->drop_graph(); }
    return MCI_OK;
}

mci_status mci_batch_run(mci_batch *b, double *avg, double *err)
{
    CHECKB(b);
    if (!avg || !err) { return fail(MCI_E_INVALID_ARGUMENT, "null output arrays"); }
    if (mcib::batch_lds_bytes(b->dmax) > 65536) { return fail(MCI_E_INVALID_ARGUMENT, "dimension too large for one wave's LDS"); }
    TRY(run_once(b));
    TRY(download(b, avg, b->d_avg, b->nint));
    TRY(download(b, err, b->d_err, b->nint));
    return MCI_OK;
}

mci_status mci_batch_get_steps(mci_batch *b, double *steps, double *rate)
{
    CHECKB(b);
    TRY(need_run(b));
    if (steps) { TRY(download(b, steps, b->d_step, b->nint)); }
    if (rate) { TRY(download(b, rate, b->d_rate, b->nint)); }
    return MCI_OK;
}

mci_status mci_batch_get_acceptance(mci_batch *b, double *rate)
{
    CHECKB(b);
    TRY(need_run(b));
    return download(b, rate, b->d_accr, b->nint);
}

mci_status mci_batch_get_walker_results(mci_batch *b, double *avg, double *err, int64_t n)
{
    CHECKB(b);
    TRY(need_run(b));
    if (n != b->nwalkers) { return fail(MCI_E_INVALID_ARGUMENT, "expected %lld walkers", (long long)b->nwalkers); }
    if (avg) { TRY(download(b, avg, b->d_wavg, (size_t)n)); }
    if (err) { TRY(download(b, err, b->d_werr, (size_t)n)); }
    return MCI_OK;
}

mci_status mci_batch_get_block_means(mci_batch *b, double *data, int64_t n)
{
    CHECKB(b);
    TRY(need_run(b));
    if (n != (int64_t)b->nblk * b->nwalkers) { return fail(MCI_E_INVALID_ARGUMENT, "expected nblocks x nwalkers values"); }
    return download(b, data, b->d_bm, (size_t)n);
}

mci_status mci_batch_get_timings(mci_batch *b, double *ms, int n)
{
    CHECKB(b);
    for (int i = 0; i < n && i < 4; ++i) { ms[i] = b->timings[i]; }
    return MCI_OK;
}

} // extern "C"
