// mci_kernels.hip — HIP kernels of the Metropolis path (gfx950).
//
//  k_seed_mt          per-walker std::mt19937_64 seeding (libstdc++ seed sequence)
//  k_sample           MCI::sample: init + nsteps Metropolis steps (+ accumulation)   MCIntegrator.cpp:252-321
//  k_calibrate_indep  MCI::findMRT2Step, one independent loop per walker              MCIntegrator.cpp:87-168
//  k_scale_steps      the step update of findMRT2Step for a shared (rank-averaged) rate
//  k_estimate         estimators per (walker, observable component)                  Estimators.cpp, MJBlocker.cpp
//  k_chunk_sum / k_combine   deterministic cross-walker reduction, MPIMCI combine     MPIMCI.cpp:85-92
//  k_move_x / k_new_random_x  MCI::moveX / MCI::newRandomX                          MCIntegrator.cpp:606-619
#include "mci_kernels.hpp"
#include "mci_step_kernels.hpp"

#include <float.h>

namespace mcid
{
// std::mt19937_64::seed(s) (libstdc++ bits/random.tcc): x0 = s, x_i = f*(x_{i-1} ^ (x_{i-1} >> 62)) + i
__global__ void k_seed_mt(uint64_t *mt, int32_t *mtp, const uint64_t *seeds, int64_t W)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) { return; }
    uint64_t v = seeds[w];
    mt[w] = v;
    for (int i = 1; i < MT_N; ++i) {
        v ^= v >> 62;
        v = v * 6364136223846793005ULL + (uint64_t)i;
        mt[(int64_t)i * W + w] = v;
    }
    mtp[w] = 0;
}

// the bookkeeping of one findMRT2Step iteration for independent walkers (MCIntegrator.cpp:125-166,
// the body of k_calibrate_indep's loop after its MIN_STAT steps): the walker's rate from the
// counters of the sampling launch that just ran, the consecutive count, the factor and the clamps,
// then the loop's exit test; a walker that leaves the loop is marked inactive, and any[0] is set
// while some walker stays in it.  cons / count start at 0 and active at 1.
__global__ void k_calib_indep_round(Chain c, uint8_t *active, int32_t *cons, int32_t *count, int nfind, double target,
                                    const double *dimSizes, int32_t *any)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W || !active[w]) { return; }
    const int64_t acc = c.acc[w], rej = c.rej[w];
    const double rate = (acc > 0) ? (double)acc / ((double)acc + (double)rej) : 0.;
    int cs = cons[w], ct = count[w];
    if (fabs(rate - target) < 0.05) { ++cs; }
    else { cs = 0; }
    const double fact = fmin(2., fmax(0.5, rate / target));
    scale_and_clamp(c, w, fact, dimSizes);
    ++ct;
    const bool stay = !(nfind < 0 && ct >= -nfind) && ((nfind < 0 && cs < 5) || ct < nfind);
    cons[w] = cs;
    count[w] = ct;
    if (stay) { any[0] = 1; }
    else { active[w] = 0; }
}

// Rotates each walker's mt19937_64 state so that its next index is 0: element j of the new state is
// element (j + p) mod 312 of the old one, p the old next index.  The lazy twist regenerates the
// element it outputs from the elements 1 and 156 places after it (mod 312), so a rotated state with
// its index rotated alike produces the same outputs.  Walkers that went through per-walker loops
// (independent calibration and decorrelation) stand at different indices; once realigned, every
// state access of a wave is one row again.  Each lane rotates its own column through LDS.
constexpr int MT_ALIGN_TPB = 64;
__global__ __launch_bounds__(MT_ALIGN_TPB) void k_mt_align(uint64_t *mt, int32_t *mtp, int64_t W)
{
    extern __shared__ uint64_t mt_col[];  // [MT_N][MT_ALIGN_TPB]
    const int64_t w = (int64_t)blockIdx.x * MT_ALIGN_TPB + threadIdx.x;
    if (w >= W) { return; }
    const int p = mtp[w];
    if (p == 0) { return; }
    uint64_t *col = mt_col + threadIdx.x;
#pragma unroll 8
    for (int j = 0; j < MT_N; ++j) {
        const int src = (j + p < MT_N) ? j + p : j + p - MT_N;
        col[j * MT_ALIGN_TPB] = mt[(int64_t)src * W + w];
    }
#pragma unroll 8
    for (int j = 0; j < MT_N; ++j) { mt[(int64_t)j * W + w] = col[j * MT_ALIGN_TPB]; }
    mtp[w] = 0;
}

// The walkers whose flag is set, listed in the first *count entries of list (*count starts at 0; the
// order follows the atomic's, and no result depends on it: each walker's columns move with it).
__global__ void k_compact_active(const uint8_t *active, int64_t W, int32_t *list, int32_t *count)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W || !active[w]) { return; }
    list[atomicAdd(count, 1)] = (int32_t)w;
}

// The per-walker columns a specialised sampling launch reads and writes, moved between a chain's
// arrays and compacted ones (stride W in both): gather (to_dense) moves walker list[i] to column i
// (positions, the step size, the generator state), scatter moves column i back to walker list[i]
// (positions, the generator state, the acceptance counters).
__global__ void k_walker_columns(Chain c, Chain d, const int32_t *list, const int32_t *count, int to_dense)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) { return; }
    const int64_t W = c.W, wl = list[i];
    const int64_t a = to_dense ? wl : i, b = to_dense ? i : wl;  // source and destination columns
    const Chain &src = to_dense ? c : d;
    const Chain &dst = to_dense ? d : c;
    for (int r = 0; r < c.ndim; ++r) { dst.xold[(int64_t)r * W + b] = src.xold[(int64_t)r * W + a]; }
    if (to_dense) { dst.steps[b] = src.steps[a]; }
    else {
        dst.acc[b] = src.acc[a];
        dst.rej[b] = src.rej[a];
    }
    if (c.rng == MCI_RNG_MT19937_64) {
        for (int r = 0; r < MT_N; ++r) { dst.mt[(int64_t)r * W + b] = src.mt[(int64_t)r * W + a]; }
        dst.mtp[b] = src.mtp[a];
    }
    else { dst.ctr[b] = src.ctr[a]; }
}

__global__ void k_scale_steps(Chain c, const CalibCtrl *ctrl, const double *dimSizes)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W || !ctrl->apply) { return; }
    scale_and_clamp(c, w, ctrl->fact, dimSizes);
}

// ordered sum of the chunk sums (the last stage of the deterministic walker sum): the chunk
// sums are staged through LDS by one wave, then added in order by lane 0
__device__ inline double ordered_chunk_sum(const double *chunks, int64_t nch, double *sh)
{
    double s = 0.0;
    bool first = true;
    for (int64_t base = 0; base < nch; base += 1024) {
        const int64_t m = (nch - base < 1024) ? nch - base : 1024;
        for (int64_t i = threadIdx.x; i < m; i += blockDim.x) { sh[i] = chunks[base + i]; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int64_t i = 0; i < m; ++i) {
                if (first) { s = sh[i]; first = false; }
                else { s = s + sh[i]; }
            }
        }
        __syncthreads();
    }
    return s;
}

// one findMRT2Step iteration's bookkeeping (MCIntegrator.cpp:124-167) on the device for walkers
// sharing the rank-averaged rate: rate = (sum of walker rates)/W, consecutive-hit counter, the
// factor clamp(rate/target, 0.5, 2) and the loop condition
__global__ void k_calib_control(const double *chunks, int64_t nch, int64_t W, double target, int nfind, CalibCtrl *ctrl)
{
    __shared__ double sh[1024];
    if (ctrl->done) {
        if (threadIdx.x == 0) { ctrl->apply = 0; }
        return;
    }
    const double sum = ordered_chunk_sum(chunks, nch, sh);
    if (threadIdx.x != 0) { return; }
    const double rate = sum / (double)W;
    ctrl->rate = rate;
    if (fabs(rate - target) < 0.05) { ++ctrl->cons; }
    else { ctrl->cons = 0; }
    ctrl->fact = fmin(2., fmax(0.5, rate / target));
    ctrl->apply = 1;
    ++ctrl->counter;
    const bool stop_cap = (nfind < 0 && ctrl->counter >= -nfind);
    const bool again = (nfind < 0 && ctrl->cons < 5) || ctrl->counter < nfind;
    ctrl->done = (stop_cap || !again) ? 1 : 0;
}

// k_walker_rates + k_chunk_sum + k_calib_control in one launch: each workgroup sums the rates of
// its chunk of 256 walkers (the same tree), and the workgroup that arrives last sums the chunks in
// order and updates the control state.  The arrival counter (ctrl->arrive, zeroed with the
// control state at the start of each calibration) is a device-scope atomic between device-scope
// fences (the chunk sums may sit in another XCD's L2); the last workgroup resets it.
__global__ __launch_bounds__(256) void k_calib_round(const int64_t *acc, const int64_t *rej, int64_t W, double *chunks,
                                                     int64_t nch, double target, int nfind, CalibCtrl *ctrl)
{
    __shared__ double buf[256];
    __shared__ double sh[1024];
    __shared__ int last;
    if (ctrl->done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { ctrl->apply = 0; }
        return;
    }
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double x = 0.0;
    if (i < W) { x = (acc[i] > 0) ? (double)acc[i] / ((double)acc[i] + (double)rej[i]) : 0.; }
    buf[threadIdx.x] = x;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { buf[threadIdx.x] = buf[threadIdx.x] + buf[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        chunks[blockIdx.x] = buf[0];
        __threadfence();
        last = (atomicAdd(&ctrl->arrive, 1u) == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (!last) { return; }
    __threadfence();
    const double sum = ordered_chunk_sum(chunks, nch, sh);
    if (threadIdx.x != 0) { return; }
    ctrl->arrive = 0u;
    const double rate = sum / (double)W;
    ctrl->rate = rate;
    if (fabs(rate - target) < 0.05) { ++ctrl->cons; }
    else { ctrl->cons = 0; }
    ctrl->fact = fmin(2., fmax(0.5, rate / target));
    ctrl->apply = 1;
    ++ctrl->counter;
    const bool stop_cap = (nfind < 0 && ctrl->counter >= -nfind);
    const bool again = (nfind < 0 && ctrl->cons < 5) || ctrl->counter < nfind;
    ctrl->done = (stop_cap || !again) ? 1 : 0;
}

__global__ void k_walker_rates(const int64_t *acc, const int64_t *rej, double *rate, int64_t W)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) { return; }
    rate[w] = (acc[w] > 0) ? (double)acc[w] / ((double)acc[w] + (double)rej[w]) : 0.;
}

// ---- estimators (one thread per walker and observable component) ----
struct Col
{
    const double *p;
    int64_t stride;
    __device__ double operator[](int64_t i) const { return p[i * stride]; }
};

// OneDim/MultiDimUncorrelatedEstimator (Estimators.cpp:37-57 / 125-155): identical sums,
// the 1-D form divides by (n-1.), the N-D form multiplies by 1./(n-1.)
template <class X>
__device__ inline void uncorr(int64_t n, const X &x, bool onedim, double &avg, double &err)
{
    double a = 0.0, e = 0.0;
    for (int64_t i = 0; i < n; ++i) { const double v = x[i]; a = a + v; e = e + v * v; }
    const double norm = 1. / (double)n;
    a = a * norm;
    e = e * norm - a * a;
    if (e > 1.e-300) { e = onedim ? sqrt(e / ((double)n - 1.)) : sqrt(e * (1. / ((double)n - 1.))); }
    else { e = 0.; }
    avg = a;
    err = e;
}

struct Arr
{
    const double *p;
    __device__ double operator[](int64_t i) const { return p[i]; }
};

__device__ inline double err_delta(int mode, const double *e)
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

// FCBlocker (Estimators.cpp:83-122, 191-246): 45 block estimates (OneDim/MultiDimBlockEstimator,
// Estimators.cpp:60-77 / 158-185, with k = 6..50 blocks of p = n / k values), the plateau by the
// smallest summed central difference, a 5-point average around it.
//
// The estimate with k blocks sums each block from 0 in index order and then sums the block means
// (and their squares) from 0 in block order.  Both are prefix computations: every k with the same
// block length p sees the same block means, and its sums over the first k of them are the running
// sums after k blocks.  So each distinct p is swept once, and every k of its range (kmin..kmax) is
// finished as the sweep passes it -- the same operations in the same order as 45 separate
// estimates, with about n values read per distinct p instead of per k (13 sweeps instead of 45 at
// n = 100).  The estimates arrive in increasing k, so the plateau search runs as they arrive on a
// window of the last nine: each central-difference sum and its 5-point averages are formed in the
// reference's order of operations, and no array is indexed at run time.
constexpr int FC_NB = 45;   // block counts 6..50
__constant__ double c_fc_rk[FC_NB] = {1. / 6., 1. / 7., 1. / 8., 1. / 9., 1. / 10., 1. / 11., 1. / 12., 1. / 13., 1. / 14., 1. / 15., 1. / 16., 1. / 17., 1. / 18., 1. / 19., 1. / 20., 1. / 21., 1. / 22., 1. / 23., 1. / 24., 1. / 25., 1. / 26., 1. / 27., 1. / 28., 1. / 29., 1. / 30., 1. / 31., 1. / 32., 1. / 33., 1. / 34., 1. / 35., 1. / 36., 1. / 37., 1. / 38., 1. / 39., 1. / 40., 1. / 41., 1. / 42., 1. / 43., 1. / 44., 1. / 45., 1. / 46., 1. / 47., 1. / 48., 1. / 49., 1. / 50.};
__constant__ double c_fc_rk1[FC_NB] = {1. / 5., 1. / 6., 1. / 7., 1. / 8., 1. / 9., 1. / 10., 1. / 11., 1. / 12., 1. / 13., 1. / 14., 1. / 15., 1. / 16., 1. / 17., 1. / 18., 1. / 19., 1. / 20., 1. / 21., 1. / 22., 1. / 23., 1. / 24., 1. / 25., 1. / 26., 1. / 27., 1. / 28., 1. / 29., 1. / 30., 1. / 31., 1. / 32., 1. / 33., 1. / 34., 1. / 35., 1. / 36., 1. / 37., 1. / 38., 1. / 39., 1. / 40., 1. / 41., 1. / 42., 1. / 43., 1. / 44., 1. / 45., 1. / 46., 1. / 47., 1. / 48., 1. / 49.};

struct FcState
{
    double wa[9], we[9];    // the last nine estimates (averages, errors)
    double best, bav, ber;  // the smallest |central-difference sum| so far and its 5-point averages
};

// finishes the estimate with k blocks (running sums a, e of its k block means: the tail of
// uncorr()) and passes it to the plateau search; k is uniform
__device__ inline void fc_push(FcState &st, int64_t k, double aa, double ee);

__device__ inline void fc_emit(FcState &st, int64_t k, double a, double e, bool onedim)
{
    const double norm = c_fc_rk[k - 6];
    const double aa = a * norm;
    double ee = e * norm - aa * aa;
    if (ee > 1.e-300) { ee = onedim ? sqrt(ee / ((double)k - 1.)) : sqrt(ee * c_fc_rk1[k - 6]); }
    else { ee = 0.; }
    fc_push(st, k, aa, ee);
}

// the estimate with k blocks (average aa, error ee) into the window and the plateau search
__device__ inline void fc_push(FcState &st, int64_t k, double aa, double ee)
{
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        st.wa[t] = st.wa[t + 1];
        st.we[t] = st.we[t + 1];
    }
    st.wa[8] = aa;
    st.we[8] = ee;
    if (k >= 14) {  // the window holds the estimates k-8..k: the central differences around k-4
        double acd = 0.;
#pragma unroll
        for (int i1 = 1; i1 <= 4; ++i1) { acd = acd + err_delta(i1, st.we); }
        const double fa = fabs(acd);
        const double sa = 0.2 * (st.wa[2] + st.wa[3] + st.wa[4] + st.wa[5] + st.wa[6]);
        const double se = 0.2 * (st.we[2] + st.we[3] + st.we[4] + st.we[5] + st.we[6]);
        if (k == 14 || fa < st.best) {
            st.best = fa;
            st.bav = sa;
            st.ber = se;
        }
    }
}

// The sweep of blocks 0..kmax-1 of length p (values 0..kmax*p-1), finishing k = kmin..kmax.  The
// values are read in batches of eight, one batch ahead of the sums (X::load8 reads any batch start
// below n + 16 in bounds); a uniform countdown closes the blocks.  Values past the sweep's end
// (the rest of its last batches) go into blocks that are never finished.
template <class X>
__device__ inline void fc_sweep(const X &x, int64_t p, int64_t kmin, int64_t kmax, bool onedim, FcState &st)
{
    const double nrm = 1. / (double)p;
    const int64_t iend = kmax * p;
    double a = 0., e = 0., s = 0.0;
    using I = typename X::index;  // block counters: 32 bits where n fits
    I left = (I)p, b = 0;
    double v0[8], v1[8];
    x.load8(0, v0);
    x.load8(8, v1);
    auto batch = [&](double (&v)[8], int64_t inext) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            s = s + v[t];
            if (--left == 0) {
                const double m = s * nrm;
                a = a + m;
                e = e + m * m;
                s = 0.0;
                left = (I)p;
                ++b;
                if (b >= kmin && b <= kmax) { fc_emit(st, b, a, e, onedim); }
            }
        }
        x.load8(inext, v);
    };
    for (int64_t i = 0; i < iend; i += 16) {
        batch(v0, i + 16);
        if (i + 8 < iend) { batch(v1, i + 24); }
    }
}

template <class X>
__device__ inline void fcblocker(int64_t n, const X &x, bool onedim, double &avg, double &err)
{
    FcState st;
#pragma unroll
    for (int t = 0; t < 9; ++t) { st.wa[t] = 0.; st.we[t] = 0.; }
    st.best = 0.;
    st.bav = 0.;
    st.ber = 0.;
    for (int64_t kmin = 6; kmin <= 50;) {
        const int64_t p = n / kmin;
        const int64_t kmax = (n / p < 50) ? n / p : 50;  // the largest k with n / k == p
        fc_sweep(x, p, kmin, kmax, onedim, st);
        kmin = kmax + 1;
    }
    avg = st.bav;
    err = st.ber;
}

// The same for a compile-time n (FcReg<N, 0, 0, ONEDIM>::run): the series in registers and every sweep
// unrolled, so that the 45 estimates are straight-line code -- the sweeps of different block
// lengths interleave freely, the window's shifts are register renames and no scalar bookkeeping is
// left.  n = 100 is the series length of the automatic decorrelation with many walkers.
template <int N, int Q, int BDONE, bool ONEDIM>
struct FcReg
{
    static constexpr int K = Q + 6;
    static constexpr int P = N / K;
    static constexpr bool NEWP = (Q == 0) || (N / (K - 1) != P);  // the first k of its block length
    static constexpr int B0 = NEWP ? 0 : BDONE;
    __device__ static void run(const double (&y)[N], double a, double e, FcState &st)
    {
        if constexpr (NEWP) {
            a = 0.;
            e = 0.;
        }
        const double nrm = 1. / (double)P;
#pragma unroll
        for (int b = B0; b < K; ++b) {
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < P; ++t) { s = s + y[b * P + t]; }
            const double m = s * nrm;
            a = a + m;
            e = e + m * m;
        }
        fc_emit(st, K, a, e, ONEDIM);
        if constexpr (Q + 1 < FC_NB) { FcReg<N, Q + 1, K, ONEDIM>::run(y, a, e, st); }
    }
};

template <int N, bool ONEDIM>
__device__ inline void fcblocker_reg(const Col &x, double &avg, double &err)
{
    static_assert(N >= 50, "FCBlocker needs at least 50 values");
    double y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) { y[i] = x[i]; }
    FcState st;
#pragma unroll
    for (int t = 0; t < 9; ++t) { st.wa[t] = 0.; st.we[t] = 0.; }
    st.best = 0.;
    st.bav = 0.;
    st.ber = 0.;
    FcReg<N, 0, 0, ONEDIM>::run(y, 0., 0., st);
    avg = st.bav;
    err = st.ber;
}

// The same for longer series, streaming each value past running block sums.  Every block count k
// keeps its own running block sum and every value is added to each of them (independent chains);
// the block ends come from a table built per workgroup in LDS: for each position, the mask of the
// block counts that close a block there and the mask of those that close their last block there.
// Each block count tests its bit (groups of eight without a close are skipped together), so the
// sums stay in registers and no per-value bookkeeping runs on the scalar unit.  Per block count the operations and their order
// are those of OneDim/MultiDimBlockEstimator: each block summed from 0 in index order, its mean
// added to the running sums of means and of their squares.  After a block count's last block its
// running sum takes the rest of the series, which no close reads.  The 45 block counts take three
// passes over the series (15 each: their three running sums each fit the registers); each pass
// hands its finished estimates to the windowed plateau search in increasing k.  The values reach LDS in batches of eight through a two-batch ring per lane, the
// loads two batches ahead; each value and its masks are read one position ahead.
constexpr int64_t FC_STREAM_MAX = 2048;
constexpr int FC_STREAM_TPB = 256;
constexpr int FC_SB = 16;
__host__ __device__ inline size_t fc_stream_lds(int64_t n)
{
    return 16 * (size_t)n + 48 * 8 + (FC_STREAM_TPB / 64) * (2 * FC_SB) * 64 * 8;
}

template <bool ONEDIM>
__device__ inline void fc_finish(int k, double &a, double &e)
{
    const double norm = c_fc_rk[k - 6];
    const double aa = a * norm;
    double ee = e * norm - aa * aa;
    if (ee > 1.e-300) { ee = ONEDIM ? sqrt(ee / ((double)k - 1.)) : sqrt(ee * c_fc_rk1[k - 6]); }
    else { ee = 0.; }
    a = aa;
    e = ee;
}

__device__ inline uint64_t uniform64(uint64_t v)
{
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) | (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// rows i0 .. i0+FC_SB-1 of a series (indices past the end clamped to its last value)
__device__ inline void fc_load_batch(const double *xp, int64_t stride, int n, int i0, double (&v)[FC_SB])
{
    if (i0 + FC_SB <= n) {
        const double *q = xp + (int64_t)i0 * stride;
#pragma unroll
        for (int t = 0; t < FC_SB; ++t) {
            v[t] = *q;
            q += stride;
        }
    }
    else {
#pragma unroll
        for (int t = 0; t < FC_SB; ++t) { v[t] = xp[(int64_t)((i0 + t < n) ? i0 + t : n - 1) * stride]; }
    }
}

// block counts k = Q0+6 .. Q1+5 over the whole series
template <bool ONEDIM, int Q0, int Q1>
__device__ inline void fc_stream_pass(const double *xp, int64_t stride, int n, const uint64_t *cmask, const uint64_t *fmask,
                                      const double *nrm, double *ring, FcState &st)
{
    constexpr int NQ = Q1 - Q0;
    constexpr uint64_t RANGE = ((1ull << NQ) - 1) << Q0;
    double s[NQ], a[NQ], e[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        s[q] = 0.0;
        a[q] = 0.;
        e[q] = 0.;
    }
    double v[FC_SB];
    fc_load_batch(xp, stride, n, 0, v);
#pragma unroll
    for (int t = 0; t < FC_SB; ++t) { ring[t * 64] = v[t]; }
    fc_load_batch(xp, stride, n, FC_SB, v);
    double x = ring[0];
    uint64_t cm = cmask[0], fm = fmask[0];
    for (int i0 = 0; i0 < n; i0 += FC_SB) {
        // batch j + 1 into the ring's other half, batch j + 2's loads in flight
        const int half = ((i0 / FC_SB + 1) & 1) * FC_SB;
#pragma unroll
        for (int t = 0; t < FC_SB; ++t) { ring[(half + t) * 64] = v[t]; }
        fc_load_batch(xp, stride, n, i0 + 2 * FC_SB, v);
        const int iend = (i0 + FC_SB < n) ? i0 + FC_SB : n;
#pragma unroll 1
        for (int i = i0; i < iend; ++i) {
            const double xc = x;
            const uint64_t cc = cm, ff = fm;
            if (i + 1 < n) {
                x = ring[((i + 1) & (2 * FC_SB - 1)) * 64];
                cm = cmask[i + 1];
                fm = fmask[i + 1];
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) { s[q] = s[q] + xc; }
            const uint64_t c = uniform64(cc) & RANGE;
            if (c == 0) { continue; }
            const uint64_t f = uniform64(ff);
            // each block count that closes here updates its own registers (no dispatch, no merge);
            // groups of eight block counts with no close skipped at once
#pragma unroll
            for (int g = Q0; g < Q1; g += 8) {
                if ((c >> g) & 0xFFull) {
#pragma unroll
                    for (int q = g; q < g + 8 && q < Q1; ++q) {
                        if ((c >> q) & 1u) {
                            const double m = s[q - Q0] * nrm[q];
                            a[q - Q0] = a[q - Q0] + m;
                            e[q - Q0] = e[q - Q0] + m * m;
                            s[q - Q0] = 0.0;
                            if ((f >> q) & 1u) { fc_finish<ONEDIM>(q + 6, a[q - Q0], e[q - Q0]); }
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) { fc_push(st, Q0 + q + 6, a[q], e[q]); }
}

template <bool ONEDIM>
__global__ __launch_bounds__(FC_STREAM_TPB) void k_fc_stream(EstimJob j)
{
    extern __shared__ double fs_sh[];
    const int n = (int)j.n;
    uint64_t *cmask = reinterpret_cast<uint64_t *>(fs_sh);
    uint64_t *fmask = cmask + n;
    double *nrm = reinterpret_cast<double *>(fmask + n);
    double *ring = nrm + 48 + (threadIdx.x >> 6) * (2 * FC_SB * 64) + (threadIdx.x & 63);  // [16][64] per wave
    for (int i = threadIdx.x; i < n; i += FC_STREAM_TPB) {
        cmask[i] = 0;
        fmask[i] = 0;
    }
    if (threadIdx.x < FC_NB) { nrm[threadIdx.x] = 1. / (double)(n / ((int)threadIdx.x + 6)); }
    __syncthreads();
    for (int r = threadIdx.x; r < FC_NB * 50; r += FC_STREAM_TPB) {
        const int q = r / 50, b = r % 50, k = q + 6;
        if (b < k) {
            const int pos = (b + 1) * (n / k) - 1;
            atomicOr((unsigned long long *)&cmask[pos], 1ull << q);
            if (b == k - 1) { atomicOr((unsigned long long *)&fmask[pos], 1ull << q); }
        }
    }
    __syncthreads();
    const int64_t tid = (int64_t)blockIdx.x * FC_STREAM_TPB + threadIdx.x;
    if (tid >= j.W * j.nobs) { return; }
    const int64_t w = tid % j.W;
    const int comp = (int)(tid / j.W);
    if (j.wcount && w >= *j.wcount) { return; }
    const int64_t wo = j.wmap ? j.wmap[w] : w;  // the walker of this column
    if (j.active && !j.active[wo]) { return; }
    const double *xp = j.data + (int64_t)comp * j.W + w;
    const int64_t stride = (int64_t)j.nobs * j.W;
    FcState st;
#pragma unroll
    for (int t = 0; t < 9; ++t) { st.wa[t] = 0.; st.we[t] = 0.; }
    st.best = 0.;
    st.bav = 0.;
    st.ber = 0.;
    fc_stream_pass<ONEDIM, 0, 15>(xp, stride, n, cmask, fmask, nrm, ring, st);
    fc_stream_pass<ONEDIM, 15, 30>(xp, stride, n, cmask, fmask, nrm, ring, st);
    fc_stream_pass<ONEDIM, 30, FC_NB>(xp, stride, n, cmask, fmask, nrm, ring, st);
    j.avg[(int64_t)(j.off + comp) * j.W + wo] = st.bav * j.scale;
    j.err[(int64_t)(j.off + comp) * j.W + wo] = st.ber * j.scale;
}

// a thread's series staged in LDS, one column per lane of a 64-thread workgroup, rows allocated to a
// multiple of 8 (>= n): a batch that starts at or past the allocation reads rows 0..7 instead
struct LdsCol
{
    using index = int;
    const double *p;
    int64_t rows;
    __device__ void load8(int64_t i0, double (&v)[8]) const
    {
        const double *q = p + ((i0 < rows) ? i0 : 0) * 64;
#pragma unroll
        for (int t = 0; t < 8; ++t) { v[t] = q[t * 64]; }
    }
};

// a series read from memory: indices clamped to the last value
struct MemCol
{
    using index = int64_t;
    const double *p;
    int64_t stride, n;
    __device__ void load8(int64_t i0, double (&v)[8]) const
    {
#pragma unroll
        for (int t = 0; t < 8; ++t) { v[t] = p[((i0 + t < n) ? i0 + t : n - 1) * stride]; }
    }
};

__constant__ double c_chi2_q95[64] = {
    3.841459, 5.991465, 7.814728, 9.487729, 11.070498, 12.591587, 14.067140, 15.507313,
    16.918978, 18.307038, 19.675138, 21.026070, 22.362032, 23.684791, 24.995790, 26.296228,
    27.587112, 28.869299, 30.143527, 31.410433, 32.670573, 33.924438, 35.172462, 36.415029,
    37.652484, 38.885139, 40.113272, 41.337138, 42.556968, 43.772972, 44.985343, 46.194260,
    47.399884, 48.602367, 49.801850, 50.998460, 52.192320, 53.383541, 54.572228, 55.758479,
    56.942387, 58.124038, 59.303512, 60.480887, 61.656233, 62.829620, 64.001112, 65.170769,
    66.338649, 67.504807, 68.669294, 69.832160, 70.993453, 72.153216, 73.311493, 74.468324,
    75.623748, 76.777803, 77.930524, 79.081944, 80.232098, 81.381015, 82.528727, 83.675261};

// The first MJ_LB blocking levels of a block of 2^MJ_LB values, in registers: each level's terms
// added to its sums in index order (FIRST: the series' first block, whose first value has no
// predecessor), then the block halved in place; v[0] ends as the block's value at level MJ_LB.
constexpr int MJ_LB = 4, MJ_BS = 1 << MJ_LB;
template <bool FIRST>
__device__ inline void mj_block_levels(double (&v)[MJ_BS], double mean, double (&g0)[MJ_LB], double (&g1)[MJ_LB], double (&pv)[MJ_LB])
{
#pragma unroll
    for (int k = 0; k < MJ_LB; ++k) {
        const int len = MJ_BS >> k;
#pragma unroll
        for (int t = 0; t < MJ_BS; ++t) {
            if (t < len) {
                const double c = v[t] - mean;
                g0[k] = g0[k] + c * c;
                if (!(FIRST && t == 0)) { g1[k] = g1[k] + pv[k] * c; }
                pv[k] = c;
            }
        }
#pragma unroll
        for (int t = 0; t < MJ_BS / 2; ++t) {
            if (t < len / 2) { v[t] = 0.5 * (v[2 * t] + v[2 * t + 1]); }
        }
    }
}

__device__ inline void mj_load_block(const Col &x, int64_t i0, double (&v)[MJ_BS])
{
#pragma unroll
    for (int t = 0; t < MJ_BS; ++t) { v[t] = x[i0 + t]; }
}

// Automatic blocking of Jonsson (PRE 98, 043304 (2018)) as organised in MJBlocker.cpp:127-154.
// The blocking transform runs in place on this thread's column of the scratch copy `y`.
// A level whose variance is exactly zero contributes M = 0 (the reference reads out of
// bounds there).
__device__ inline void mjblocker(int64_t n, const Col &x, double *y, int64_t ystride, double &avg, double &err)
{
    int npow = 0;
    while (((int64_t)1 << npow) < n) { ++npow; }
    double mean = 0.0;
    int64_t i = 0;
    for (; i + MJ_BS <= n; i += MJ_BS) {  // batches of loads, the sum in index order
        double v[MJ_BS];
        mj_load_block(x, i, v);
#pragma unroll
        for (int t = 0; t < MJ_BS; ++t) { mean = mean + v[t]; }
    }
    for (; i < n; ++i) { mean = mean + x[i]; }
    mean = mean / (double)n;
    double var[64], gam[64];
    int64_t nred = n;
    int k0 = 0;
    if (n >= 2 * MJ_BS && (n & (n - 1)) == 0) {
        // Levels 0 .. MJ_LB-1 in one pass over the data: blocks of MJ_BS values (every pair of
        // every one of these levels lies inside one block) in registers; each level's sums run
        // over its values in index order as in the per-level passes below, and only the
        // level-MJ_LB values (n / MJ_BS) go to the scratch copy.  About 2n values of traffic
        // instead of 4n; few registers, so that many waves keep loads in flight.
        double g0[MJ_LB], g1[MJ_LB], pv[MJ_LB];
#pragma unroll
        for (int k = 0; k < MJ_LB; ++k) {
            g0[k] = 0.0;
            g1[k] = 0.0;
            pv[k] = 0.0;
        }
        const int64_t nb = n / MJ_BS;
        {
            double v[MJ_BS];
            mj_load_block(x, 0, v);
            mj_block_levels<true>(v, mean, g0, g1, pv);
            y[0] = v[0];
        }
        for (int64_t b = 1; b < nb; ++b) {
            double v[MJ_BS];
            mj_load_block(x, b * MJ_BS, v);
            mj_block_levels<false>(v, mean, g0, g1, pv);
            y[b * ystride] = v[0];
        }
#pragma unroll
        for (int k = 0; k < MJ_LB; ++k) {
            var[k] = g0[k] / (double)(n >> k);
            gam[k] = g1[k] / (double)(n >> k);
        }
        nred = n >> MJ_LB;
        k0 = MJ_LB;
    }
    // One pass per level: the sums of the level and the pair means of the next one, written in
    // place (index i/2 is written after i has been read).  Level 0 reads the data itself, so the
    // copy holds at most n/2 values and every level is read once: about 4n values of traffic
    // where a copy followed by separate sum and halving passes moves 7n.
    auto level = [&](auto src, int k) {
        double g0 = 0.0, g1 = 0.0;
        double even = src(0);
        double prev = even - mean;
        g0 = g0 + prev * prev;
        for (int64_t i = 1; i < nred; ++i) {
            const double raw = src(i);
            const double cur = raw - mean;
            g0 = g0 + cur * cur;
            g1 = g1 + prev * cur;
            prev = cur;
            if (i & 1) { y[(i >> 1) * ystride] = 0.5 * (even + raw); }
            else { even = raw; }
        }
        var[k] = g0 / (double)nred;
        gam[k] = g1 / (double)nred;
        nred = nred / 2;
    };
    for (int k = k0; k < npow; ++k) {
        if (k == 0) { level([&](int64_t i) { return x[i]; }, k); }
        else { level([&](int64_t i) { return y[i * ystride]; }, k); }
    }
    // M (reversed) and its cumulative sum; largest k with Msum_k < chi^2_{0.95}(k+1)
    double M[64];
    for (int i = 0; i < npow; ++i) {
        const double v = var[i], g = gam[i];
        const double r = g / v;
        M[npow - i - 1] = (v == 0.0) ? 0.0 : (r * r) * ldexp(1.0, npow - i);
    }
    int kk = -1;
    double msum = 0.0;
    double ms[64];
    for (int i = 0; i < npow; ++i) { msum = msum + M[i]; ms[i] = msum; }
    for (kk = npow - 1; kk >= 0; --kk) {
        if (ms[kk] < c_chi2_q95[kk]) { break; }
    }
    int k = npow - (kk + 1);
    if (npow == 0) { avg = mean; err = 0.0; return; }
    if (k >= npow) { k = npow - 1; }
    avg = mean;
    err = sqrt(var[k] / ldexp(1.0, npow - k));
}

// the same for n <= 64 (at most 6 blocking levels): identical operations in identical order,
// with the per-level arrays in registers (fixed-size loops, runtime indices as selects) instead
// of private memory
template <int L>
__device__ inline double pick(const double (&a)[L], int i)
{
    double v = a[0];
#pragma unroll
    for (int q = 1; q < L; ++q) { v = (i == q) ? a[q] : v; }
    return v;
}

__device__ inline void mjblocker_small(int64_t n, const Col &x, double *y, int64_t ystride, double &avg, double &err)
{
    constexpr int LMAX = 6; // log2(MJ_LDS_MAX)
    int npow = 0;
    while (((int64_t)1 << npow) < n) { ++npow; }
    double mean = 0.0;
    // sixteen loads in flight, then the copies and the sum in order
    for (int64_t i0 = 0; i0 < n; i0 += 16) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) { v[q] = (i0 + q < n) ? x[i0 + q] : 0.0; }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            if (i0 + q < n) {
                y[(i0 + q) * ystride] = v[q];
                mean = mean + v[q];
            }
        }
    }
    mean = mean / (double)n;
    double var[LMAX], gam[LMAX];
    int64_t nred = n;
#pragma unroll
    for (int k = 0; k < LMAX; ++k) {
        var[k] = 0.0;
        gam[k] = 0.0;
        if (k < npow) {
            double g0 = 0.0, g1 = 0.0;
            double prev = y[0] - mean;
            g0 = g0 + prev * prev;
            for (int64_t i = 1; i < nred; ++i) {
                const double cur = y[i * ystride] - mean;
                g0 = g0 + cur * cur;
                g1 = g1 + prev * cur;
                prev = cur;
            }
            var[k] = g0 / (double)nred;
            gam[k] = g1 / (double)nred;
            for (int64_t i = 0; i < nred / 2; ++i) {
                y[i * ystride] = 0.5 * (y[(2 * i) * ystride] + y[(2 * i + 1) * ystride]);
            }
            nred = nred / 2;
        }
    }
    if (npow == 0) { avg = mean; err = 0.0; return; }
    // M of level i goes to reversed position npow - 1 - i; cumulative sums in that order
    double Ml[LMAX];
#pragma unroll
    for (int i = 0; i < LMAX; ++i) {
        const double v = var[i], g = gam[i];
        const double r = g / v;
        Ml[i] = (v == 0.0) ? 0.0 : (r * r) * ldexp(1.0, npow - i);
    }
    double ms[LMAX];
    double msum = 0.0;
#pragma unroll
    for (int q = 0; q < LMAX; ++q) {
        if (q < npow) { msum = msum + pick(Ml, npow - 1 - q); }
        ms[q] = msum;
    }
    int kk = -1;
#pragma unroll
    for (int q = LMAX - 1; q >= 0; --q) {
        if (kk < 0 && q < npow && ms[q] < c_chi2_q95[q]) { kk = q; }
    }
    int k = npow - (kk + 1);
    if (k >= npow) { k = npow - 1; }
    avg = mean;
    err = sqrt(pick(var, k) / ldexp(1.0, npow - k));
}

// the same for n = 2^P <= 64: the series in registers, every loop unrolled at compile time, so
// that all n loads are in flight at once and no LDS round trip sits in the sums' dependency
// chains (identical operations in identical order)
template <int P>
__device__ inline void mjblocker_reg(const Col &x, double &avg, double &err)
{
    constexpr int N = 1 << P;
    double y[N];
#pragma unroll
    for (int i = 0; i < N; ++i) { y[i] = x[i]; }
    double mean = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) { mean = mean + y[i]; }
    mean = mean / (double)N;
    if (P == 0) { avg = mean; err = 0.0; return; }
    double var[P > 0 ? P : 1], gam[P > 0 ? P : 1];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        const int nred = N >> k;
        double g0 = 0.0, g1 = 0.0;
        double prev = y[0] - mean;
        g0 = g0 + prev * prev;
#pragma unroll
        for (int i = 1; i < nred; ++i) {
            const double cur = y[i] - mean;
            g0 = g0 + cur * cur;
            g1 = g1 + prev * cur;
            prev = cur;
        }
        var[k] = g0 / (double)nred;
        gam[k] = g1 / (double)nred;
#pragma unroll
        for (int i = 0; i < nred / 2; ++i) { y[i] = 0.5 * (y[2 * i] + y[2 * i + 1]); }
    }
    // M of level i at reversed position P - 1 - i; cumulative sums in that order
    double ms[P > 0 ? P : 1];
    double msum = 0.0;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int i = P - 1 - q;
        const double v = var[i], g = gam[i];
        const double r = g / v;
        msum = msum + ((v == 0.0) ? 0.0 : (r * r) * ldexp(1.0, P - i));
        ms[q] = msum;
    }
    int kk = -1;
#pragma unroll
    for (int q = P - 1; q >= 0; --q) {
        if (kk < 0 && ms[q] < c_chi2_q95[q]) { kk = q; }
    }
    int k = P - (kk + 1);
    if (k >= P) { k = P - 1; }
    avg = mean;
    err = sqrt(pick(var, k) / ldexp(1.0, P - k));
}

// one instantiation per estimator (E: MCI_ESTIM_NOOP, _UNCORRELATED, _FCBLOCKER, _MJBLOCKER, or
// the staged forms below; the host resolves CORRELATED by n), so that the private arrays of the
// general MJBlocker (scratch memory) and FCBlocker's 90 registers of estimates weigh on no other
constexpr int EST_MJ_SMALL = 5;
constexpr int EST_MJ_REG = 6;  // MJBlocker on 2^P <= 64 values in registers
constexpr int EST_FC_LDS = 7;  // FCBlocker on at most FC_LDS_MAX values staged in LDS
constexpr int EST_FC_REG = 8;  // FCBlocker on FC_REG_N values in registers
constexpr int64_t FC_REG_N = 100;
constexpr int64_t FC_LDS_MAX = 128;
__host__ __device__ inline int64_t fc_rows(int64_t n) { return (n + 7) / 8 * 8; }
template <int E, int P = 0>
__global__ __launch_bounds__(256) void k_estimate(EstimJob j)
{
    extern __shared__ double est_sh[];
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= j.W * j.nobs) { return; }
    const int64_t w = tid % j.W;
    const int comp = (int)(tid / j.W);
    if (j.wcount && w >= *j.wcount) { return; }
    const int64_t wo = j.wmap ? j.wmap[w] : w;  // the walker of this column
    if (j.active && !j.active[wo]) { return; }
    const Col x{j.data + (int64_t)comp * j.W + w, (int64_t)j.nobs * j.W};
    const bool onedim = (j.nobs == 1);
    double a = 0., e = 0.;
    if (E == MCI_ESTIM_NOOP) { a = x[0]; e = 0.; }
    if (E == MCI_ESTIM_UNCORRELATED) { uncorr(j.n, x, onedim, a, e); }
    if (E == MCI_ESTIM_FCBLOCKER) { fcblocker(j.n, MemCol{x.p, x.stride, j.n}, onedim, a, e); }
    if (E == EST_FC_REG) { fcblocker_reg<FC_REG_N, P == 1>(x, a, e); }  // P: one component
    if (E == EST_FC_LDS) {
        // each lane writes and reads only its own column: no barrier.  64 loads in flight per lane
        // (the indices past the series clamped to its last value, the rows past the allocation not
        // written), so that a series of up to 128 values is staged in two memory round trips
        double *col = est_sh + threadIdx.x;
        const int64_t rows = fc_rows(j.n);
        for (int64_t c0 = 0; c0 < rows; c0 += 64) {
            double v[64];
#pragma unroll
            for (int t = 0; t < 64; ++t) { v[t] = x[(c0 + t < j.n) ? c0 + t : j.n - 1]; }
#pragma unroll
            for (int t = 0; t < 64; ++t) {
                if (c0 + t < rows) { col[(c0 + t) * 64] = v[t]; }
            }
        }
        fcblocker(j.n, LdsCol{col, rows}, onedim, a, e);
    }
    if (E == MCI_ESTIM_MJBLOCKER) { mjblocker(j.n, x, j.scratch + (int64_t)comp * j.W + w, (int64_t)j.nobs * j.W, a, e); }
    if (E == EST_MJ_SMALL) { mjblocker_small(j.n, x, est_sh + threadIdx.x, (int64_t)blockDim.x, a, e); }
    if (E == EST_MJ_REG) { mjblocker_reg<P>(x, a, e); }
    j.avg[(int64_t)(j.off + comp) * j.W + wo] = a * j.scale;
    j.err[(int64_t)(j.off + comp) * j.W + wo] = e * j.scale;
}

// ---- deterministic cross-walker sums: chunks of 256 (pairwise tree), chunks in order ----
__global__ void k_chunk_sum(const double *v, int64_t W, int ncomp, int square, double *chunks, int64_t nchunks)
{
    __shared__ double buf[256];
    const int64_t chunk = blockIdx.x;
    const int comp = blockIdx.y;
    const int64_t i = chunk * 256 + threadIdx.x;
    double x = (i < W) ? v[(int64_t)comp * W + i] : 0.0;
    if (square) { x = x * x; }
    buf[threadIdx.x] = x;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { buf[threadIdx.x] = buf[threadIdx.x] + buf[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { chunks[(int64_t)comp * nchunks + chunk] = buf[0]; }
}

// MPIMCI::integrate combine (MPIMCI.cpp:85-92): avg = sum/W, err = sqrt(sum err^2)/W.  One
// workgroup per component: the chunk sums are staged through LDS (coalesced), then added in
// order by lane 0 (one dependent add per chunk, no memory round trip inside the chain)
__global__ void k_combine(const double *csum_avg, const double *csum_err2, int64_t nchunks, int ncomp, int64_t W,
                          double *avg, double *err)
{
    __shared__ double sa[1024], se[1024];
    const int comp = blockIdx.x;
    if (comp >= ncomp) { return; }
    double ta = 0.0, te = 0.0;
    for (int64_t base = 0; base < nchunks; base += 1024) {
        const int64_t m = (nchunks - base < 1024) ? nchunks - base : 1024;
        for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
            sa[i] = csum_avg[(int64_t)comp * nchunks + base + i];
            se[i] = csum_err2[(int64_t)comp * nchunks + base + i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int64_t i = 0; i < m; ++i) {
                if (base == 0 && i == 0) { ta = sa[0]; te = se[0]; }
                else { ta = ta + sa[i]; te = te + se[i]; }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        avg[comp] = ta / (double)W;
        err[comp] = sqrt(te) / (double)W;
    }
}

// plain ordered sum of the chunk sums (rates), result / W
__global__ void k_mean_of_chunks(const double *chunks, int64_t nchunks, int64_t W, double *out)
{
    __shared__ double sh[1024];
    const double s = ordered_chunk_sum(chunks, nchunks, sh);
    if (threadIdx.x == 0) { *out = s / (double)W; }
}

// convergence test of the automatic decorrelation per independent walker
// (MCIntegrator.cpp:229-240): keep sampling while any |old-new| > 2 sqrt(olderr^2 + newerr^2)
__global__ void k_decorr_compare(double *oa, double *oe, const double *na, const double *ne, int nd, int64_t W, uint8_t *active)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W || !active[w]) { return; }
    bool again = false;
    for (int i = 0; i < nd; ++i) {
        const double o = oa[(int64_t)i * W + w], n = na[(int64_t)i * W + w];
        const double eo = oe[(int64_t)i * W + w], en = ne[(int64_t)i * W + w];
        if (fabs(o - n) > 2 * sqrt(eo * eo + en * en)) { again = true; break; }
    }
    for (int i = 0; i < nd; ++i) {
        oa[(int64_t)i * W + w] = na[(int64_t)i * W + w];
        oe[(int64_t)i * W + w] = ne[(int64_t)i * W + w];
    }
    active[w] = again ? 1 : 0;
}

// MCI::moveX (MCIntegrator.cpp:606-612)
__global__ void k_move_x(Chain c)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) { return; }
    Lane L;
    load_lane(c, L, w);
    const int64_t W = c.W;
    for (int i = 0; i < c.ndim; ++i) { c.xnew[(int64_t)i * W + w] = c.xold[(int64_t)i * W + w]; }
    step_begin(L);
    trial_move(c, L);
    step_end(c, L);
    for (int j = 0; j < c.ndim; ++j) {
        double x = c.xnew[(int64_t)j * W + w];
        if (c.periodic && j >= L.cfirst && j < L.cfirst + L.nchanged) { wrap(c, j, x); }
        c.xold[(int64_t)j * W + w] = x;
    }
    if (c.rng == MCI_RNG_MT19937_64) { c.mtp[w] = L.mtp; } else { c.ctr[w] = L.t; }
    if (c.nsf != nullptr) { // the move's distribution keeps its saved normals
        c.nsf[w] = L.sf;
        c.nsv[w] = L.sv[0];
        c.nsv[W + w] = L.sv[1];
    }
}

// MCI::newRandomX (MCIntegrator.cpp:614-619)
__global__ void k_new_random_x(Chain c)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) { return; }
    Lane L;
    load_lane(c, L, w);
    const int64_t W = c.W;
    step_begin(L);
    for (int i = 0; i < c.ndim; ++i) { c.xnew[(int64_t)i * W + w] = draw_unit(c, L); }
    step_end(c, L);
    for (int i = 0; i < c.ndim; ++i) {
        double x = c.xnew[(int64_t)i * W + w];
        if (c.periodic) { x = c.lb[i] + x * (c.ub[i] - c.lb[i]); }
        else { x = -(double)FLT_MAX + x * ((double)FLT_MAX + (double)FLT_MAX); }
        c.xold[(int64_t)i * W + w] = x;
    }
    if (c.rng == MCI_RNG_MT19937_64) { c.mtp[w] = L.mtp; } else { c.ctr[w] = L.t; }
}

// periodic wrap of every coordinate of xold (MCI::setX/setIRange apply the domain)
__global__ void k_apply_domain(Chain c)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W || !c.periodic) { return; }
    for (int i = 0; i < c.ndim; ++i) { wrap(c, i, c.xold[(int64_t)i * c.W + w]); }
}

// ---------------------------------------------------------------------------
// launchers (host)
// ---------------------------------------------------------------------------
static inline dim3 grid1(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

hipError_t launch_seed_mt(uint64_t *mt, int32_t *mtp, const uint64_t *seeds, int64_t W, hipStream_t s)
{
    hipLaunchKernelGGL(k_seed_mt, grid1(W, 256), dim3(256), 0, s, mt, mtp, seeds, W);
    return hipGetLastError();
}

hipError_t launch_compact_active(const uint8_t *active, int64_t W, int32_t *list, int32_t *count, hipStream_t s)
{
    hipLaunchKernelGGL(k_compact_active, grid1(W, 256), dim3(256), 0, s, active, W, list, count);
    return hipGetLastError();
}

hipError_t launch_walker_columns(const Chain &c, const Chain &dense, const int32_t *list, const int32_t *count, int to_dense,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_walker_columns, grid1(c.W, 256), dim3(256), 0, s, c, dense, list, count, to_dense);
    return hipGetLastError();
}

hipError_t launch_mt_align(uint64_t *mt, int32_t *mtp, int64_t W, hipStream_t s)
{
    const size_t lds = sizeof(uint64_t) * MT_N * MT_ALIGN_TPB;
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_mt_align), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL(k_mt_align, dim3((unsigned)((W + MT_ALIGN_TPB - 1) / MT_ALIGN_TPB)), dim3(MT_ALIGN_TPB), lds, s, mt, mtp, W);
    return hipGetLastError();
}

hipError_t launch_sample(const Chain &c, int64_t nsteps, int acc, const uint8_t *active, hipStream_t s, const CalibCtrl *skip)
{
    hipLaunchKernelGGL(k_sample, grid1(c.W, 64), dim3(64), 0, s, c, nsteps, acc, active, skip);
    return hipGetLastError();
}

hipError_t launch_calibrate_indep(const Chain &c, int nfind, double target, int64_t min_stat, const double *dimSizes, hipStream_t s)
{
    hipLaunchKernelGGL(k_calibrate_indep, grid1(c.W, 64), dim3(64), 0, s, c, nfind, target, min_stat, dimSizes);
    return hipGetLastError();
}

hipError_t launch_scale_steps(const Chain &c, const CalibCtrl *ctrl, const double *dimSizes, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_steps, grid1(c.W, 256), dim3(256), 0, s, c, ctrl, dimSizes);
    return hipGetLastError();
}

hipError_t launch_calib_indep_round(const Chain &c, uint8_t *active, int32_t *cons, int32_t *count, int nfind, double target,
                                    const double *dimSizes, int32_t *any, hipStream_t s)
{
    hipLaunchKernelGGL(k_calib_indep_round, grid1(c.W, 256), dim3(256), 0, s, c, active, cons, count, nfind, target, dimSizes, any);
    return hipGetLastError();
}

hipError_t launch_calib_round(const int64_t *acc, const int64_t *rej, int64_t W, double *chunks, double target, int nfind,
                              CalibCtrl *ctrl, hipStream_t s)
{
    const int64_t nch = (W + 255) / 256;
    hipLaunchKernelGGL(k_calib_round, dim3((unsigned)nch), dim3(256), 0, s, acc, rej, W, chunks, nch, target, nfind, ctrl);
    return hipGetLastError();
}

hipError_t launch_calib_control(const double *chunks, int64_t W, double target, int nfind, CalibCtrl *ctrl, hipStream_t s)
{
    const int64_t nch = (W + 255) / 256;
    hipLaunchKernelGGL(k_calib_control, dim3(1), dim3(256), 0, s, chunks, nch, W, target, nfind, ctrl);
    return hipGetLastError();
}

hipError_t launch_walker_rates(const int64_t *acc, const int64_t *rej, double *rate, int64_t W, hipStream_t s)
{
    hipLaunchKernelGGL(k_walker_rates, grid1(W, 256), dim3(256), 0, s, acc, rej, rate, W);
    return hipGetLastError();
}

hipError_t launch_estimate(const EstimJob &j, hipStream_t s)
{
    const bool mj = j.estim == MCI_ESTIM_MJBLOCKER || (j.estim == MCI_ESTIM_CORRELATED && (j.n & (j.n - 1)) == 0);
    const size_t sh = (mj && j.n <= MJ_LDS_MAX) ? sizeof(double) * 64 * (size_t)j.n : 0;
    const dim3 g = grid1(j.W * j.nobs, 64);
    const bool pow2 = j.n >= 1 && (j.n & (j.n - 1)) == 0;
    if (mj && pow2 && j.n <= 64) {
        const dim3 g256 = grid1(j.W * j.nobs, 256);
        switch (j.n) {
        case 1: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 0>), g256, dim3(256), 0, s, j); break;
        case 2: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 1>), g256, dim3(256), 0, s, j); break;
        case 4: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 2>), g256, dim3(256), 0, s, j); break;
        case 8: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 3>), g256, dim3(256), 0, s, j); break;
        case 16: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 4>), g256, dim3(256), 0, s, j); break;
        case 32: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 5>), g256, dim3(256), 0, s, j); break;
        default: hipLaunchKernelGGL((k_estimate<EST_MJ_REG, 6>), g256, dim3(256), 0, s, j); break;
        }
    }
    else if (mj && j.n <= MJ_LDS_MAX) { hipLaunchKernelGGL(k_estimate<EST_MJ_SMALL>, g, dim3(64), sh, s, j); }
    else if (mj) { hipLaunchKernelGGL(k_estimate<MCI_ESTIM_MJBLOCKER>, g, dim3(64), 0, s, j); }
    else if ((j.estim == MCI_ESTIM_CORRELATED || j.estim == MCI_ESTIM_FCBLOCKER) && j.n == FC_REG_N) {
        if (j.nobs == 1) { hipLaunchKernelGGL((k_estimate<EST_FC_REG, 1>), g, dim3(64), 0, s, j); }
        else { hipLaunchKernelGGL((k_estimate<EST_FC_REG, 0>), g, dim3(64), 0, s, j); }
    }
    else if ((j.estim == MCI_ESTIM_CORRELATED || j.estim == MCI_ESTIM_FCBLOCKER) && j.n <= FC_LDS_MAX) {
        hipLaunchKernelGGL(k_estimate<EST_FC_LDS>, g, dim3(64), sizeof(double) * 64 * (size_t)fc_rows(j.n), s, j);
    }
    else if ((j.estim == MCI_ESTIM_CORRELATED || j.estim == MCI_ESTIM_FCBLOCKER) && j.n <= FC_STREAM_MAX) {
        const dim3 gs = grid1(j.W * j.nobs, FC_STREAM_TPB);
        const size_t lds = fc_stream_lds(j.n);
        if (lds > 65536) {
            // the largest table once per process (the first such call; later calls, under a graph
            // capture too, make no runtime call besides the launch)
            static const hipError_t attr = [] {
                const int mx = (int)fc_stream_lds(FC_STREAM_MAX);
                hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_fc_stream<true>), hipFuncAttributeMaxDynamicSharedMemorySize, mx);
                if (e == hipSuccess) { e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_fc_stream<false>), hipFuncAttributeMaxDynamicSharedMemorySize, mx); }
                return e;
            }();
            if (attr != hipSuccess) { return attr; }
        }
        if (j.nobs == 1) { hipLaunchKernelGGL(k_fc_stream<true>, gs, dim3(FC_STREAM_TPB), lds, s, j); }
        else { hipLaunchKernelGGL(k_fc_stream<false>, gs, dim3(FC_STREAM_TPB), lds, s, j); }
    }
    else if (j.estim == MCI_ESTIM_CORRELATED || j.estim == MCI_ESTIM_FCBLOCKER) { hipLaunchKernelGGL(k_estimate<MCI_ESTIM_FCBLOCKER>, g, dim3(64), 0, s, j); }
    else if (j.estim == MCI_ESTIM_UNCORRELATED) { hipLaunchKernelGGL(k_estimate<MCI_ESTIM_UNCORRELATED>, g, dim3(64), 0, s, j); }
    else { hipLaunchKernelGGL(k_estimate<MCI_ESTIM_NOOP>, g, dim3(64), 0, s, j); }
    return hipGetLastError();
}

hipError_t launch_chunk_sum(const double *v, int64_t W, int ncomp, int square, double *chunks, hipStream_t s)
{
    const int64_t nch = (W + 255) / 256;
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)nch, (unsigned)ncomp), dim3(256), 0, s, v, W, ncomp, square, chunks, nch);
    return hipGetLastError();
}

hipError_t launch_combine(const double *ca, const double *ce, int64_t W, int ncomp, double *avg, double *err, hipStream_t s)
{
    const int64_t nch = (W + 255) / 256;
    if (ncomp <= 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_combine, dim3((unsigned)ncomp), dim3(256), 0, s, ca, ce, nch, ncomp, W, avg, err);
    return hipGetLastError();
}

hipError_t launch_mean_of_chunks(const double *chunks, int64_t W, double *out, hipStream_t s)
{
    const int64_t nch = (W + 255) / 256;
    hipLaunchKernelGGL(k_mean_of_chunks, dim3(1), dim3(256), 0, s, chunks, nch, W, out);
    return hipGetLastError();
}

hipError_t launch_decorr_compare(double *oa, double *oe, const double *na, const double *ne, int nd, int64_t W, uint8_t *active,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_decorr_compare, grid1(W, 256), dim3(256), 0, s, oa, oe, na, ne, nd, W, active);
    return hipGetLastError();
}

hipError_t launch_move_x(const Chain &c, hipStream_t s)
{
    hipLaunchKernelGGL(k_move_x, grid1(c.W, 64), dim3(64), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_new_random_x(const Chain &c, hipStream_t s)
{
    hipLaunchKernelGGL(k_new_random_x, grid1(c.W, 64), dim3(64), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_apply_domain(const Chain &c, hipStream_t s)
{
    hipLaunchKernelGGL(k_apply_domain, grid1(c.W, 256), dim3(256), 0, s, c);
    return hipGetLastError();
}
} // namespace mcid
