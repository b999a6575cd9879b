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
// FP64 for the FP64 estimators.  Every FP32 operation is correctly rounded and exp is the
// library's own bexp (mci_f32math.hpp), so the CPU oracle reproduces the chain bit for bit.
#include "mci_f32math.hpp"
#include "mci_fast.hpp"
#include "mci_fast_plan.hpp"
#include "mci_kernels.hpp"

#include <float.h>
#include <type_traits>

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
    double *steps;         // [W] (one type; rescaled in place by a launch with a pending factor)
    int64_t *acc, *rej;    // [W]
    int64_t blocksize;     // B (Simple: nsteps)
    int64_t nblocks;
    double *d_x2;          // X2 data [nblocks][ndim][W] or nullptr
    double *d_x2sum;       // X2_SUM data [nblocks][1][W] or nullptr
    double *d_exp;         // EXP_X2 data [nblocks][1][W] or nullptr
    double x2sum_div;      // X2_SUM divisor
    double exp_div;        // EXP_X2 divisor
    int nvecs;
    uint32_t lemire_thr;   // 2^32 mod nvecs
    int ibits;             // index bits m in the shared word (nvecs = 2^m), -1: Lemire index
    int vshift;            // 32 - m: the index is the shared word's bits [vshift, 32) (none when m = 0)
    uint32_t ubits;        // max(m, 0): the acceptance uniform's bits start below the index
    uint32_t umask;        // FP64: the shared word's n = 32 - ubits acceptance bits
    float uexp;            // FP64: n, the screening exponent's offset
    const CalibCtrl *pend; // shared calibration: pending step factor, loop finished (FP64 kernel only)
    const double *dimsizes;
    uint64_t *mt;          // reference streams: [312][W] mt19937_64 states (F_MT launches)
    int32_t *mtp;          // [W] next state index
    const uint8_t *active; // [W] or nullptr: walkers whose flag is 0 do nothing (the independent
                           // decorrelation rounds of walkers that have finished theirs)
    const int32_t *wmap;   // or nullptr: the walker each lane's columns belong to (its Philox stream),
    const int32_t *wcount; // and the number of lanes in use, when the walkers are compacted
};

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t &o0, uint32_t &o1, uint32_t &o2, uint32_t &o3)
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
    o0 = c0; o1 = c1; o2 = c2; o3 = c3;
}

// a or b by a lane's flag as arithmetic on the values: a select of two elements of a register array
// can otherwise become a select of their addresses, which puts the array in scratch memory
__device__ __forceinline__ uint32_t pick(bool on, uint32_t a, uint32_t b) { return b ^ ((a ^ b) & (0u - (uint32_t)on)); }

// Lemire index on [0, range) from word w0 (SRRDVecMove.hpp:78 through uniform_int_distribution), for
// a range whose index the shared word does not hold: low = (w0 * range) mod 2^32 < thr = 2^32 mod
// range rejects and draws from the extra-word stream of step t.  thr is computed on the host; a wave
// vote skips the rare path.
__device__ __forceinline__ int draw_lemire(uint32_t w0, uint32_t range, uint32_t thr, uint32_t wid, uint64_t t, uint32_t k0, uint32_t k1)
{
    uint64_t prod = (uint64_t)w0 * range;
    if (__builtin_expect(__any((uint32_t)prod < thr), 0)) {
        uint32_t low = (uint32_t)prod;
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

// observable set of one launch (template flags: no per-step branches on them)
// F_P2: the index is the shared word's top bits (a power-of-two range); F_FULL: Full accumulators (every
// step's values stored)
// F_MT: the reference streams (mt19937_64 per walker) instead of the native ones
enum { F_X2 = 1, F_SUM = 2, F_EXP = 4, F_P2 = 8, F_FULL = 16, F_MT = 32 };

// FP64 chain.  One wave runs alone on its SIMD at the benchmark size (65,536 walkers =
// 1,024 waves), so a step costs about 4 cycles per instruction of any kind and 16 per
// v_mad_u64_u32 (quarter rate; a Philox call holds 20 of them).  The step is written to issue few
// instructions:
//  - steps go in pairs, as the native stream does: one Philox block holds the shared word
//    (index bits, then the acceptance uniform's top bits) and the proposal word of two
//    consecutive steps, so a pair body reads both steps' coordinates from LDS at once, computes
//    the next pair's block while the reads are in flight, then runs the two steps (the second
//    reads what the first wrote when both picked the same coordinate);
//  - no per-step branch on the observable set (template), the steps of one block as an inner
//    loop with a 32-bit accept count;
//  - with X2 the coordinate and its block sum are one 16-byte LDS record (one ds_read_b128 and
//    one ds_write_b128 per step);
//  - the proposal (2 w1 + 1 - 2^32) 2^-32 as one exact fused multiply-add;
//  - the decision screened in FP32 (F = 2^(n - ex log2 e) against the n acceptance bits of the
//    shared word), the exact FP64 test (53-bit u from the side-stream word, FP64 exp) only for
//    lanes inside the band, under a wave vote;
//  - with X2 and X2_SUM both present, the block sum of |x|^2 is the sum of the coordinates' block
//    sums, taken at the block's end (no per-step update).
//
// With Full accumulators (F_FULL, FullAccumulator.cpp:14-18: the reference's default for
// addObservable(obs)) every step stores the whole observable vector: row s of X2 is x_i^2 for all
// i, of X2_SUM the ordered sum (0 + x_0^2 + x_1^2 + ...)/c and of EXP_X2 exp(-that/c), the
// expressions of the generic kernel, so the traces are bit-identical to it.  Each row of a wave is
// one 512-byte store per component; the coordinates are read from LDS, the exp is recomputed only
// when the walker moved.  These launches are bound by HBM writes (136 bytes per walker-step at
// d = 16 with X2 and X2_SUM).

// the running sums of one walker in k_fast64
struct Run64
{
    double r;     // |x|^2
    double Sr;    // block sum of |x|^2 (dwell-weighted)
    double Se, e; // block sum of exp(-|x|^2/c), its current value
    double mult;  // steps of the block a value set now counts for (B - j)
    uint32_t na;  // accepted steps of the chunk
};

// canonical53 of the reference streams (generate_canonical<double, 53> on mt19937_64): w / 2^64,
// clamped below 1 (as in mci_device.hpp)
__device__ __forceinline__ double mt_canonical(uint64_t w)
{
    double r = (double)w / 18446744073709551616.0;
    if (r >= 1.0) { r = 0x1.fffffffffffffp-1; }
    return r;
}

// One step on coordinate v (LDS slot `slot`) from its value xo and block sum So, which the caller
// read and which hold the values after the step on return.  Native streams: w0 is the step's
// shared word, w1 its proposal word, t its counter (the exact test's side-stream word is keyed by
// it).  Reference streams (F_MT): w1 is the proposal's output and w0 the acceptance uniform's.  row:
// the step's index in the run (Full traces).
template <int FL>
__device__ __forceinline__ bool step64(const FastArgs &A, double2 *rec, double *xs, const double *as, Run64 &R, int v, int slot,
                                       double &xo, double &So, float cv, uint64_t w0, uint64_t w1, uint64_t t, int64_t row,
                                       uint32_t wid, double step, int lane, int64_t w)
{
    constexpr bool FULL = (FL & F_FULL) != 0;
    constexpr bool HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    constexpr bool MT = (FL & F_MT) != 0;
    constexpr bool ACC = (FL & 7) != 0 && !FULL && !MT;  // dwell-time sums (the reference streams sum per step)
    constexpr bool REC = HX2 && !FULL;
    // native proposal (2 w1 + 1 - 2^32) 2^-32 = w1 2^-31 - (1 - 2^-32), one exact fused
    // multiply-add; reference streams: uniform_real_distribution(-1, 1) as u 2 + (-1)
    const double rr = MT ? (mt_canonical(w1) * 2.0) + (-1.0) : fma((double)(uint32_t)w1, 0x1p-31, -(1.0 - 0x1p-32));
    const double xn = xo + step * rr;
    const double xn2 = xn * xn, xo2 = xo * xo;
    const double dx2n = xn2 - xo2;  // the change of x_v^2 if accepted
    // The FP64 test is u <= exp(-ex) with ex = a_v xn^2 - a_v xo^2 (pdf_acceptance's form).  The
    // shared word's low n = 32 - m bits are u's top bits: u lies in [U 2^-n, (U + 1) 2^-n).  The
    // screen takes F = 2^(n - a_v log2 e (xn^2 - xo^2)) in FP32: relative error below 2e-5 wherever
    // F >= 1 (the FP64 rounding of ex is far smaller), and U rounded to FP32.  U < 0.9999 F - 2
    // makes the FP64 test certain to accept, U > 1.0001 F certain to reject; between the two the
    // FP64 test decides.
    // Reference streams: u = w0 / 2^64 rounded to 53 bits lies in [(U - 1) 2^-32, (U + 2) 2^-32)
    // for U the top word of w0, hence the margins 3 and 1.
    const float U = MT ? (float)(uint32_t)(w0 >> 32) : (float)((uint32_t)w0 & A.umask);
    const float F = __builtin_amdgcn_exp2f(fmaf((float)dx2n, cv, MT ? 32.0f : A.uexp));
    const float Facc = MT ? fmaf(F, 0.9999f, -3.0f) : fmaf(F, 0.9999f, -2.0f);
    const float Frej = MT ? fmaf(F, 1.0001f, 1.0f) : F * 1.0001f;
    const bool acc = U < Facc;
    const bool band = !acc && U <= Frej;
    if (acc) { ++R.na; }
    bool moved = acc;
    double xsel = acc ? xn : xo, Snew = So;
    if (ACC) {
        // the new value counts for the block's remaining steps (one fused multiply-add: the block
        // means agree with the per-step sums of the generic kernel to about 1e-15 relative)
        const double dx2 = acc ? dx2n : 0.0;
        Snew = fma(dx2, R.mult, So);
        if (REC) { rec[slot] = make_double2(xsel, Snew); }
        else { xs[slot] = xsel; }
        R.r = R.r + dx2;
        if (!(HX2 && HSUM)) { R.Sr = fma(dx2, R.mult, R.Sr); }
        if (HEXP && acc) { const double en = exp(-(R.r / A.exp_div)); R.Se = R.Se + (en - R.e) * R.mult; R.e = en; }
    }
    else if (REC) { rec[slot] = make_double2(xsel, So); }  // the reference streams' per-step sums: the caller adds
    else { xs[slot] = xsel; }
    // Lanes inside the band went through the update above as rejections (which leaves every number
    // as it was: +0.0 added to non-negative sums).  When the wave holds any, the exact test runs and
    // the lanes it accepts are updated as accepted steps, from the values read at the start of the
    // step.  (two ballots of plain compares: each is the compare's own lane mask)
    if (__builtin_expect((__builtin_amdgcn_ballot_w64(U <= Frej) & ~__builtin_amdgcn_ballot_w64(acc)) != 0, 0)) {
        double u;
        if (MT) { u = (mt_canonical(w0) * 1.0) + 0.0; }
        else {
            uint32_t l0, l1, l2, l3;
            philox(0x30000000u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, l0, l1, l2, l3);
            const uint64_t uw = (((uint64_t)(uint32_t)w0 << 32) | l0) << A.ubits;
            u = (double)(uw >> 11) * 0x1p-53;
        }
        const double av = as[v];
        const double ex = xn2 * av - xo2 * av;
        if (band && u <= exp(-ex)) {
            R.na += 1u;
            moved = true;
            xsel = xn;
            if (ACC) {
                Snew = fma(dx2n, R.mult, So);
                if (REC) { rec[slot] = make_double2(xn, Snew); }
                else { xs[slot] = xn; }
                R.r = R.r + dx2n;
                if (!(HX2 && HSUM)) { R.Sr = fma(dx2n, R.mult, R.Sr); }
                if (HEXP) { const double en = exp(-(R.r / A.exp_div)); R.Se = R.Se + (en - R.e) * R.mult; R.e = en; }
            }
            else if (REC) { rec[slot] = make_double2(xn, So); }
            else { xs[slot] = xn; }
        }
    }
    if (ACC) { R.mult = R.mult - 1.0; }
    if (FULL) {
        // this step's row of every trace (the observable values after the step)
        double q2s = 0.0;
        for (int i = 0; i < A.ndim; ++i) {
            const double xi = xs[i * WAVE + lane];
            const double q2 = xi * xi;
            if (HX2) { A.d_x2[(row * A.ndim + i) * A.W + w] = q2; }
            q2s = q2s + q2;
        }
        if (HSUM) { A.d_x2sum[row * A.W + w] = q2s / A.x2sum_div; }
        if (HEXP) {
            if (moved) { R.e = exp(-(q2s / A.exp_div)); }
            A.d_exp[row * A.W + w] = R.e;
        }
    }
    xo = xsel;
    So = Snew;
    return moved;
}

// Steps t and t + 1 (t even) of k_fast64: c = blocks 0 (and 1 without a shared index) of their
// pair, n receives the next pair's.  Both steps' records are read at once, the next blocks are
// computed while the reads are in flight, then the two steps run; the second reads what the first
// wrote when both picked the same coordinate.
template <int FL>
__device__ __forceinline__ void pair64(const FastArgs &A, double2 *rec, double *xs, const double *as, const float *cs, Run64 &R,
                                       const uint32_t (&c)[6], uint32_t (&n)[6], uint64_t t, int64_t row, uint32_t wid, double step,
                                       int lane, int64_t w)
{
    constexpr bool REC = (FL & F_X2) != 0 && (FL & F_FULL) == 0;
    constexpr bool P2 = (FL & F_P2) != 0;
    // c[0..3]: block 0 (shared and next word of step t, then of step t + 1); c[4], c[5]: words 0
    // and 2 of block 1 (the proposal words when the index is drawn by Lemire's method)
    int vA, vB;
    if (P2) {
        vA = (int)__builtin_amdgcn_ubfe(c[0], A.vshift, A.ibits);
        vB = (int)__builtin_amdgcn_ubfe(c[2], A.vshift, A.ibits);
    }
    else {
        vA = draw_lemire(c[1], (uint32_t)A.nvecs, A.lemire_thr, wid, t, A.key0, A.key1);
        vB = draw_lemire(c[3], (uint32_t)A.nvecs, A.lemire_thr, wid, t + 1, A.key0, A.key1);
    }
    const int sA = vA * WAVE + lane, sB = vB * WAVE + lane;
    double xA, SA = 0.0, xB, SB = 0.0;
    if (REC) {
        const double2 ra = rec[sA], rb = rec[sB];
        xA = ra.x; SA = ra.y; xB = rb.x; SB = rb.y;
    }
    else { xA = xs[sA]; xB = xs[sB]; }
    const float cA = cs[vA], cB = cs[vB];
    const uint64_t tn = (t + 2) >> 1;
    uint32_t m1, m3;
    philox(0u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, n[0], n[1], n[2], n[3]);
    if (!P2) { philox(1u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, n[4], m1, n[5], m3); }
    // the next blocks are needed here, so that their computation fills the reads' latency (instead
    // of sinking below the first step's branch)
    asm volatile("" ::"v"(n[0]), "v"(n[1]), "v"(n[2]), "v"(n[3]), "v"(n[4]), "v"(n[5]));
    step64<FL>(A, rec, xs, as, R, vA, sA, xA, SA, cA, c[0], P2 ? c[1] : c[4], t, row, wid, step, lane, w);
    if (vB == vA) { xB = xA; SB = SA; }  // the second step reads what the first wrote
    step64<FL>(A, rec, xs, as, R, vB, sB, xB, SB, cB, c[2], P2 ? c[3] : c[5], t + 1, row + 1, wid, step, lane, w);
}

template <int FL>
__global__ __launch_bounds__(WAVE) void k_fast64(FastArgs A)
{
    constexpr bool FULL = (FL & F_FULL) != 0;
    constexpr bool HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    constexpr bool ACC = (FL & 7) != 0 && !FULL;  // block sums (dwell time; per step with MT)
    constexpr bool REC = HX2 && !FULL;            // 16-byte {x_i, block sum} records
    constexpr bool P2 = (FL & F_P2) != 0;         // the index is the shared word's top bits
    constexpr bool MT = (FL & F_MT) != 0;         // reference streams
    extern __shared__ double lds[];
    const int d = A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    const bool live = w < A.W && (!A.wcount || w < *A.wcount);
    if (A.pend) {
        // the previous calibration iteration's factor (k_scale_steps with one step type), then
        // nothing more once the loop has finished
        if (A.pend->apply && live) {
            double st = A.steps[w] * A.pend->fact;
            for (int i = 0; i < d; ++i) {
                if (st > 0.5 * A.dimsizes[i]) { st = 0.5 * A.dimsizes[i]; }
            }
            if (st < (double)FLT_MIN) { st = (double)FLT_MIN; }
            A.steps[w] = st;
        }
        if (A.pend->done) { return; }
    }
    const int64_t B = A.blocksize;
    double2 *rec = reinterpret_cast<double2 *>(lds);  // X2 blocks: [d][64] {x_i, block sum of x_i^2}
    double *xs = lds;                                  // otherwise [d][64] x_i
    double *as = lds + (REC ? 2 : 1) * d * WAVE;       // [d] a_i
    float *cs = reinterpret_cast<float *>(as + d);     // [d] -a_i log2 e in FP32 (screening exponent)
    for (int i = lane; i < d; i += WAVE) {
        as[i] = A.a[i];
        cs[i] = (float)(-A.a[i] * 0x1.71547652b82fep+0);
    }
    Run64 R{0.0, 0.0, 0.0, 0.0, (double)B, 0u};
    if (live) {
        for (int i = 0; i < d; ++i) {
            const double v = A.x64[(int64_t)i * A.W + w];
            if (REC) { rec[i * WAVE + lane] = make_double2(v, MT ? 0.0 : (double)B * (v * v)); }
            else { xs[i * WAVE + lane] = v; }
            R.r = R.r + v * v;
        }
    }
    __syncthreads();
    if (!live || (A.active && !A.active[w])) { return; }
    const double step = A.steps[w];
    [[maybe_unused]] const uint32_t wid = A.wmap ? (uint32_t)A.wmap[w] : (uint32_t)w;
    [[maybe_unused]] const uint32_t range = (uint32_t)A.nvecs, thr = A.lemire_thr;
    [[maybe_unused]] uint64_t t = MT ? 0 : A.ctr[w];
    int64_t nacc = 0;
    int64_t b = 0;
    if (ACC && !MT) { R.Sr = (double)B * R.r; }
    if (HEXP && ACC) { R.e = exp(-(R.r / A.exp_div)); R.Se = MT ? 0.0 : (double)B * R.e; }
    if (HEXP && FULL) {
        double q0 = 0.0;
        for (int i = 0; i < d; ++i) { const double v = xs[i * WAVE + lane]; q0 = q0 + v * v; }
        R.e = exp(-(q0 / A.exp_div));
    }
    int64_t inblk = 0;  // steps done in the current block
    // a block's end: its means out, |x|^2 recomputed exactly, the sums restarted
    auto flush_block = [&]() {
        const double norm = 1. / (double)B;
        double rx = 0.0, sS = 0.0;
        for (int i = 0; i < d; ++i) {
            double v2;
            if (REC) {
                const double2 rv = rec[i * WAVE + lane];
                v2 = rv.x;
                sS = sS + rv.y;
                A.d_x2[(b * d + i) * A.W + w] = rv.y * norm;
                rec[i * WAVE + lane] = make_double2(v2, (double)B * (v2 * v2));
            }
            else { v2 = xs[i * WAVE + lane]; }
            rx = rx + v2 * v2;
        }
        if (HSUM) { A.d_x2sum[b * A.W + w] = ((HX2 ? sS : R.Sr) * norm) / A.x2sum_div; }
        if (HEXP) { A.d_exp[b * A.W + w] = R.Se * norm; }
        R.r = rx;  // refresh |x|^2 exactly at every block boundary
        R.Sr = (double)B * R.r;
        if (HEXP) { R.e = exp(-(R.r / A.exp_div)); R.Se = (double)B * R.e; }
        R.mult = (double)B;
        inblk = 0;
        ++b;
    };
    if constexpr (MT) {
        // Reference streams: three outputs of the walker's mt19937_64 per step (the index, the
        // proposal, the acceptance uniform; a power-of-two index never rejects).  The lazy twist
        // of mci_device.hpp's mt_next, in groups of MT_G steps: the loads of a group's state words
        // are issued a group ahead, so they are in flight while the group before runs.
        constexpr int MT_G = 4, NO = 3 * MT_G;
        const int64_t W = A.W;
        // The walkers of a wave are usually at the same state position (every step takes three
        // outputs; seeding starts all at 0); then the position is wave-uniform and every state
        // access is a uniform row pointer plus the lane (scalar address arithmetic).
        const int kw = A.mtp[w];
        const int kf = __builtin_amdgcn_readfirstlane(kw);
        const bool uniform = __builtin_amdgcn_ballot_w64(kw != kf) == 0;
        const int64_t wbase = (int64_t)blockIdx.x * WAVE;
        auto run = [&](auto uni_c) {
            constexpr bool UNI = decltype(uni_c)::value;
            auto at = [&](int kk) -> uint64_t * {
                if constexpr (UNI) { return (A.mt + ((int64_t)kk * W + wbase)) + lane; }
                else { return A.mt + ((int64_t)kk * W + w); }
            };
            int k = UNI ? kf : kw;
            uint64_t xk = *at(k);   // the current value of element k
            uint64_t x1[NO], xm[NO], o[NO];
            auto issue = [&](int kk) {
#pragma unroll
                for (int j = 0; j < NO; ++j) {
                    const int k1 = (kk + 1 == MT_N) ? 0 : kk + 1;
                    const int km = (kk < MT_N - MT_M) ? kk + MT_M : kk + MT_M - MT_N;
                    x1[j] = *at(k1);
                    xm[j] = *at(km);
                    kk = k1;
                }
            };
            issue(k);
            for (int64_t s = 0; s < A.nsteps; s += MT_G) {
                const int ng = (int)(A.nsteps - s < MT_G ? A.nsteps - s : MT_G);
                // this group's outputs, exactly the 3 ng it consumes (the state advances by those)
#pragma unroll
                for (int j = 0; j < NO; ++j) {
                    if (j < 3 * ng) {
                        const uint64_t nv = mt_twist(j == 0 ? xk : x1[j - 1], x1[j], xm[j]);
                        *at(k) = nv;
                        o[j] = mt_temper(nv);
                        k = (k + 1 == MT_N) ? 0 : k + 1;
                    }
                }
#pragma unroll
                for (int j = 0; j < NO; ++j) {
                    if (j == 3 * ng - 1) { xk = x1[j]; }
                }
                if (s + MT_G < A.nsteps) { issue(k); }
                R.na = 0;
#pragma unroll
                for (int j = 0; j < MT_G; ++j) {
                    if (j < ng) {
                        const int v = A.ibits > 0 ? (int)(o[3 * j] >> (64 - A.ibits)) : 0;
                        const int sl = v * WAVE + lane;
                        double xo, So = 0.0;
                        if (REC) { const double2 rv = rec[sl]; xo = rv.x; So = rv.y; }
                        else { xo = xs[sl]; }
                        const bool moved = step64<FL>(A, rec, xs, as, R, v, sl, xo, So, cs[v], o[3 * j + 2], o[3 * j + 1], 0, s + j, 0u,
                                                      step, lane, w);
                        if (ACC) {
                            // the reference's per-step sums (BlockAccumulator / SimpleAccumulator): every
                            // component's value after the step added in step order, so the block means
                            // are bit-identical to the reference's
                            double q = 0.0;
                            for (int i = 0; i < d; ++i) {
                                if (REC) {
                                    double2 rv = rec[i * WAVE + lane];
                                    const double q2 = rv.x * rv.x;
                                    q = q + q2;
                                    rv.y = rv.y + q2;
                                    rec[i * WAVE + lane] = rv;
                                }
                                else {
                                    const double xv = xs[i * WAVE + lane];
                                    q = q + xv * xv;
                                }
                            }
                            if (HSUM) { R.Sr = R.Sr + q / A.x2sum_div; }
                            if (HEXP) {
                                if (moved) { R.e = exp(-(q / A.exp_div)); }
                                R.Se = R.Se + R.e;
                            }
                            if (++inblk == B) {
                                const double norm = 1. / (double)B;
                                if (REC) {
                                    for (int i = 0; i < d; ++i) {
                                        const double2 rv = rec[i * WAVE + lane];
                                        A.d_x2[(b * d + i) * A.W + w] = rv.y * norm;
                                        rec[i * WAVE + lane] = make_double2(rv.x, 0.0);
                                    }
                                }
                                if (HSUM) { A.d_x2sum[b * A.W + w] = R.Sr * norm; }
                                if (HEXP) { A.d_exp[b * A.W + w] = R.Se * norm; }
                                R.Sr = 0.0;
                                R.Se = 0.0;
                                inblk = 0;
                                ++b;
                            }
                        }
                    }
                }
                nacc += R.na;
            }
            A.mtp[w] = k;
        };
        if (uniform) { run(std::integral_constant<bool, true>{}); }
        else { run(std::integral_constant<bool, false>{}); }
    }
    else {
        // Blocks 0 and (without a shared index) 1 of the current step pair t / 2, as pair64 takes them:
        // p[0], p[1] are step t's shared and next word when t is even, p[2], p[3] when t is odd; p[4] /
        // p[5] the word after (block 1's words 0 / 2).
        uint32_t p[6] = {0u, 0u, 0u, 0u, 0u, 0u}, nx[6] = {0u, 0u, 0u, 0u, 0u, 0u};
        uint32_t g1, g3;
        philox(0u, wid, (uint32_t)(t >> 1), (uint32_t)(t >> 33), A.key0, A.key1, p[0], p[1], p[2], p[3]);
        if (!P2) { philox(1u, wid, (uint32_t)(t >> 1), (uint32_t)(t >> 33), A.key0, A.key1, p[4], g1, p[5], g3); }
        // The lanes of a wave step in pairs when their counters agree in parity (always, unless the
        // counters were set so); otherwise every step runs alone, each lane on its own parity.
        const uint64_t podd = __builtin_amdgcn_ballot_w64(((uint32_t)t & 1u) != 0);
        const bool uni = podd == 0 || podd == __builtin_amdgcn_ballot_w64(true);
        // Blocks start at step 0.  The steps run in chunks of at most 2^30 that never cross a block
        // end, so the inner loop's counters are 32-bit whatever the block length (a Simple
        // accumulator's block is the whole run).
        for (int64_t s = 0; s < A.nsteps;) {
            const int len = fast_chunk_len(A.nsteps, s, B, inblk, ACC);
            R.na = 0;
            // one step alone, on the lane's own parity
            auto single = [&](int q) {
                const bool odd = ((uint32_t)t & 1u) != 0;
                const uint32_t ws = pick(odd, p[2], p[0]), wn = pick(odd, p[3], p[1]);
                int v;
                uint32_t wp;
                if (P2) { v = (int)__builtin_amdgcn_ubfe(ws, A.vshift, A.ibits); wp = wn; }
                else {
                    v = draw_lemire(wn, range, thr, wid, t, A.key0, A.key1);
                    wp = pick(odd, p[5], p[4]);
                }
                const int sl = v * WAVE + lane;
                double xo, So = 0.0;
                if (REC) { const double2 rv = rec[sl]; xo = rv.x; So = rv.y; }
                else { xo = xs[sl]; }
                step64<FL>(A, rec, xs, as, R, v, sl, xo, So, cs[v], ws, wp, t, s + q, wid, step, lane, w);
                ++t;
                if (odd) {
                    const uint64_t tn = t >> 1;
                    philox(0u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, p[0], p[1], p[2], p[3]);
                    if (!P2) { philox(1u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, p[4], g1, p[5], g3); }
                }
            };
            int q = 0;
            if (uni) {
                // an odd counter first takes one step; then groups of four steps (two pairs, the blocks
                // alternating between p and nx), then the rest alone
                if (((uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)t) & 1u) != 0u && q < len) { single(q); ++q; }
                for (; q + 4 <= len; q += 4) {
                    pair64<FL>(A, rec, xs, as, cs, R, p, nx, t, s + q, wid, step, lane, w);
                    pair64<FL>(A, rec, xs, as, cs, R, nx, p, t + 2, s + q + 2, wid, step, lane, w);
                    t += 4;
                }
            }
            for (; q < len; ++q) { single(q); }
            nacc += R.na;
            s += len;
            inblk += len;
            if (ACC && inblk == B) { flush_block(); }
        }
        A.ctr[w] = t;
    }
    for (int i = 0; i < d; ++i) { A.x64[(int64_t)i * A.W + w] = REC ? rec[i * WAVE + lane].x : xs[i * WAVE + lane]; }
    A.acc[w] = nacc;
    A.rej[w] = A.nsteps - nacc;
}

// All-coordinate moves (SRRDAllMove with one step size: MoveType::All, the reference's default
// move), FP64, both stream modes.  Every step proposes every coordinate, so the state is read and
// written whole: [d][64] positions (with X2 16-byte {x_i, block sum} records), the proposal in
// [d][64] beside them, the proto sum of the accepted state in a register.  The step follows the
// generic kernel's arithmetic: x'_i = x_i + step r_i in coordinate order, p'_i = (x'_i x'_i) a_i,
// e = (0 + sum p_i) - (0 + sum p'_i), accept u <= exp(e); block sums are the reference's per-step
// sums, so every number is bit-identical to the generic kernel.  Random draws: native streams --
// the shared word 0 carries the acceptance uniform's top 32 bits (no index shares it), words 1 .. d
// the proposals; reference streams -- d proposal outputs, then the acceptance uniform.  The screen
// of k_fast64 decides all but a band of lanes, which take the exact FP64 test.
template <int FL>
__global__ __launch_bounds__(WAVE) void k_fast64_all(FastArgs A)
{
    constexpr bool FULL = (FL & F_FULL) != 0;
    constexpr bool HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    constexpr bool ACC = (FL & 7) != 0 && !FULL;  // per-step block sums
    constexpr bool REC = HX2 && !FULL;
    constexpr bool MT = (FL & F_MT) != 0;
    extern __shared__ double lds[];
    const int d = A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    const bool live = w < A.W && (!A.wcount || w < *A.wcount);
    if (A.pend) {
        if (A.pend->apply && live) {
            double st = A.steps[w] * A.pend->fact;
            for (int i = 0; i < d; ++i) {
                if (st > 0.5 * A.dimsizes[i]) { st = 0.5 * A.dimsizes[i]; }
            }
            if (st < (double)FLT_MIN) { st = (double)FLT_MIN; }
            A.steps[w] = st;
        }
        if (A.pend->done) { return; }
    }
    const int64_t B = A.blocksize;
    double2 *rec = reinterpret_cast<double2 *>(lds);  // X2: [d][64] {x_i, block sum of x_i^2}
    double *xs = lds;                                  // otherwise [d][64] x_i
    double *xn = lds + (REC ? 2 : 1) * d * WAVE;       // [d][64] the proposal
    double *as = xn + d * WAVE;                        // [d] a_i, then the reference streams' ring
    for (int i = lane; i < d; i += WAVE) { as[i] = A.a ? A.a[i] : 1.0; }
    auto xget = [&](int i) -> double { return REC ? rec[i * WAVE + lane].x : xs[i * WAVE + lane]; };
    double Sold = 0.0, e = 0.0, Sr = 0.0, Se = 0.0;
    if (live) {
        for (int i = 0; i < d; ++i) {
            const double v = A.x64[(int64_t)i * A.W + w];
            if (REC) { rec[i * WAVE + lane] = make_double2(v, 0.0); }
            else { xs[i * WAVE + lane] = v; }
        }
    }
    __syncthreads();
    if (!live || (A.active && !A.active[w])) { return; }
    double q0 = 0.0;
    for (int i = 0; i < d; ++i) {
        const double v = xget(i);
        Sold = Sold + (v * v) * as[i];
        q0 = q0 + v * v;
    }
    if (HEXP) { e = exp(-(q0 / A.exp_div)); }
    const double step = A.steps[w];
    const uint32_t wid = A.wmap ? (uint32_t)A.wmap[w] : (uint32_t)w;
    const int64_t W = A.W;
    [[maybe_unused]] uint64_t t = MT ? 0 : A.ctr[w];
    // Reference streams: the walker's mt19937_64 outputs in chunks of MT_C into a per-lane LDS ring
    // (a step takes d + 1 of them, in order).  The lazy twist of mci_device.hpp's mt_next, with the
    // state words of a chunk loaded a chunk ahead; the chunks produce exactly the outputs the launch
    // consumes, so the state ends where the generic kernel's does.
    constexpr int MT_C = 16;
    [[maybe_unused]] uint64_t *ring = reinterpret_cast<uint64_t *>(as + d);  // [MT_C][64]
    [[maybe_unused]] int k = 0, pos = MT_C;
    [[maybe_unused]] uint64_t xk = 0;
    [[maybe_unused]] int64_t need = 0;
    [[maybe_unused]] uint64_t x1[MT_C], xm[MT_C];
    auto issue = [&](int kk) {
#pragma unroll
        for (int j = 0; j < MT_C; ++j) {
            const int k1 = (kk + 1 == MT_N) ? 0 : kk + 1;
            const int km = (kk < MT_N - MT_M) ? kk + MT_M : kk + MT_M - MT_N;
            x1[j] = A.mt[(int64_t)k1 * W + w];
            xm[j] = A.mt[(int64_t)km * W + w];
            kk = k1;
        }
    };
    auto mt_out = [&]() -> uint64_t {
        if (pos == MT_C) {
            const int64_t c = need < MT_C ? need : MT_C;
#pragma unroll
            for (int j = 0; j < MT_C; ++j) {
                if (j < c) {
                    const uint64_t nv = mt_twist(j == 0 ? xk : x1[j - 1], x1[j], xm[j]);
                    A.mt[(int64_t)k * W + w] = nv;
                    ring[j * WAVE + lane] = mt_temper(nv);
                    k = (k + 1 == MT_N) ? 0 : k + 1;
                }
            }
#pragma unroll
            for (int j = 0; j < MT_C; ++j) {
                if (j == c - 1) { xk = x1[j]; }
            }
            need -= c;
            if (need > 0) { issue(k); }
            pos = 0;
        }
        return ring[(pos++) * WAVE + lane];
    };
    if constexpr (MT) {
        k = A.mtp[w];
        xk = A.mt[(int64_t)k * W + w];
        need = A.nsteps * (int64_t)(d + 1);
        if (need > 0) { issue(k); }
    }
    int64_t nacc = 0, inblk = 0, b = 0, s = 0;
    // a step's decision from the proto sum of its proposal (in xn) and its acceptance word (native:
    // the shared word; reference streams: the acceptance output), then the observables after it
    auto tail = [&](double Snew, uint64_t uword) {
        const double ee = Sold - Snew;  // log of the acceptance ratio, as the generic kernel forms it
        const float U = MT ? (float)(uint32_t)(uword >> 32) : (float)(uint32_t)uword;
        const float F = __builtin_amdgcn_exp2f(fmaf((float)ee, 0x1.715476p+0f, 32.0f));
        const float Facc = MT ? fmaf(F, 0.9999f, -3.0f) : fmaf(F, 0.9999f, -2.0f);
        const float Frej = MT ? fmaf(F, 1.0001f, 1.0f) : F * 1.0001f;
        bool acc = U < Facc;
        const bool band = !acc && U <= Frej;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(band) != 0, 0)) {
            double u;
            if (MT) { u = (mt_canonical(uword) * 1.0) + 0.0; }
            else {
                uint32_t l0, l1, l2, l3;
                philox(0x30000000u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, l0, l1, l2, l3);
                u = (double)((((uint64_t)(uint32_t)uword << 32) | l0) >> 11) * 0x1p-53;
            }
            if (band) { acc = u <= exp(ee); }
        }
        if (!MT) { ++t; }
        if (acc) {
            ++nacc;
            Sold = Snew;
#pragma unroll 4
            for (int i = 0; i < d; ++i) {
                const double v = xn[i * WAVE + lane];
                if (REC) { rec[i * WAVE + lane].x = v; }
                else { xs[i * WAVE + lane] = v; }
            }
        }
        // the observables after the step: the reference's per-step sums, or the trace row
        if (ACC || FULL) {
            double q = 0.0;
#pragma unroll 4
            for (int i = 0; i < d; ++i) {
                if (REC) {
                    double2 rv = rec[i * WAVE + lane];
                    const double q2 = rv.x * rv.x;
                    q = q + q2;
                    rv.y = rv.y + q2;
                    rec[i * WAVE + lane] = rv;
                }
                else {
                    const double xv = xs[i * WAVE + lane];
                    const double q2 = xv * xv;
                    if (FULL && HX2) { A.d_x2[(s * d + i) * W + w] = q2; }
                    q = q + q2;
                }
            }
            if (HEXP && acc) { e = exp(-(q / A.exp_div)); }
            if (FULL) {
                if (HSUM) { A.d_x2sum[s * W + w] = q / A.x2sum_div; }
                if (HEXP) { A.d_exp[s * W + w] = e; }
            }
            else {
                if (HSUM) { Sr = Sr + q / A.x2sum_div; }
                if (HEXP) { Se = Se + e; }
                if (++inblk == B) {
                    const double norm = 1. / (double)B;
                    if (REC) {
                        for (int i = 0; i < d; ++i) {
                            const double2 rv = rec[i * WAVE + lane];
                            A.d_x2[(b * d + i) * W + w] = rv.y * norm;
                            rec[i * WAVE + lane] = make_double2(rv.x, 0.0);
                        }
                    }
                    if (HSUM) { A.d_x2sum[b * W + w] = Sr * norm; }
                    if (HEXP) { A.d_exp[b * W + w] = Se * norm; }
                    Sr = 0.0;
                    Se = 0.0;
                    inblk = 0;
                    ++b;
                }
            }
        }
        ++s;
    };
    // coordinate i of the proposal from the random value r, its proto added to Sn in coordinate order
    auto prop = [&](int i, double r, double &Sn) {
        const double v = xget(i) + step * r;
        xn[i * WAVE + lane] = v;
        Sn = Sn + (v * v) * as[i];
    };
    if constexpr (MT) {
        while (s < A.nsteps) {
            double Snew = 0.0;
            for (int i = 0; i < d; ++i) { prop(i, (mt_canonical(mt_out()) * 2.0) + (-1.0), Snew); }
            tail(Snew, mt_out());
        }
    }
    else {
        // main word j of step t: word 2 (t mod 2) + (j mod 2) of block j / 2 of pair t / 2.  Steps go
        // in pairs: the first computes the blocks and keeps words 2 and 3 of each for the second.
        uint32_t *wb = reinterpret_cast<uint32_t *>(as + d);  // [d + 2][64] the second step's words
        auto rnd = [](uint32_t v) { return fma((double)v, 0x1p-31, -(1.0 - 0x1p-32)); };
        // MODE 0: a step alone (either parity), 1: the first of a pair, 2: the second of a pair
        // the words j and j + 1 of the step (block j / 2)
        auto words = [&](auto mode_c, int j, uint32_t par, uint64_t tp, uint32_t &wa, uint32_t &wv) {
            constexpr int MODE = decltype(mode_c)::value;
            if constexpr (MODE == 2) {
                wa = wb[j * WAVE + lane];
                wv = wb[(j + 1) * WAVE + lane];
            }
            else {
                uint32_t o0, o1, o2, o3;
                philox((uint32_t)(j >> 1), wid, (uint32_t)tp, (uint32_t)(tp >> 32), A.key0, A.key1, o0, o1, o2, o3);
                if constexpr (MODE == 1) {
                    wa = o0;
                    wv = o1;
                    wb[j * WAVE + lane] = o2;
                    wb[(j + 1) * WAVE + lane] = o3;
                }
                else {
                    wa = par ? o2 : o0;
                    wv = par ? o3 : o1;
                }
            }
        };
        auto step_words = [&](auto mode_c) {
            const uint32_t par = (uint32_t)(t & 1u);
            const uint64_t tp = t >> 1;
            double Sn = 0.0;
            uint32_t uw, wa, wv;
            // block 0: the acceptance word and coordinate 0
            words(mode_c, 0, par, tp, uw, wv);
            prop(0, rnd(wv), Sn);
            // then two coordinates a block, their LDS reads issued a block ahead (the next block's
            // coordinates, clamped into range, are read while this block's Philox rounds run)
            int j = 2;
            double x0n = xget(1 < d ? 1 : d - 1), x1n = xget(2 < d ? 2 : d - 1), a0n = as[1 < d ? 1 : d - 1], a1n = as[2 < d ? 2 : d - 1];
            for (; j + 1 <= d; j += 2) {
                const double x0 = x0n, x1 = x1n, a0 = a0n, a1 = a1n;
                const int n0 = j + 1 < d ? j + 1 : d - 1, n1 = j + 2 < d ? j + 2 : d - 1;
                x0n = xget(n0);
                x1n = xget(n1);
                a0n = as[n0];
                a1n = as[n1];
                words(mode_c, j, par, tp, wa, wv);
                const double v0 = x0 + step * rnd(wa);
                xn[(j - 1) * WAVE + lane] = v0;
                Sn = Sn + (v0 * v0) * a0;
                const double v1 = x1 + step * rnd(wv);
                xn[j * WAVE + lane] = v1;
                Sn = Sn + (v1 * v1) * a1;
            }
            if (j == d) { // d even: word d alone (coordinate d - 1)
                words(mode_c, j, par, tp, wa, wv);
                prop(d - 1, rnd(wa), Sn);
            }
            tail(Sn, uw);
        };
        // pairs need the wave's counters to agree in parity (always, unless a restored state set them apart)
        const uint64_t podd = __builtin_amdgcn_ballot_w64((t & 1u) != 0);
        const bool pairs = podd == 0 || podd == __builtin_amdgcn_ballot_w64(true);
        if (pairs) {
            if ((t & 1u) && s < A.nsteps) { step_words(std::integral_constant<int, 0>{}); }
            while (s + 2 <= A.nsteps) {
                step_words(std::integral_constant<int, 1>{});
                step_words(std::integral_constant<int, 2>{});
            }
        }
        while (s < A.nsteps) { step_words(std::integral_constant<int, 0>{}); }
    }
    for (int i = 0; i < d; ++i) { A.x64[(int64_t)i * W + w] = xget(i); }
    if constexpr (MT) { A.mtp[w] = k; }
    else { A.ctr[w] = t; }
    A.acc[w] = nacc;
    A.rej[w] = A.nsteps - nacc;
}

// The all-coordinate kernel with the walker in registers, for d <= DM: positions, the proposal and
// the block sums are DM-entry register arrays, every loop over the coordinates unrolled to DM with a
// wave-uniform d guard, so a step touches no LDS.  The arithmetic, the draws and the per-step sums
// are those of k_fast64_all (bit-identical to the generic kernel).  Native streams: steps go in
// pairs, one Philox block per two words of each of the two steps (the first step keeps words 2 and
// 3 of every block for the second), so a pair computes ceil((d + 1) / 2) blocks.  Reference streams:
// the walker's mt19937_64 outputs are made 32 at a time (their state words loaded a chunk ahead)
// into a per-lane LDS ring of 64; a step reads its d + 1 outputs from the ring.
// EX: d == DM exactly (every coordinate guard folds away, the step is straight-line code).
template <int FL, int DM, bool EX>
__global__ __launch_bounds__(WAVE) void k_fast64_allr(FastArgs A)
{
    constexpr bool FULL = (FL & F_FULL) != 0;
    constexpr bool HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    constexpr bool ACC = (FL & 7) != 0 && !FULL;  // per-step block sums
    constexpr bool REC = HX2 && !FULL;             // per-coordinate block sums
    constexpr bool MT = (FL & F_MT) != 0;
    constexpr int NB = DM / 2 + 1;                 // Philox blocks of a step pair at most
    extern __shared__ double lds[];
    const int d = EX ? DM : A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    const bool live = w < A.W && (!A.wcount || w < *A.wcount);
    if (A.pend) {
        if (A.pend->apply && live) {
            double st = A.steps[w] * A.pend->fact;
            for (int i = 0; i < d; ++i) {
                if (st > 0.5 * A.dimsizes[i]) { st = 0.5 * A.dimsizes[i]; }
            }
            if (st < (double)FLT_MIN) { st = (double)FLT_MIN; }
            A.steps[w] = st;
        }
        if (A.pend->done) { return; }
    }
    if (!live || (A.active && !A.active[w])) { return; }
    const int64_t W = A.W, B = A.blocksize;
    double x[DM], xn[DM], sm[DM], a[DM];
    double Sold = 0.0, q = 0.0, e = 0.0, Sr = 0.0, Se = 0.0;
#pragma unroll
    for (int i = 0; i < DM; ++i) {
        x[i] = 0.0;
        xn[i] = 0.0;
        sm[i] = 0.0;
        a[i] = 0.0;
        if (i < d) {
            x[i] = A.x64[(int64_t)i * W + w];
            a[i] = A.a ? A.a[i] : 1.0;
            Sold = Sold + (x[i] * x[i]) * a[i];
            q = q + x[i] * x[i];
        }
    }
    if (HEXP) { e = exp(-(q / A.exp_div)); }
    const double step = A.steps[w];
    const uint32_t wid = A.wmap ? (uint32_t)A.wmap[w] : (uint32_t)w;
    [[maybe_unused]] uint64_t t = MT ? 0 : A.ctr[w];
    int64_t nacc = 0, inblk = 0, b = 0, s = 0;

    // the decision of a step from the proto sum Snew and |x'|^2 qn of its proposal (xn), then the
    // observables after it; uw: the native shared word or the reference acceptance output
    auto finish = [&](double Snew, double qn, uint64_t uw) {
        const double ee = Sold - Snew;
        const float U = MT ? (float)(uint32_t)(uw >> 32) : (float)(uint32_t)uw;
        const float F = __builtin_amdgcn_exp2f(fmaf((float)ee, 0x1.715476p+0f, 32.0f));
        const float Facc = MT ? fmaf(F, 0.9999f, -3.0f) : fmaf(F, 0.9999f, -2.0f);
        const float Frej = MT ? fmaf(F, 1.0001f, 1.0f) : F * 1.0001f;
        bool acc = U < Facc;
        const bool band = !acc && U <= Frej;
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(band) != 0, 0)) {
            double u;
            if (MT) { u = (mt_canonical(uw) * 1.0) + 0.0; }
            else {
                uint32_t l0, l1, l2, l3;
                philox(0x30000000u, wid, (uint32_t)t, (uint32_t)(t >> 32), A.key0, A.key1, l0, l1, l2, l3);
                u = (double)((((uint64_t)(uint32_t)uw << 32) | l0) >> 11) * 0x1p-53;
            }
            if (band) { acc = u <= exp(ee); }
        }
        if (!MT) { ++t; }
        nacc += acc ? 1 : 0;
        Sold = acc ? Snew : Sold;
        q = acc ? qn : q;
#pragma unroll
        for (int i = 0; i < DM; ++i) {
            if (i < d) { x[i] = acc ? xn[i] : x[i]; }
        }
        if (HEXP && acc) { e = exp(-(q / A.exp_div)); }
        if (FULL) {
            if (HX2) {
#pragma unroll
                for (int i = 0; i < DM; ++i) {
                    if (i < d) { A.d_x2[(s * d + i) * W + w] = x[i] * x[i]; }
                }
            }
            if (HSUM) { A.d_x2sum[s * W + w] = q / A.x2sum_div; }
            if (HEXP) { A.d_exp[s * W + w] = e; }
        }
        if (ACC) {
            if (REC) {
#pragma unroll
                for (int i = 0; i < DM; ++i) {
                    if (i < d) { sm[i] = sm[i] + x[i] * x[i]; }
                }
            }
            if (HSUM) { Sr = Sr + q / A.x2sum_div; }
            if (HEXP) { Se = Se + e; }
            if (++inblk == B) {
                const double norm = 1. / (double)B;
                if (REC) {
#pragma unroll
                    for (int i = 0; i < DM; ++i) {
                        if (i < d) {
                            A.d_x2[(b * d + i) * W + w] = sm[i] * norm;
                            sm[i] = 0.0;
                        }
                    }
                }
                if (HSUM) { A.d_x2sum[b * W + w] = Sr * norm; }
                if (HEXP) { A.d_exp[b * W + w] = Se * norm; }
                Sr = 0.0;
                Se = 0.0;
                inblk = 0;
                ++b;
            }
        }
        ++s;
    };
    // coordinate i of the proposal from the random value r (the generic kernel's xn + st * r, then
    // the proto (xn xn) a_i and |x'|^2 summed in coordinate order)
    auto propose = [&](int i, double r, double &Sn, double &qn) {
        const double v = x[i] + step * r;
        xn[i] = v;
        Sn = Sn + (v * v) * a[i];
        qn = qn + v * v;
    };

    if constexpr (MT) {
        constexpr int MT_C = 32, RING = 64;
        uint64_t *ring = reinterpret_cast<uint64_t *>(lds);  // [RING][64]
        const int kw = A.mtp[w];
        // The walkers of a wave are usually at the same state position (every step of this kernel
        // takes d + 1 outputs; seeding starts all at 0); then the position is wave-uniform and every
        // state access is a uniform row pointer plus the lane (scalar address arithmetic).
        const int kf = __builtin_amdgcn_readfirstlane(kw);
        const bool uniform = __builtin_amdgcn_ballot_w64(kw != kf) == 0;
        const int64_t wbase = (int64_t)blockIdx.x * WAVE;
        auto run = [&](auto uni_c) {
            constexpr bool UNI = decltype(uni_c)::value;
            int k = UNI ? kf : kw;
            auto at = [&](int kk) -> uint64_t * {
                if constexpr (UNI) { return (A.mt + ((int64_t)kk * W + wbase)) + lane; }
                else { return A.mt + ((int64_t)kk * W + w); }
            };
            uint64_t xk = *at(k);
            uint64_t x1[MT_C], xm[MT_C];
            int64_t need = A.nsteps * (int64_t)(d + 1);
            auto issue = [&](int kk) {
#pragma unroll
                for (int j = 0; j < MT_C; ++j) {
                    const int k1 = (kk + 1 == MT_N) ? 0 : kk + 1;
                    const int km = (kk < MT_N - MT_M) ? kk + MT_M : kk + MT_M - MT_N;
                    x1[j] = *at(k1);
                    xm[j] = *at(km);
                    kk = k1;
                }
            };
            if (need > 0) { issue(k); }
            int head = 0, tail = 0, avail = 0;
            for (int64_t st = 0; st < A.nsteps; ++st) {
                while (avail < d + 1) {
                    // the next chunk: exactly the outputs still to be consumed, at most MT_C
                    const int c = (int)(need < MT_C ? need : MT_C);
#pragma unroll
                    for (int j = 0; j < MT_C; ++j) {
                        if (j < c) {
                            const uint64_t nv = mt_twist(j == 0 ? xk : x1[j - 1], x1[j], xm[j]);
                            *at(k) = nv;
                            ring[((head + j) & (RING - 1)) * WAVE + lane] = mt_temper(nv);
                            k = (k + 1 == MT_N) ? 0 : k + 1;
                        }
                    }
#pragma unroll
                    for (int j = 0; j < MT_C; ++j) {
                        if (j == c - 1) { xk = x1[j]; }
                    }
                    head = (head + c) & (RING - 1);
                    avail += c;
                    need -= c;
                    if (need > 0) { issue(k); }
                }
                double Sn = 0.0, qn = 0.0;
#pragma unroll
                for (int i = 0; i < DM; ++i) {
                    if (i < d) { propose(i, (mt_canonical(ring[((tail + i) & (RING - 1)) * WAVE + lane]) * 2.0) + (-1.0), Sn, qn); }
                }
                const uint64_t uw = ring[((tail + d) & (RING - 1)) * WAVE + lane];
                tail = (tail + d + 1) & (RING - 1);
                avail -= d + 1;
                finish(Sn, qn, uw);
            }
            A.mtp[w] = k;
        };
        if (uniform) { run(std::integral_constant<bool, true>{}); }
        else { run(std::integral_constant<bool, false>{}); }
    }
    else {
        // main word j of step t: word 2 (t mod 2) + (j mod 2) of block j / 2 of pair t / 2; the
        // native proposal (2 w + 1 - 2^32) 2^-32 as one exact fused multiply-add
        auto rnd = [](uint32_t v) { return fma((double)v, 0x1p-31, -(1.0 - 0x1p-32)); };
        const int nb = (d + 2) >> 1;
        uint32_t sw[2 * NB];  // the second step's words of a pair
        // one step on its own (either parity)
        auto single = [&]() {
            const uint32_t par = (uint32_t)(t & 1u);
            const uint64_t tp = t >> 1;
            double Sn = 0.0, qn = 0.0;
            uint32_t uw = 0;
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) {
                if (bi < nb) {
                    uint32_t o0, o1, o2, o3;
                    philox((uint32_t)bi, wid, (uint32_t)tp, (uint32_t)(tp >> 32), A.key0, A.key1, o0, o1, o2, o3);
                    const uint32_t wa = par ? o2 : o0, wb = par ? o3 : o1;
                    if (bi == 0) { uw = wa; }
                    else { propose(2 * bi - 1, rnd(wa), Sn, qn); }
                    if (2 * bi < d) { propose(2 * bi, rnd(wb), Sn, qn); }
                }
            }
            finish(Sn, qn, uw);
        };
        if ((t & 1u) && s < A.nsteps) { single(); }
        while (s + 2 <= A.nsteps) {
            const uint64_t tp = t >> 1;
            double Sn = 0.0, qn = 0.0;
            uint32_t uw = 0;
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) {
                if (bi < nb) {
                    uint32_t o0, o1, o2, o3;
                    philox((uint32_t)bi, wid, (uint32_t)tp, (uint32_t)(tp >> 32), A.key0, A.key1, o0, o1, o2, o3);
                    sw[2 * bi] = o2;
                    sw[2 * bi + 1] = o3;
                    if (bi == 0) { uw = o0; }
                    else { propose(2 * bi - 1, rnd(o0), Sn, qn); }
                    if (2 * bi < d) { propose(2 * bi, rnd(o1), Sn, qn); }
                }
            }
            finish(Sn, qn, uw);
            Sn = 0.0;
            qn = 0.0;
#pragma unroll
            for (int bi = 0; bi < NB; ++bi) {
                if (bi < nb) {
                    if (bi == 0) { uw = sw[0]; }
                    else { propose(2 * bi - 1, rnd(sw[2 * bi]), Sn, qn); }
                    if (2 * bi < d) { propose(2 * bi, rnd(sw[2 * bi + 1]), Sn, qn); }
                }
            }
            finish(Sn, qn, uw);
        }
        if (s < A.nsteps) { single(); }
        A.ctr[w] = t;
    }
#pragma unroll
    for (int i = 0; i < DM; ++i) {
        if (i < d) { A.x64[(int64_t)i * W + w] = x[i]; }
    }
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

// Kahan step applied only where `on` holds (selects keep the common path branch-free; a Kahan
// step with y = 0 is not an identity when the compensation is nonzero)
__device__ __forceinline__ void kahan_if(float &s, float &c, float y, bool on)
{
    const float yy = y - c;
    const float tt = s + yy;
    const float cc = (tt - s) - yy;
    s = on ? tt : s;
    c = on ? cc : c;
}

// Kahan step adding the exact product a b: its rounding error (fused multiply-add, exact) goes
// into the compensation first.  Dwell weights b reach the block length, so an uncompensated
// FP32 product would lose about 2^-24 b a per accepted change.
__device__ __forceinline__ void kahan_prod(float &s, float &c, float a, float b)
{
    const float p = a * b;
    const float pe = fmaf(a, b, -p);
    const float yy = p - (c - pe);
    const float tt = s + yy;
    c = (tt - s) - yy;
    s = tt;
}

__device__ __forceinline__ void kahan_prod_if(float &s, float &c, float a, float b, bool on)
{
    const float p = a * b;
    const float pe = fmaf(a, b, -p);
    const float yy = p - (c - pe);
    const float tt = s + yy;
    const float cc = (tt - s) - yy;
    s = on ? tt : s;
    c = on ? cc : c;
}


// Compensations of the FP32 block sums are stored in BF16 (round to nearest even): a Kahan
// compensation is below half an ulp of its sum, so its top 8 bits keep the pair exact to about
// 2^-33 of the sum, and the 16 bits left over hold the chunk step of the coordinate's last change.
__device__ __forceinline__ uint32_t bf16_bits(float v)
{
    const uint32_t u = __float_as_uint(v);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

// FP32 chain.  Block sums are sums of positive terms: a value (x_i^2, |x|^2 or exp(-|x|^2/c))
// enters its block sum once, when it is replaced or its chunk ends, times the steps it was held
// (its dwell).  Each such term is rounded once and enters a Kahan pair exactly (the product's
// rounding error, by fused multiply-add, goes into the compensation), so the block mean's
// relative error stays about 1e-7 whatever the block length.  (Adding the changes x_new^2 -
// x_old^2 times the remaining steps instead would sum terms much larger than the block sum, of
// both signs, and lose precision with the block length.)  A run is cut into chunks of at most
// FAST32_REFRESH steps that never cross a block end; every chunk end closes all dwells and
// recomputes |x|^2 from the position.  With X2 the per-coordinate state is 12 bytes in three
// [coordinate][lane] arrays: x_i, the block sum of x_i^2, and (BF16 compensation, step of the last
// change in the chunk) -- 48 KiB at d = 64, so three workgroups share a CU.
// the running sums of one walker in k_fast32 (chunk steps jr, je: when the current |x|^2 and exp
// values were set)
struct Run32
{
    float r, rc;   // |x|^2, Kahan-compensated
    float Sr, Cr;  // block sum of |x|^2
    float Se, Ce;  // block sum of exp(-|x|^2/c)
    float e;
    int jr, je;
    uint32_t na;   // accepted steps of the chunk
};

// One FP32 step at chunk step q on coordinate v (LDS slot `slot`) from its value xo, block sum So
// and packed compensation / last change cj, which the caller read and which hold the values after
// the step on return.  ub: the 24 acceptance bits, wp: the proposal word.
template <int FL>
__device__ __forceinline__ void step32(const FastArgs &A, float *xs, float *Ss, uint32_t *CJ, const float *as, Run32 &R, int v, int slot,
                                       float &xo, float &So, uint32_t &cj, uint32_t ub, uint32_t wp, int q, float step, float inv_ediv)
{
    constexpr bool ACC = (FL & 7) != 0, HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    const float asv = as[v];
    // proposal on the symmetric 24-bit grid (2m + 1 - 2^24) 2^-24, acceptance u on 2^-24 (the 24
    // bits of the shared word below the index)
    const float rr = (float)(2 * (int32_t)(wp >> 8) + 1 - 16777216) * 0x1p-24f;
    const float u = (float)ub * 0x1p-24f;
    const float xn = fmaf(step, rr, xo);
    const float dx2 = (xn - xo) * (xn + xo);
    const float ex = asv * dx2;
    // u <= bexp(-ex), this library's own exp (the CPU evaluates the same operations).  The hardware
    // exp2 screens it: F = 2^(-ex log2 e) is within 1e-5 relative of bexp(-ex) wherever bexp is
    // neither 0 nor inf (the rounding of ex log2 e contributes at most 127 * 9e-8 to the exponent),
    // so u < 0.9998 F accepts and u > 1.0002 F rejects with certainty; lanes between the two (and
    // F == 0, flushed) take bexp under a wave vote.
    const float F = __builtin_amdgcn_exp2f(ex * -0x1.715476p+0f);
    bool acc = u < F * 0.9998f;
    const bool band = !acc && !(u > F * 1.0002f);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(band) != 0, 0)) {
        if (band) { acc = u <= mcif::bexp(-ex); }
    }
    R.na += acc ? 1u : 0u;
    const float xsel = acc ? xn : xo;
    xs[slot] = xsel;
    if (ACC) {
        if (HX2) {
            // the old x_v^2 was held since the chunk step in the low half of cj
            const int jl = (int)(cj & 0xFFFFu);
            float Co = __uint_as_float(cj & 0xFFFF0000u);
            kahan_prod_if(So, Co, xo * xo, (float)(q - jl), acc);
            Ss[slot] = So;
            cj = (bf16_bits(Co) << 16) | (acc ? (uint32_t)q : (uint32_t)jl);
            CJ[slot] = cj;
        }
        if (HSUM) {
            kahan_prod_if(R.Sr, R.Cr, R.r, (float)(q - R.jr), acc);
            R.jr = acc ? q : R.jr;
        }
        kahan_if(R.r, R.rc, dx2, acc);
        if (HEXP && acc) { kahan_prod(R.Se, R.Ce, R.e, (float)(q - R.je)); R.je = q; R.e = mcif::bexp(-R.r * inv_ediv); }
    }
    xo = xsel;
}

// Steps t and t + 1 (t even) of k_fast32 at chunk steps q, q + 1, as pair64 does for the FP64 chain
template <int FL>
__device__ __forceinline__ void pair32(const FastArgs &A, float *xs, float *Ss, uint32_t *CJ, const float *as, Run32 &R,
                                       const uint32_t (&c)[6], uint32_t (&n)[6], uint64_t t, int q, uint32_t wid, float step,
                                       float inv_ediv, int lane)
{
    constexpr bool HX2 = (FL & F_X2) != 0;
    constexpr bool P2 = (FL & F_P2) != 0;
    int vA, vB;
    uint32_t uA, uB, wA, wB;
    if (P2) {
        vA = (int)__builtin_amdgcn_ubfe(c[0], A.vshift, A.ibits);
        vB = (int)__builtin_amdgcn_ubfe(c[2], A.vshift, A.ibits);
        uA = (c[0] << A.ubits) >> 8;
        uB = (c[2] << A.ubits) >> 8;
        wA = c[1];
        wB = c[3];
    }
    else {
        vA = draw_lemire(c[1], (uint32_t)A.nvecs, A.lemire_thr, wid, t, A.key0, A.key1);
        vB = draw_lemire(c[3], (uint32_t)A.nvecs, A.lemire_thr, wid, t + 1, A.key0, A.key1);
        uA = c[0] >> 8;
        uB = c[2] >> 8;
        wA = c[4];
        wB = c[5];
    }
    const int sA = vA * WAVE + lane, sB = vB * WAVE + lane;
    float xA = xs[sA], xB = xs[sB], SA = 0.f, SB = 0.f;
    uint32_t jA = 0u, jB = 0u;
    if (HX2) { SA = Ss[sA]; jA = CJ[sA]; SB = Ss[sB]; jB = CJ[sB]; }
    const uint64_t tn = (t + 2) >> 1;
    uint32_t m1, m3;
    philox(0u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, n[0], n[1], n[2], n[3]);
    if (!P2) { philox(1u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, n[4], m1, n[5], m3); }
    asm volatile("" ::"v"(n[0]), "v"(n[1]), "v"(n[2]), "v"(n[3]), "v"(n[4]), "v"(n[5]));
    step32<FL>(A, xs, Ss, CJ, as, R, vA, sA, xA, SA, jA, uA, wA, q, step, inv_ediv);
    if (vB == vA) { xB = xA; SB = SA; jB = jA; }  // the second step reads what the first wrote
    step32<FL>(A, xs, Ss, CJ, as, R, vB, sB, xB, SB, jB, uB, wB, q + 1, step, inv_ediv);
}

template <int FL>
__global__ __launch_bounds__(WAVE) void k_fast32(FastArgs A)
{
    constexpr bool ACC = (FL & 7) != 0, HX2 = (FL & F_X2) != 0, HSUM = (FL & F_SUM) != 0, HEXP = (FL & F_EXP) != 0;
    extern __shared__ float lds32[];
    const int d = A.ndim;
    const int lane = threadIdx.x;
    const int64_t w = (int64_t)blockIdx.x * WAVE + lane;
    const bool live = w < A.W && (!A.wcount || w < *A.wcount);
    float *xs = lds32;                                                // [d][64] x_i
    float *Ss = lds32 + d * WAVE;                                     // X2: [d][64] block sums of x_i^2
    uint32_t *CJ = reinterpret_cast<uint32_t *>(lds32 + 2 * d * WAVE); // X2: [d][64] BF16 compensation << 16 | last change
    float *as = lds32 + (HX2 ? 3 : 1) * d * WAVE;                     // [d]
    for (int i = lane; i < d; i += WAVE) { as[i] = (float)A.a[i]; }
    float r = 0.f, rc = 0.f;                                          // |x|^2, Kahan-compensated
    if (live) {
        for (int i = 0; i < d; ++i) {
            const float v = (float)A.x64[(int64_t)i * A.W + w];
            xs[i * WAVE + lane] = v;
            if (HX2) { Ss[i * WAVE + lane] = 0.f; CJ[i * WAVE + lane] = 0u; }
            kahan(r, rc, v * v);
        }
    }
    __syncthreads();
    if (!live || (A.active && !A.active[w])) { return; }
    const float step = (float)A.steps[w];
    uint64_t t = A.ctr[w];
    const uint32_t wid = A.wmap ? (uint32_t)A.wmap[w] : (uint32_t)w;
    const uint32_t range = (uint32_t)A.nvecs, thr = A.lemire_thr;
    const int B = (int)A.blocksize;  // at most 2^31 - 1 (checked by the host)
    const float inv_ediv = (float)(1.0 / A.exp_div);
    Run32 R{r, rc, 0.f, 0.f, 0.f, 0.f, 0.f, 0, 0, 0u};
    int64_t nacc = 0;
    int j = 0;  // steps done in the block
    int64_t b = 0;
    if (HEXP) { R.e = mcif::bexp(-R.r * inv_ediv); }
    // Blocks 0 and (without a shared index) 1 of the current step pair t / 2, laid out as for the
    // FP64 kernel (pair64)
    constexpr bool P2 = (FL & F_P2) != 0;
    uint32_t p[6] = {0u, 0u, 0u, 0u, 0u, 0u}, nx[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    uint32_t g1, g3;
    philox(0u, wid, (uint32_t)(t >> 1), (uint32_t)(t >> 33), A.key0, A.key1, p[0], p[1], p[2], p[3]);
    if (!P2) { philox(1u, wid, (uint32_t)(t >> 1), (uint32_t)(t >> 33), A.key0, A.key1, p[4], g1, p[5], g3); }
    const uint64_t podd = __builtin_amdgcn_ballot_w64(((uint32_t)t & 1u) != 0);
    const bool uni = podd == 0 || podd == __builtin_amdgcn_ballot_w64(true);
    for (int64_t s = 0; s < A.nsteps;) {
        const int len = fast32_chunk_len(A.nsteps, s, B, j, ACC);
        auto single = [&](int q) {
            const bool odd = ((uint32_t)t & 1u) != 0;
            const uint32_t ws = pick(odd, p[2], p[0]), wn = pick(odd, p[3], p[1]);
            int v;
            uint32_t wp, ub;
            if (P2) {
                v = (int)__builtin_amdgcn_ubfe(ws, A.vshift, A.ibits);
                wp = wn;
                ub = (ws << A.ubits) >> 8;
            }
            else {
                v = draw_lemire(wn, range, thr, wid, t, A.key0, A.key1);
                wp = pick(odd, p[5], p[4]);
                ub = ws >> 8;
            }
            const int slot = v * WAVE + lane;
            float xo = xs[slot], So = 0.f;
            uint32_t cj = 0u;
            if (HX2) { So = Ss[slot]; cj = CJ[slot]; }
            step32<FL>(A, xs, Ss, CJ, as, R, v, slot, xo, So, cj, ub, wp, q, step, inv_ediv);
            ++t;
            if (odd) {
                const uint64_t tn = t >> 1;
                philox(0u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, p[0], p[1], p[2], p[3]);
                if (!P2) { philox(1u, wid, (uint32_t)tn, (uint32_t)(tn >> 32), A.key0, A.key1, p[4], g1, p[5], g3); }
            }
        };
        int q = 0;
        if (uni) {
            if (((uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)t) & 1u) != 0u && q < len) { single(q); ++q; }
            for (; q + 4 <= len; q += 4) {
                pair32<FL>(A, xs, Ss, CJ, as, R, p, nx, t, q, wid, step, inv_ediv, lane);
                pair32<FL>(A, xs, Ss, CJ, as, R, nx, p, t + 2, q + 2, wid, step, inv_ediv, lane);
                t += 4;
            }
        }
        for (; q < len; ++q) { single(q); }
        nacc += R.na;  // chunks hold at most FAST32_REFRESH steps
        R.na = 0;
        s += len;
        if (ACC) {
            // the chunk ends: every value held counts until here; at a block end the sums flush
            // to FP64 block means, else they carry on into the next chunk
            j += len;
            const bool bend = (j == B);
            const double norm = 1. / B;
            float rx = 0.f, rxc = 0.f;
            for (int i = 0; i < d; ++i) {
                const float v2 = xs[i * WAVE + lane];
                if (HX2) {
                    float Si = Ss[i * WAVE + lane];
                    const uint32_t c2 = CJ[i * WAVE + lane];
                    float Ci = __uint_as_float(c2 & 0xFFFF0000u);
                    kahan_prod(Si, Ci, v2 * v2, (float)(len - (int)(c2 & 0xFFFFu)));
                    if (bend) {
                        A.d_x2[(b * d + i) * A.W + w] = ((double)Si - (double)Ci) * norm;
                        Ss[i * WAVE + lane] = 0.f;
                        CJ[i * WAVE + lane] = 0u;
                    }
                    else {
                        Ss[i * WAVE + lane] = Si;
                        CJ[i * WAVE + lane] = bf16_bits(Ci) << 16;
                    }
                }
                kahan(rx, rxc, v2 * v2);
            }
            if (HSUM) { kahan_prod(R.Sr, R.Cr, R.r, (float)(len - R.jr)); }
            if (HEXP) { kahan_prod(R.Se, R.Ce, R.e, (float)(len - R.je)); }
            if (bend) {
                if (HSUM) { A.d_x2sum[b * A.W + w] = (((double)R.Sr - (double)R.Cr) * norm) / A.x2sum_div; }
                if (HEXP) { A.d_exp[b * A.W + w] = ((double)R.Se - (double)R.Ce) * norm; }
                R.Sr = R.Cr = R.Se = R.Ce = 0.f;
                j = 0;
                ++b;
            }
            R.jr = R.je = 0;
            R.r = rx;
            R.rc = rxc;
            if (HEXP) { R.e = mcif::bexp(-R.r * inv_ediv); }
        }
    }
    for (int i = 0; i < d; ++i) { A.x64[(int64_t)i * A.W + w] = (double)xs[i * WAVE + lane]; }
    A.ctr[w] = t;
    A.acc[w] = nacc;
    A.rej[w] = A.nsteps - nacc;
}

size_t lds_bytes(int prec, int ndim, bool x2)
{
    if (prec == MCI_PREC_FP64) { return sizeof(double) * ((size_t)(x2 ? 2 : 1) * ndim * WAVE + ndim) + sizeof(float) * ndim; }
    return sizeof(float) * ((size_t)(x2 ? 3 : 1) * ndim * WAVE + ndim);
}

// the all-coordinate kernels: registers up to ALL_DM coordinates (LDS: the reference streams' ring);
// beyond, positions (X2 blocks: records), proposals, a_i and the ring in LDS
constexpr int ALL_DM = 16;
size_t lds_bytes_all(int ndim, bool x2, bool mt)
{
    if (ndim <= ALL_DM) { return mt ? sizeof(uint64_t) * 64 * WAVE : 0; }
    return sizeof(double) * ((size_t)(x2 ? 2 : 1) * ndim * WAVE + (size_t)ndim * WAVE + ndim) +
           (mt ? sizeof(uint64_t) * 16 * WAVE : sizeof(uint32_t) * (size_t)(ndim + 2) * WAVE);
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

template <int FL>
void launch_fl(int prec, dim3 grid, size_t sh, hipStream_t s, const FastArgs &A)
{
    if (prec == MCI_PREC_FP64) { hipLaunchKernelGGL(k_fast64<FL>, grid, dim3(WAVE), sh, s, A); }
    else if constexpr ((FL & (F_FULL | F_MT)) == 0) { hipLaunchKernelGGL(k_fast32<FL>, grid, dim3(WAVE), sh, s, A); }
}

template <int FL>
void launch_all(dim3 grid, size_t sh, hipStream_t s, const FastArgs &A)
{
    if constexpr ((FL & F_MT) == 0) {
        if (A.ndim == ALL_DM) {
            hipLaunchKernelGGL((k_fast64_allr<FL, ALL_DM, true>), grid, dim3(WAVE), sh, s, A);
            return;
        }
    }
    if (A.ndim <= ALL_DM) { hipLaunchKernelGGL((k_fast64_allr<FL, ALL_DM, false>), grid, dim3(WAVE), sh, s, A); }
    else { hipLaunchKernelGGL(k_fast64_all<FL>, grid, dim3(WAVE), sh, s, A); }
}
} // namespace

bool fast_eligible(int prec, int rng, int movetype, int veclen, int ntypes, int periodic, int npdf, int pdfkind, int nvecs)
{
    // the reference streams: the FP64 chain, and a power-of-two coordinate count, whose index draw
    // (uniform_int_distribution, Lemire on 64-bit words) never rejects, so a step takes three outputs
    const bool streams = rng == MCI_RNG_PHILOX || (rng == MCI_RNG_MT19937_64 && prec == MCI_PREC_FP64 && nvecs >= 1 && (nvecs & (nvecs - 1)) == 0);
    const bool target = ntypes == 1 && !periodic && npdf == 1 && pdfkind == MCI_PDF_GAUSS_DIAG;
    // all-coordinate moves (one step size): FP64, either stream
    // (exp(-|x|^2) is the diagonal form with a_i = 1: the same protos, sums and ratio bit for bit)
    if (movetype == MCI_MOVE_ALL) {
        return ntypes == 1 && !periodic && npdf == 1 && (pdfkind == MCI_PDF_GAUSS_DIAG || pdfkind == MCI_PDF_GAUSS_ISO) &&
               prec == MCI_PREC_FP64 && (rng == MCI_RNG_PHILOX || rng == MCI_RNG_MT19937_64);
    }
    return streams && movetype == MCI_MOVE_VEC && veclen == 1 && target;
}

bool fast_obs_eligible(const ObsDev *obs, int nobs, int ndim, int prec, int movetype, int rng)
{
    int B = -1;
    int seen = 0;
    bool x2 = false;
    for (int i = 0; i < nobs; ++i) {
        const ObsDev &o = obs[i];
        if (o.nskip != 1 || (o.acc == ACC_FULL && prec != MCI_PREC_FP64)) { return false; }
        if (o.kind != MCI_OBS_X2 && o.kind != MCI_OBS_X2_SUM && o.kind != MCI_OBS_EXP_X2) { return false; }
        const int bit = 1 << o.kind;
        if (seen & bit) { return false; }
        seen |= bit;
        x2 = x2 || o.kind == MCI_OBS_X2;
        const int bs = (int)(o.acc == ACC_SIMPLE ? -2 : o.blocksize);
        if (B == -1) { B = bs; }
        else if (B != bs) { return false; }
    }
    // B == 1: Full accumulators, no block-sum records
    if (movetype == MCI_MOVE_ALL) { return lds_bytes_all(ndim, x2 && B != 1, rng == MCI_RNG_MT19937_64) <= lds_limit(); }
    return lds_bytes(prec, ndim, x2 && B != 1) <= lds_limit();
}

int fast_sample(FastState &fs, const Chain &c, int prec, int64_t nsteps, const ObsDev *obs, int nobs, hipStream_t s,
                const CalibCtrl *pend, const double *dimsizes, const uint8_t *active, const int32_t *wmap,
                const int32_t *wcount)
{
    (void)fs;
    if (pend && (prec != MCI_PREC_FP64 || !dimsizes || active)) { return MCI_E_INVALID_ARGUMENT; }
    FastArgs A;
    A.active = active;
    A.wmap = wmap;
    A.wcount = wcount;
    A.pend = pend;
    A.dimsizes = dimsizes;
    A.W = c.W;
    A.ndim = c.ndim;
    A.nsteps = nsteps;
    A.key0 = c.key0;
    A.key1 = c.key1;
    A.ctr = c.ctr;
    A.x64 = c.xold;
    A.a = c.pdf[0].kind == MCI_PDF_GAUSS_ISO ? nullptr : c.pdf[0].a;  // exp(-|x|^2): a_i = 1
    A.steps = c.steps;
    A.acc = c.acc;
    A.rej = c.rej;
    A.nvecs = c.nvecs;
    A.lemire_thr = c.nvecs > 0 ? (0u - (uint32_t)c.nvecs) % (uint32_t)c.nvecs : 0u;  // all-coordinate moves: no index
    // the native stream's shared word: the index's bits (FP32: at most 8), then the acceptance bits;
    // the reference streams take the index from the top log2(nvecs) bits of its own output
    const bool mt = c.rng == MCI_RNG_MT19937_64;
    A.ibits = prec == MCI_PREC_FP64 ? c.ibits : shared_index_bits32(c.ibits);
    if (mt) {
        A.ibits = 0;
        while ((1 << A.ibits) < c.nvecs) { ++A.ibits; }
    }
    A.mt = c.mt;
    A.mtp = c.mtp;
    A.vshift = A.ibits >= 0 ? 32 - A.ibits : 0;
    A.ubits = A.ibits > 0 ? (uint32_t)A.ibits : 0u;
    A.umask = 0xFFFFFFFFu >> A.ubits;
    A.uexp = (float)(32 - (int)A.ubits);
    A.blocksize = 1;
    A.nblocks = 0;
    A.d_x2 = A.d_x2sum = A.d_exp = nullptr;
    A.x2sum_div = 1.0;
    A.exp_div = (double)c.ndim;
    int fl = 0;
    for (int i = 0; i < nobs && obs; ++i) {
        const ObsDev &o = obs[i];
        if (o.acc == ACC_FULL) { fl |= F_FULL; }
        A.blocksize = (o.acc == ACC_SIMPLE) ? nsteps : (int64_t)o.blocksize;
        A.nblocks = o.nstore;
        if (o.kind == MCI_OBS_X2) { A.d_x2 = o.data; fl |= F_X2; }
        if (o.kind == MCI_OBS_X2_SUM) { A.d_x2sum = o.data; A.x2sum_div = o.p0; fl |= F_SUM; }
        if (o.kind == MCI_OBS_EXP_X2) { A.d_exp = o.data; A.exp_div = o.p0; fl |= F_EXP; }
    }
    if (!A.a && c.movetype != MCI_MOVE_ALL) { return MCI_E_INVALID_ARGUMENT; } // single-coordinate moves: explicit weights
    if (prec != MCI_PREC_FP64 && A.blocksize > INT32_MAX) { return MCI_E_INVALID_ARGUMENT; }
    const dim3 grid((unsigned)((c.W + WAVE - 1) / WAVE));
    if (c.movetype == MCI_MOVE_ALL) {
        if (prec != MCI_PREC_FP64 || (mt && (!c.mt || !c.mtp))) { return MCI_E_INVALID_ARGUMENT; }
        if (mt) { fl |= F_MT; }
        const size_t sha = lds_bytes_all(c.ndim, (fl & F_X2) != 0 && !(fl & F_FULL), mt);
        switch (fl) {
        case 0: launch_all<0>(grid, sha, s, A); break;
        case 1: launch_all<1>(grid, sha, s, A); break;
        case 2: launch_all<2>(grid, sha, s, A); break;
        case 3: launch_all<3>(grid, sha, s, A); break;
        case 4: launch_all<4>(grid, sha, s, A); break;
        case 5: launch_all<5>(grid, sha, s, A); break;
        case 6: launch_all<6>(grid, sha, s, A); break;
        case 7: launch_all<7>(grid, sha, s, A); break;
        case 17: launch_all<17>(grid, sha, s, A); break;
        case 18: launch_all<18>(grid, sha, s, A); break;
        case 19: launch_all<19>(grid, sha, s, A); break;
        case 20: launch_all<20>(grid, sha, s, A); break;
        case 21: launch_all<21>(grid, sha, s, A); break;
        case 22: launch_all<22>(grid, sha, s, A); break;
        case 23: launch_all<23>(grid, sha, s, A); break;
        case 32: launch_all<32>(grid, sha, s, A); break;
        case 33: launch_all<33>(grid, sha, s, A); break;
        case 34: launch_all<34>(grid, sha, s, A); break;
        case 35: launch_all<35>(grid, sha, s, A); break;
        case 36: launch_all<36>(grid, sha, s, A); break;
        case 37: launch_all<37>(grid, sha, s, A); break;
        case 38: launch_all<38>(grid, sha, s, A); break;
        case 39: launch_all<39>(grid, sha, s, A); break;
        case 49: launch_all<49>(grid, sha, s, A); break;
        case 50: launch_all<50>(grid, sha, s, A); break;
        case 51: launch_all<51>(grid, sha, s, A); break;
        case 52: launch_all<52>(grid, sha, s, A); break;
        case 53: launch_all<53>(grid, sha, s, A); break;
        case 54: launch_all<54>(grid, sha, s, A); break;
        case 55: launch_all<55>(grid, sha, s, A); break;
        default: return MCI_E_INVALID_ARGUMENT;
        }
        return hipGetLastError() == hipSuccess ? MCI_OK : MCI_E_DEVICE;
    }
    const size_t sh = lds_bytes(prec, c.ndim, (fl & F_X2) != 0 && !(fl & F_FULL));
    if (A.ibits >= 0) { fl |= F_P2; }
    if (mt) {
        if (prec != MCI_PREC_FP64 || (c.nvecs & (c.nvecs - 1)) != 0 || !c.mt || !c.mtp) { return MCI_E_INVALID_ARGUMENT; }
        fl |= F_MT;
    }
    switch (fl) {
    case 0: launch_fl<0>(prec, grid, sh, s, A); break;
    case 1: launch_fl<1>(prec, grid, sh, s, A); break;
    case 2: launch_fl<2>(prec, grid, sh, s, A); break;
    case 3: launch_fl<3>(prec, grid, sh, s, A); break;
    case 4: launch_fl<4>(prec, grid, sh, s, A); break;
    case 5: launch_fl<5>(prec, grid, sh, s, A); break;
    case 6: launch_fl<6>(prec, grid, sh, s, A); break;
    case 7: launch_fl<7>(prec, grid, sh, s, A); break;
    case 8: launch_fl<8>(prec, grid, sh, s, A); break;
    case 9: launch_fl<9>(prec, grid, sh, s, A); break;
    case 10: launch_fl<10>(prec, grid, sh, s, A); break;
    case 11: launch_fl<11>(prec, grid, sh, s, A); break;
    case 12: launch_fl<12>(prec, grid, sh, s, A); break;
    case 13: launch_fl<13>(prec, grid, sh, s, A); break;
    case 14: launch_fl<14>(prec, grid, sh, s, A); break;
    case 15: launch_fl<15>(prec, grid, sh, s, A); break;
    // Full accumulators (FP64 only)
    case 17: launch_fl<17>(prec, grid, sh, s, A); break;
    case 18: launch_fl<18>(prec, grid, sh, s, A); break;
    case 19: launch_fl<19>(prec, grid, sh, s, A); break;
    case 20: launch_fl<20>(prec, grid, sh, s, A); break;
    case 21: launch_fl<21>(prec, grid, sh, s, A); break;
    case 22: launch_fl<22>(prec, grid, sh, s, A); break;
    case 23: launch_fl<23>(prec, grid, sh, s, A); break;
    case 25: launch_fl<25>(prec, grid, sh, s, A); break;
    case 26: launch_fl<26>(prec, grid, sh, s, A); break;
    case 27: launch_fl<27>(prec, grid, sh, s, A); break;
    case 28: launch_fl<28>(prec, grid, sh, s, A); break;
    case 29: launch_fl<29>(prec, grid, sh, s, A); break;
    case 30: launch_fl<30>(prec, grid, sh, s, A); break;
    case 31: launch_fl<31>(prec, grid, sh, s, A); break;
    // reference streams (FP64; the index is always a shift)
    case 40: launch_fl<40>(prec, grid, sh, s, A); break;
    case 41: launch_fl<41>(prec, grid, sh, s, A); break;
    case 42: launch_fl<42>(prec, grid, sh, s, A); break;
    case 43: launch_fl<43>(prec, grid, sh, s, A); break;
    case 44: launch_fl<44>(prec, grid, sh, s, A); break;
    case 45: launch_fl<45>(prec, grid, sh, s, A); break;
    case 46: launch_fl<46>(prec, grid, sh, s, A); break;
    case 47: launch_fl<47>(prec, grid, sh, s, A); break;
    case 57: launch_fl<57>(prec, grid, sh, s, A); break;
    case 58: launch_fl<58>(prec, grid, sh, s, A); break;
    case 59: launch_fl<59>(prec, grid, sh, s, A); break;
    case 60: launch_fl<60>(prec, grid, sh, s, A); break;
    case 61: launch_fl<61>(prec, grid, sh, s, A); break;
    case 62: launch_fl<62>(prec, grid, sh, s, A); break;
    case 63: launch_fl<63>(prec, grid, sh, s, A); break;
    default: return MCI_E_INVALID_ARGUMENT;
    }
    return hipGetLastError() == hipSuccess ? MCI_OK : MCI_E_DEVICE;
}

void fast_release(FastState &fs) { (void)fs; }
} // namespace mcid
