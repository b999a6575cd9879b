// mci_step_kernels.hpp — the generic chain's kernels that evaluate sampling functions and
// observables: k_sample (MCI::sample: init + nsteps Metropolis steps + accumulation,
// MCIntegrator.cpp:252-321) and k_calibrate_indep (MCI::findMRT2Step, one independent loop per
// walker, MCIntegrator.cpp:87-168).  They are compiled into the library, and again, at run time,
// into a module with the caller's device functors (MCI_RTC_USER: the functor dispatch of
// mci_device.hpp is then defined ahead of this header).
#pragma once

#include "mci_kernels.hpp"

#include <float.h>

namespace mcid
{
__global__ void k_sample(Chain c, int64_t nsteps, int accumulate_obs, const uint8_t *active, const CalibCtrl *skip)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) { return; }
    if (skip && skip->done) { return; }
    if (active && !active[w]) { return; }
    Lane L;
    load_lane(c, L, w);
    init_sampling(c, L);
    uint32_t dirty = 0xffffffffu;
    if (accumulate_obs) { zero_accu(c, w); }
    for (int64_t s = 0; s < nsteps; ++s) {
        if (c.npdf > 0) { step_mrt2(c, L); }
        else { step_random(c, L); }
        if (accumulate_obs) {
            accumulate(c, L, s, dirty);
            // file output (MCIntegrator.cpp:315-316, 509-516, 533-542): the last observable values
            // and the position after steps 0, f, 2f, ...
            if (c.rec_obs != nullptr && s % c.rec_fobs == 0) {
                const int64_t r = s / c.rec_fobs;
                for (int j = 0; j < c.rec_nobsdim; ++j) { c.rec_obs[(r * c.rec_nobsdim + j) * c.W + w] = c.obsval[(int64_t)j * c.W + w]; }
            }
            if (c.rec_x != nullptr && s % c.rec_fx == 0) {
                const int64_t r = s / c.rec_fx;
                for (int i = 0; i < c.ndim; ++i) { c.rec_x[(r * c.ndim + i) * c.W + w] = c.xold[(int64_t)i * c.W + w]; }
            }
        }
    }
    if (accumulate_obs) { finalize_accu(c, w, nsteps); }
    store_lane(c, L);
}

// step clamps of findMRT2Step (MCIntegrator.cpp:147-161)
__device__ inline void scale_and_clamp(const Chain &c, int64_t w, double fact, const double *dimSizes)
{
    const int64_t W = c.W;
    for (int t = 0; t < c.ntypes; ++t) { double &s = c.steps[(int64_t)t * W + w]; s = s * fact; }
    for (int i = 0; i < c.ndim; ++i) {
        int t = 0;
        while (t < c.ntypes && !(i < c.typeEnds[t])) { ++t; }
        double &s = c.steps[(int64_t)t * W + w];
        if (s > 0.5 * dimSizes[i]) { s = 0.5 * dimSizes[i]; }
    }
    for (int t = 0; t < c.ntypes; ++t) {
        double &s = c.steps[(int64_t)t * W + w];
        if (s < (double)FLT_MIN) { s = (double)FLT_MIN; }
    }
}

__global__ void k_calibrate_indep(Chain c, int nfind, double target, int64_t min_stat, const double *dimSizes)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= c.W) { return; }
    Lane L;
    load_lane(c, L, w);
    int cons = 0, counter = 0;
    while ((nfind < 0 && cons < 5) || counter < nfind) {
        init_sampling(c, L);
        for (int64_t s = 0; s < min_stat; ++s) { step_mrt2(c, L); }
        const double rate = (L.acc > 0) ? (double)L.acc / ((double)L.acc + (double)L.rej) : 0.;
        if (fabs(rate - target) < 0.05) { ++cons; }
        else { cons = 0; }
        const double fact = fmin(2., fmax(0.5, rate / target));
        scale_and_clamp(c, w, fact, dimSizes);
        ++counter;
        if (nfind < 0 && counter >= -nfind) { break; }
    }
    store_lane(c, L);
}
} // namespace mcid
