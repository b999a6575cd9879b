// mci_fast.hpp — specialised sampling kernels for the hot configuration family:
// single-coordinate uniform moves on one diagonal Gaussian (MCI_PDF_GAUSS_DIAG), unbounded
// domain, observables from {X2, X2_SUM, EXP_X2} with Block/Simple accumulators.
// Walker coordinates live in LDS, block sums are accumulated by dwell time, and the
// FP64 acceptance test is pre-screened with the FP32 exponential.
#pragma once

#include "mci_device.hpp"

namespace mcid
{
struct FastState
{
    float *x32 = nullptr;   // FP32 walker coordinates [ndim][W]
    size_t nx32 = 0;
};

bool fast_eligible(int prec, int rng, int movetype, int veclen, int ntypes, int periodic, int npdf, int pdfkind, int nvecs);
// the observables of the main sampling fit the specialised kernel (and its LDS)
bool fast_obs_eligible(const ObsDev *obs, int nobs, int ndim, int prec, int movetype, int rng);
struct CalibCtrl;
// nsteps Metropolis steps; obs == nullptr: no accumulation (decorrelation).  pend (FP64 only):
// the shared calibration's control state; the launch first applies its pending step factor
// (pend->apply, clamped by dimsizes as k_scale_steps does) and then does nothing more if
// pend->done.  active (not with pend): per-walker flags, walkers whose flag is 0 do nothing, as in
// the generic kernel.  wmap / wcount (device): the chain's per-walker columns hold a compacted set of
// walkers -- lane i < *wcount is walker wmap[i] (its Philox stream); the lanes past *wcount do nothing.
int fast_sample(FastState &fs, const Chain &c, int prec, int64_t nsteps, const ObsDev *obs, int nobs, hipStream_t s,
                const CalibCtrl *pend = nullptr, const double *dimsizes = nullptr, const uint8_t *active = nullptr,
                const int32_t *wmap = nullptr, const int32_t *wcount = nullptr);
void fast_release(FastState &fs);
} // namespace mcid
