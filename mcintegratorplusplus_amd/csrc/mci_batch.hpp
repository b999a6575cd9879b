// mci_batch.hpp — kernels of the batched-integral path (many independent Gaussian-mixture
// integrals, FP32 walkers, FP64 reductions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mcib
{
constexpr int WAVE = 64;
constexpr int MAXK = 4;
constexpr int CALR = 64;  // calibration lanes (pooled ranks) per integral: one wave (16 lanes, four
                          // integrals per wave, measured slower: the wave count, not the lane-steps, sets the time)

// per-integral parameters, struct-of-arrays over integrals (integral index order)
struct BatchParams
{
    int nint;
    const int *ndim;          // [nint]
    const int *ncomp;         // [nint]
    const float *lw;          // [nint][4] log w_k - d/2 log(2 pi s_k^2)
    const float *h;           // [nint][4] 1 / (2 s_k^2)
    const float *sig;         // [nint][4] s_k
    const float *cw;          // [nint][4] cumulative weights (last = 1)
    const float *mu;          // concatenated [K_i][d_i] blocks
    const int64_t *mu_off;    // [nint]
    uint32_t key0, key1;      // Philox key (seed)
};

struct CalibArgs
{
    BatchParams P;
    const int *ilist;         // integrals of this launch (CALR lanes each)
    int nfind;
    double target;
    const double *step0;      // [nint] initial step
    double *step;             // [nint] calibrated step
    double *rate;             // [nint] out: last pooled acceptance rate
};

struct SampleArgs
{
    BatchParams P;
    const double *step;       // [nint]
    const int *wint;          // [nwalkers] integral of each walker
    const int64_t *wsteps;    // [nint] steps per walker T_i
    int64_t nwalkers;
    int64_t ndecorr;
    int nblocks;
    const int *items;         // 64-walker items of this launch (item q = walkers 64q .. 64q+63)
    int64_t nitems;
    unsigned int *queue;      // work-queue head (zeroed by k_batch_zero before the launch)
    double *bmeans;           // [nblocks][nwalkers]
    int32_t *acc;             // [nwalkers] accepted moves while accumulating
};

struct CombineArgs
{
    int nint;
    const int64_t *woff;      // [nint] first walker of integral i
    const int64_t *wcount;    // [nint] R_i
    const int64_t *wsteps;    // [nint] T_i
    const double *wavg, *werr;// [nwalkers] per-walker MJBlocker estimates
    const int32_t *acc;       // [nwalkers]
    int mode;                 // MCI_BATCH_ERR_*
    double *avg, *err, *accrate; // [nint]
};

size_t batch_lds_bytes(int dmax);
hipError_t launch_batch_zero(unsigned int *q, int n, hipStream_t s);
hipError_t launch_batch_calibrate(const CalibArgs &a, int nlist, int dmax, hipStream_t s);
hipError_t launch_batch_sample(const SampleArgs &a, int dmax, int nwaves, hipStream_t s);
hipError_t launch_batch_combine(const CombineArgs &a, hipStream_t s);
// resident 64-lane workgroups of k_batch_sample per CU at the given LDS size
int batch_sample_waves_per_cu(int dmax);
} // namespace mcib
