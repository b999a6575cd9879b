// mci_kernels.hpp — launch interface between the host orchestration (mci_api.cpp) and the
// HIP kernels (mci_kernels.hip, mci_fast.hip).
#pragma once

#include "mci_device.hpp"

namespace mcid
{
// series up to this length run MJBlocker in LDS (64 x MJ_LDS_MAX doubles per workgroup)
constexpr int64_t MJ_LDS_MAX = 64;

// device-side state of the shared (rank-averaged) step calibration loop
struct CalibCtrl
{
    double rate, fact;
    int cons, counter, done, apply;
    unsigned arrive;  // k_calib_round's arrival counter (zeroed with the rest at each calibration)
};

// one estimator pass over one observable's accumulator data [n][nobs][W]
// rows of the general MJBlocker's working copy for a series of n values: the blocked levels,
// n/2 values at most (the data itself is level 0)
inline int64_t mj_scratch_rows(int64_t n) { return n / 2 > 0 ? n / 2 : 1; }

struct EstimJob
{
    const double *data;
    double *scratch;      // [mj_scratch_rows(n)] rows of the data's layout (MJBlocker's in-place blocking)
    int64_t n, W;
    int nobs, off, estim;
    double scale;         // volume factor (no sampling function), else 1
    const uint8_t *active;
    double *avg, *err;    // [nobsdim][W]
    // or nullptr: the data's columns hold a compacted set of walkers -- column i < *wcount belongs to
    // walker wmap[i], whose results go to column wmap[i]; the other columns are not read
    const int32_t *wmap = nullptr, *wcount = nullptr;
};

#ifndef MCI_RTC_USER // the host's launch interface (not part of the run-time build of the step kernels)
hipError_t launch_seed_mt(uint64_t *mt, int32_t *mtp, const uint64_t *seeds, int64_t W, hipStream_t s);
// the walkers with a set flag listed (count starts at 0); their sampling columns moved to / from
// compacted arrays (dense: the chain with the compacted arrays)
hipError_t launch_compact_active(const uint8_t *active, int64_t W, int32_t *list, int32_t *count, hipStream_t s);
hipError_t launch_walker_columns(const Chain &c, const Chain &dense, const int32_t *list, const int32_t *count, int to_dense,
                                 hipStream_t s);
// every walker's mt19937_64 state rotated to next index 0 (the same generator)
hipError_t launch_mt_align(uint64_t *mt, int32_t *mtp, int64_t W, hipStream_t s);
hipError_t launch_sample(const Chain &c, int64_t nsteps, int acc, const uint8_t *active, hipStream_t s,
                         const CalibCtrl *skip = nullptr);
hipError_t launch_calibrate_indep(const Chain &c, int nfind, double target, int64_t min_stat, const double *dimSizes, hipStream_t s);
hipError_t launch_scale_steps(const Chain &c, const CalibCtrl *ctrl, const double *dimSizes, hipStream_t s);
hipError_t launch_calib_control(const double *chunks, int64_t W, double target, int nfind, CalibCtrl *ctrl, hipStream_t s);
// rates + chunk sums + control of one calibration iteration in one launch
hipError_t launch_calib_round(const int64_t *acc, const int64_t *rej, int64_t W, double *chunks, double target, int nfind,
                              CalibCtrl *ctrl, hipStream_t s);
hipError_t launch_walker_rates(const int64_t *acc, const int64_t *rej, double *rate, int64_t W, hipStream_t s);
// one independent calibration iteration's bookkeeping after its sampling launch (active walkers)
hipError_t launch_calib_indep_round(const Chain &c, uint8_t *active, int32_t *cons, int32_t *count, int nfind, double target,
                                    const double *dimSizes, int32_t *any, hipStream_t s);
hipError_t launch_estimate(const EstimJob &j, hipStream_t s);
hipError_t launch_chunk_sum(const double *v, int64_t W, int ncomp, int square, double *chunks, hipStream_t s);
hipError_t launch_combine(const double *ca, const double *ce, int64_t W, int ncomp, double *avg, double *err, hipStream_t s);
hipError_t launch_mean_of_chunks(const double *chunks, int64_t W, double *out, hipStream_t s);
hipError_t launch_decorr_compare(double *oa, double *oe, const double *na, const double *ne, int nd, int64_t W, uint8_t *active,
                                 hipStream_t s);
hipError_t launch_move_x(const Chain &c, hipStream_t s);
hipError_t launch_new_random_x(const Chain &c, hipStream_t s);
hipError_t launch_apply_domain(const Chain &c, hipStream_t s);
#endif
} // namespace mcid
