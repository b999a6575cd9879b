// mci_rtc.hpp — the caller's device functors (MCI_PDF_DEVICE, MCI_OBS_DEVICE): the generic step
// kernels (mci_step_kernels.hpp) built at run time together with the functors' sources.
#pragma once

#include "mci_kernels.hpp"

#include <string>
#include <vector>

namespace mcid
{
// a functor's name (its functions are NAME_proto, ... or NAME_observable) and source
struct RtcFunctor
{
    std::string name, source;
};

// the generic step kernels of one module (loaded on one device)
struct RtcKernels
{
    hipFunction_t sample = nullptr, calibrate = nullptr;
};

// true when name is a C identifier
bool rtc_valid_name(const std::string &name);
// the program text: the functors' sources (one copy per name), their dispatch by index (pdfs[i] is
// sampling function index i, obs[i] observable index i; the move, if any), then the step kernels.
// Returns false with a message when two functors share a name but not a source.
// move: the caller's trial move (or null)
bool rtc_program(const std::vector<RtcFunctor> &pdfs, const std::vector<RtcFunctor> &obs, const RtcFunctor *move, std::string &text,
                 std::string &msg);
// compiles text for arch (e.g. "gfx950") without loading it; log: the compiler's messages
int rtc_check(const std::string &text, const char *arch, std::string &log);
// the kernels of text on the current device: compiled and loaded once per text and device
int rtc_get(const std::string &text, RtcKernels &out, std::string &log);
hipError_t rtc_launch_sample(const RtcKernels &k, const Chain &c, int64_t nsteps, int acc, const uint8_t *active, hipStream_t s,
                             const CalibCtrl *skip);
hipError_t rtc_launch_calibrate_indep(const RtcKernels &k, const Chain &c, int nfind, double target, int64_t min_stat,
                                      const double *dimSizes, hipStream_t s);
} // namespace mcid
