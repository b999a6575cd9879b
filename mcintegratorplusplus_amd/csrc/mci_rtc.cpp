// mci_rtc.cpp — run-time builds of the generic step kernels with the caller's device functors.
//
// The reference's users subclass SamplingFunctionInterface / ObservableFunctionInterface with their
// own C++ (SamplingFunctionInterface.hpp:74-97, ObservableFunctionInterface.hpp:48-61).  A GPU
// library cannot call host code per walker and step, so the caller gives the functor as HIP C++
// source instead; it is compiled by hiprtc together with the library's own step kernels (their
// headers are embedded in the library at build time), so the functor runs inlined in the chain.
// Modules are cached per program text and device for the life of the process.
#include "mci_rtc.hpp"

#include <hip/hiprtc.h>

#include <map>
#include <mutex>
#include <set>
#include <utility>

namespace mcid
{
// the embedded headers (written by the build into mci_rtc_embed.cpp)
extern const int rtc_nheaders;
extern const char *const rtc_header_names[];
extern const char *const rtc_header_texts[];

namespace
{
std::mutex g_mu;
struct Loaded
{
    hipModule_t mod = nullptr;
    RtcKernels k;
};
std::map<std::pair<std::string, int>, Loaded> g_cache;

const char *kSample = "mcid::k_sample", *kCalib = "mcid::k_calibrate_indep";

// compiles text; on success code holds the code object and the kernels' lowered names
int compile(const std::string &text, const std::string &arch, std::vector<char> &code, std::string &nsample, std::string &ncalib,
            std::string &log)
{
    hiprtcProgram p;
    if (hiprtcCreateProgram(&p, text.c_str(), "mci_device_functors.hip", rtc_nheaders, rtc_header_texts, rtc_header_names) !=
        HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return MCI_E_DEVICE;
    }
    hiprtcAddNameExpression(p, kSample);
    hiprtcAddNameExpression(p, kCalib);
    // the library's own flags: no contraction (bit-exact FP64 chains), correctly rounded FP32
    // divide and sqrt, no packed FP32 instructions
    const std::string off = "--offload-arch=" + arch;
    const char *opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-DMCI_RTC_USER=1"};
    const hiprtcResult r = hiprtcCompileProgram(p, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t ls = 0;
    if (hiprtcGetProgramLogSize(p, &ls) == HIPRTC_SUCCESS && ls > 1) {
        log.assign(ls, '\0');
        hiprtcGetProgramLog(p, &log[0]);
        log.resize(ls - 1);
    }
    if (r != HIPRTC_SUCCESS) {
        if (log.empty()) { log = hiprtcGetErrorString(r); }
        hiprtcDestroyProgram(&p);
        return MCI_E_INVALID_ARGUMENT;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(p, &cs);
    code.resize(cs);
    hiprtcGetCode(p, code.data());
    const char *ls1 = nullptr, *ls2 = nullptr;
    hiprtcGetLoweredName(p, kSample, &ls1);
    hiprtcGetLoweredName(p, kCalib, &ls2);
    nsample = ls1 ? ls1 : "";
    ncalib = ls2 ? ls2 : "";
    hiprtcDestroyProgram(&p);
    return (cs > 0 && !nsample.empty() && !ncalib.empty()) ? MCI_OK : MCI_E_DEVICE;
}

void dispatch(std::string &t, const char *sig, const char *call, const char *ret, const std::vector<RtcFunctor> &f, const char *suffix)
{
    t += "__device__ inline ";
    t += sig;
    t += "\n{\n    switch (uid) {\n";
    for (size_t i = 0; i < f.size(); ++i) {
        t += "    case " + std::to_string(i) + ": " + (ret[0] ? "return " : "") + f[i].name + suffix + call + ";" +
             (ret[0] ? "" : " break;") + "\n";
    }
    t += "    default: break;\n    }\n";
    if (ret[0]) { t += std::string("    return ") + ret + ";\n"; }
    t += "}\n";
}
} // namespace

bool rtc_valid_name(const std::string &n)
{
    if (n.empty() || n.size() > 64) { return false; }
    for (size_t i = 0; i < n.size(); ++i) {
        const char ch = n[i];
        const bool alpha = (ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || ch == '_';
        if (!(alpha || (i > 0 && ch >= '0' && ch <= '9'))) { return false; }
    }
    return true;
}

bool rtc_program(const std::vector<RtcFunctor> &pdfs, const std::vector<RtcFunctor> &obs, const RtcFunctor *move, std::string &text,
                 std::string &msg)
{
    std::map<std::string, const std::string *> src;
    const std::vector<RtcFunctor> mv = move ? std::vector<RtcFunctor>{*move} : std::vector<RtcFunctor>{};
    for (const auto *v : {&pdfs, &obs, &mv}) {
        for (const RtcFunctor &f : *v) {
            auto it = src.find(f.name);
            if (it != src.end() && *it->second != f.source) {
                msg = "two device functors are named " + f.name + " with different sources";
                return false;
            }
            src[f.name] = &f.source;
        }
    }
    text = "#include \"../../include/mci_amd/device_functor.hpp\"\n";
    std::set<std::string> done;
    for (const auto *v : {&pdfs, &obs, &mv}) {
        for (const RtcFunctor &f : *v) {
            if (!done.insert(f.name).second) { continue; }
            text += "#line 1 \"" + f.name + "\"\n" + f.source + "\n";
        }
    }
    text += "#line 1 \"mci_device_functor_dispatch\"\n";
    dispatch(text, "void mci_user_pdf_proto(int uid, int ndim, MciVec x, const double *par, MciVec proto)", "(ndim, x, par, proto)", "", pdfs,
             "_proto");
    dispatch(text, "double mci_user_pdf_acceptance(int uid, int nproto, MciVec po, MciVec pn, const double *par)", "(nproto, po, pn, par)",
             "0.0", pdfs, "_acceptance");
    dispatch(text, "double mci_user_pdf_sampling(int uid, int nproto, MciVec pv, const double *par)", "(nproto, pv, par)", "0.0", pdfs,
             "_sampling");
    dispatch(text, "void mci_user_observable(int uid, int ndim, MciVec x, const double *par, MciVec out)", "(ndim, x, par, out)", "", obs,
             "_observable");
    // updatedAcceptance: the caller's overload when the source defines one (a non-template function
    // is preferred to this template of the same signature), else the interface's default
    std::set<std::string> upd;
    for (const RtcFunctor &f : pdfs) {
        if (!upd.insert(f.name).second) { continue; }
        text += "template <int = 0>\n__device__ inline double " + f.name +
                "_updated_acceptance(int ndim, int nproto, MciVec xnew, MciChanged, MciVec po, MciVec pn, const double *par)\n{\n    " +
                f.name + "_proto(ndim, xnew, par, pn);\n    return " + f.name + "_acceptance(nproto, po, pn, par);\n}\n";
    }
    dispatch(text, "double mci_user_pdf_updated(int uid, int ndim, int nproto, MciVec x, MciChanged ch, MciVec po, MciVec pn, const double *par)",
             "(ndim, nproto, x, ch, po, pn, par)", "0.0", pdfs, "_updated_acceptance");
    // the trial move (trialMove, and protoFunction when it has proto values: without one of its own a
    // template that computes nothing is taken)
    text += "__device__ inline double mci_user_move(int ndim, MciVec x, MciMoved &moved, MciVec po, MciVec pn, MciVec steps, MciRng &rng, "
            "const double *par)\n{\n";
    text += move ? "    return " + move->name + "_move(ndim, x, moved, po, pn, steps, rng, par);\n}\n" : "    return 1.0;\n}\n";
    if (move && upd.count(move->name) == 0) {
        text += "template <int = 0>\n__device__ inline void " + move->name + "_proto(int, MciVec, const double *, MciVec) {}\n";
    }
    text += "__device__ inline void mci_user_move_proto(int ndim, MciVec x, const double *par, MciVec proto)\n{\n";
    text += move ? "    " + move->name + "_proto(ndim, x, par, proto);\n}\n" : "}\n";
    text += "#include \"mci_step_kernels.hpp\"\n";
    return true;
}

int rtc_check(const std::string &text, const char *arch, std::string &log)
{
    std::vector<char> code;
    std::string a, b;
    return compile(text, arch ? arch : "gfx950", code, a, b, log);
}

int rtc_get(const std::string &text, RtcKernels &out, std::string &log)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { return MCI_E_DEVICE; }
    std::lock_guard<std::mutex> lk(g_mu);
    auto key = std::make_pair(text, dev);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
        out = it->second.k;
        return MCI_OK;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) { return MCI_E_DEVICE; }
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) { arch.resize(colon); }
    std::vector<char> code;
    std::string ns, nc;
    const int st = compile(text, arch, code, ns, nc, log);
    if (st != MCI_OK) { return st; }
    Loaded L;
    if (hipModuleLoadData(&L.mod, code.data()) != hipSuccess) {
        log = "hipModuleLoadData failed";
        return MCI_E_DEVICE;
    }
    if (hipModuleGetFunction(&L.k.sample, L.mod, ns.c_str()) != hipSuccess ||
        hipModuleGetFunction(&L.k.calibrate, L.mod, nc.c_str()) != hipSuccess) {
        log = "hipModuleGetFunction failed";
        return MCI_E_DEVICE;
    }
    g_cache.emplace(key, L);
    out = L.k;
    return MCI_OK;
}

hipError_t rtc_launch_sample(const RtcKernels &k, const Chain &c, int64_t nsteps, int acc, const uint8_t *active, hipStream_t s,
                             const CalibCtrl *skip)
{
    Chain cc = c;
    void *args[] = {&cc, &nsteps, &acc, &active, &skip};
    const unsigned grid = (unsigned)((c.W + 63) / 64);
    return hipModuleLaunchKernel(k.sample, grid, 1, 1, 64, 1, 1, 0, s, args, nullptr);
}

hipError_t rtc_launch_calibrate_indep(const RtcKernels &k, const Chain &c, int nfind, double target, int64_t min_stat,
                                      const double *dimSizes, hipStream_t s)
{
    Chain cc = c;
    void *args[] = {&cc, &nfind, &target, &min_stat, &dimSizes};
    const unsigned grid = (unsigned)((c.W + 63) / 64);
    return hipModuleLaunchKernel(k.calibrate, grid, 1, 1, 64, 1, 1, 0, s, args, nullptr);
}
} // namespace mcid
