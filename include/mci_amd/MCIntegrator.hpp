// MCIntegrator.hpp — C++ facade with the interface of the reference's mci::MCI
// (include/mci/MCIntegrator.hpp:26-226), running on the MI355X through the C-ABI of
// include/mci_amd.h.  Header-only; link against mcintegratorplusplus_amd/lib/libmci_amd.so.
//
// Differences a user of the reference sees:
//  - Sampling functions and observables are device functors: instead of deriving from
//    SamplingFunctionInterface / ObservableFunctionInterface, pass a descriptor from the
//    catalogue below (kind + parameters).  The catalogue's expressions are those of the
//    reference's test fixtures (test/common/TestMCIFunctions.hpp) and of the benchmark targets.
//  - A context may hold many walkers (nwalkers > 1): independent MCI chains on one GPU, or MPI-
//    like coupled ranks (setWalkerCoupling).  integrate() returns the walker-combined result as
//    MPIMCI::integrate does (src/MPIMCI.cpp:85-92); with one walker it is MCI::integrate.
//  - Exceptions: std::invalid_argument, std::domain_error and std::runtime_error as in the
//    reference; HIP failures raise std::runtime_error.
#ifndef MCI_AMD_MCINTEGRATOR_HPP
#define MCI_AMD_MCINTEGRATOR_HPP

#include "../mci_amd.h"

#include <cstdint>
#include <stdexcept>
#include <exception>
#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace mci
{
// --- enums of include/mci/Factories.hpp (same order and names) ---
enum class EstimatorType { Noop, Uncorrelated, Correlated, FCBlocker, MJBlocker };
enum class MoveType { All, Vec, MultiStep };
enum class SRRDType { Uniform, Gaussian, Student, Cauchy, Exponential, Gamma, Weibull, Lognormal, Chisq, Fisher };
enum class RandomMode { MT19937_64 = MCI_RNG_MT19937_64, Philox = MCI_RNG_PHILOX };
enum class Precision { FP64 = MCI_PREC_FP64, FP32 = MCI_PREC_FP32 };
enum class WalkerCoupling { Independent = MCI_COUPLING_INDEPENDENT, Shared = MCI_COUPLING_SHARED };

static constexpr double DEFAULT_MRT2STEP = 0.05; // Factories.hpp:145

inline void check(mci_status st)
{
    switch (st) {
    case MCI_OK: return;
    case MCI_E_INVALID_ARGUMENT: throw std::invalid_argument(mci_last_error());
    case MCI_E_DOMAIN: throw std::domain_error(mci_last_error());
    default: throw std::runtime_error(mci_last_error());
    }
}

// selectEstimatorType (Factories.hpp:61-71)
inline EstimatorType selectEstimatorType(bool flag_correlated, bool flag_error = true)
{
    if (flag_correlated) {
        if (!flag_error) {
            throw std::invalid_argument("[selectEstimatorType] Error calculation is set off, but correlated error estimation is set on.");
        }
        return EstimatorType::Correlated;
    }
    return flag_error ? EstimatorType::Uncorrelated : EstimatorType::Noop;
}

// --- device functor descriptors ---
struct SamplingFunction
{
    int kind;
    int ndim;
    std::vector<double> params;
    std::string name{}, source{}; // MCI_PDF_DEVICE: the functor's name and HIP source
    int nproto = 0;               // MCI_PDF_DEVICE: its proto values
    int getNDim() const { return ndim; }
};

// a user subclass of SamplingFunctionInterface (SamplingFunctionInterface.hpp:74-97) as device code:
// `source` defines NAME_proto, NAME_acceptance and NAME_sampling (mci_amd/device_functor.hpp), with
// nproto proto values and the given parameters
inline SamplingFunction DeviceSamplingFunction(const std::string &name, const std::string &source, int ndim, int nproto,
                                               const std::vector<double> &params = {})
{
    SamplingFunction f{MCI_PDF_DEVICE, ndim, params};
    f.name = name;
    f.source = source;
    f.nproto = nproto;
    return f;
}

// MultiStepMove (MultiStepMove.hpp): nsteps inner Metropolis steps of a contained SRRD move under
// the move's own sampling functions, then one outer acceptance; handed to MCI::setTrialMove
struct MultiStepMove
{
    int ndim, nsteps;
    SRRDType srrd = SRRDType::Uniform; // contained move: default UniformVecMove(ndim, 1, 0.05)
    int veclen = 1, ntypes = 1;
    std::vector<int> typeEnds;
    std::vector<SamplingFunction> pdfs;
    MultiStepMove(int nd, int ns) : ndim(nd), nsteps(ns) {}
    explicit MultiStepMove(int nd) : MultiStepMove(nd, nd) {}
    void setNSteps(int n) { nsteps = n; }
    int getNSteps() const { return nsteps; }
    void setTrialMove(SRRDType s, int vlen = 0, int nt = 1, const int te[] = nullptr)
    {
        srrd = s;
        veclen = vlen;
        ntypes = nt;
        typeEnds.assign(te, te ? te + nt : te);
    }
    void addSamplingFunction(const SamplingFunction &pdf)
    {
        if (pdf.getNDim() != ndim) {
            throw std::invalid_argument("[MultiStepMove::addSamplingFunction] Passed sampling function's number of inputs is not equal to number of walkers.");
        }
        pdfs.push_back(pdf);
    }
    void clearSamplingFunctions() { pdfs.clear(); }
};

// A mixed trial move (no reference counterpart; it plugs into the non-symmetric trial-move hook,
// TrialMoveInterface.hpp:63-69): with probability pmix an independence draw from a diagonal Gaussian
// mixture with factor q(x)/q(x'), e.g. the trained mixture of a learned proposal
// (mci_lp_get_mixture), else the contained SRRD move; handed to MCI::setTrialMove
struct LearnedMixtureMove
{
    int ndim;
    double pmix;
    std::vector<double> weights, means, sigmas; // [K], [K][ndim], [K][ndim]
    SRRDType srrd = SRRDType::Uniform;          // contained move, as MCI::setTrialMove(SRRDType, ...)
    int veclen = 0, ntypes = 1;
    std::vector<int> typeEnds;
    LearnedMixtureMove(int nd, double p) : ndim(nd), pmix(p) {}
    void addComponent(double w, const std::vector<double> &mu, const std::vector<double> &sigma)
    {
        if ((int)mu.size() != ndim || (int)sigma.size() != ndim) { throw std::invalid_argument("[LearnedMixtureMove] component sizes must equal ndim"); }
        weights.push_back(w);
        means.insert(means.end(), mu.begin(), mu.end());
        sigmas.insert(sigmas.end(), sigma.begin(), sigma.end());
    }
    void setTrialMove(SRRDType s, int vlen = 0, int nt = 1, const int te[] = nullptr)
    {
        srrd = s;
        veclen = vlen;
        ntypes = nt;
        typeEnds.assign(te, te ? te + nt : te);
    }
};

struct ObservableFunction
{
    int kind;
    int ndim;
    std::vector<double> params;
    int nobs_dep = 0; // MCI_OBS_DEP_SQUARE: the observable it squares has this many components
                      // MCI_OBS_DEVICE: the functor's number of values
    std::string name{}, source{}; // MCI_OBS_DEVICE: the functor's name and HIP source
    int getNDim() const { return ndim; }
    int getNObs() const
    {
        if (kind == MCI_OBS_DEP_SQUARE || kind == MCI_OBS_DEVICE) { return nobs_dep; }
        return (kind == MCI_OBS_X2 || kind == MCI_OBS_X || kind == MCI_OBS_PDF_PROTO) ? ndim : 1;
    }
};

// a user subclass of ObservableFunctionInterface (ObservableFunctionInterface.hpp:48-61) as device
// code: `source` defines NAME_observable (mci_amd/device_functor.hpp) with nobs values
inline ObservableFunction DeviceObservable(const std::string &name, const std::string &source, int ndim, int nobs,
                                           const std::vector<double> &params = {})
{
    ObservableFunction o{MCI_OBS_DEVICE, ndim, params, nobs};
    o.name = name;
    o.source = source;
    return o;
}

// a user subclass of TrialMoveInterface (TrialMoveInterface.hpp:16-70) as device code: `source`
// defines NAME_move and, with proto values, NAME_proto (mci_amd/device_functor.hpp); its step sizes
// (initial values) over typeEnds as for SRRD moves, and its change rate; handed to MCI::setTrialMove
struct DeviceTrialMove
{
    std::string name, source;
    int nproto = 0;
    std::vector<double> steps{0.05};
    std::vector<int> typeEnds{};
    double changeRate = 1.;
    std::vector<double> params{};
};

// sampling functions (TestMCIFunctions.hpp:123-256 and the benchmark targets)
inline SamplingFunction ThreeDimGaussianPDF() { return {MCI_PDF_GAUSS_ISO, 3, {}}; }       // exp(-|x|^2), d = 3
inline SamplingFunction GaussianPDF(int nd) { return {MCI_PDF_GAUSS_ISO, nd, {}}; }         // exp(-|x|^2), one proto value
inline SamplingFunction Gauss(int nd) { return {MCI_PDF_GAUSS_DIAG, nd, {}}; }              // exp(-sum x_i^2), per-coordinate
inline SamplingFunction AnisotropicGaussPDF(const std::vector<double> &sigma)               // exp(-sum x_i^2 / (2 sigma_i^2))
{
    SamplingFunction f{MCI_PDF_GAUSS_DIAG, (int)sigma.size(), {}};
    for (double s : sigma) { f.params.push_back(0.5 / (s * s)); }
    return f;
}
inline SamplingFunction ExpNDPDF(int nd) { return {MCI_PDF_EXP_DIAG, nd, {}}; }            // exp(-sum |x_i|)
inline SamplingFunction Exp1DPDF() { return {MCI_PDF_EXP_DIAG, 1, {}}; }
// isotropic Gaussian mixture: weights w, widths s, means mu[K][nd]
inline SamplingFunction GaussianMixturePDF(int nd, const std::vector<double> &w, const std::vector<double> &s,
                                           const std::vector<double> &mu)
{
    SamplingFunction f{MCI_PDF_GMM_ISO, nd, {(double)w.size()}};
    f.params.insert(f.params.end(), w.begin(), w.end());
    f.params.insert(f.params.end(), s.begin(), s.end());
    f.params.insert(f.params.end(), mu.begin(), mu.end());
    return f;
}

// observables (TestMCIFunctions.hpp:260-462)
inline ObservableFunction XSquared() { return {MCI_OBS_X2_SUM, 3, {3.0}}; }              // |x|^2/3, d = 3
inline ObservableFunction XYZSquared() { return {MCI_OBS_X2, 3, {}}; }                   // x_i^2, d = 3
inline ObservableFunction X2(int nd) { return {MCI_OBS_X2, nd, {}}; }                    // x_i^2
inline ObservableFunction X2Sum(int nd) { return {MCI_OBS_X2_SUM, nd, {}}; }             // |x|^2
inline ObservableFunction XND(int nd) { return {MCI_OBS_X, nd, {}}; }                    // x_i
inline ObservableFunction X1D() { return {MCI_OBS_X, 1, {}}; }
inline ObservableFunction Polynom(int nd) { return {MCI_OBS_X_SUM, nd, {}}; }            // sum x_i
inline ObservableFunction Constval(int nd, double c = 1.3) { return {MCI_OBS_CONST, nd, {c}}; }
inline ObservableFunction ExpX2(int nd, double c) { return {MCI_OBS_EXP_X2, nd, {c}}; }  // exp(-|x|^2/c)
inline ObservableFunction Indicator(int nd, int j) { return {MCI_OBS_INDICATOR, nd, {(double)j}}; } // 1[x_j > 0]
inline ObservableFunction GaussXSquared() { return {MCI_OBS_GAUSS_X2, 3, {}}; }        // exp(-|x|^2)|x|^2/(3 pi^1.5)
// dependent observables (DependentObservableInterface.hpp): what sampling function j computed at the
// accepted state, or the squares of the values observable k (added earlier, nskip dividing this one's)
// last computed
inline ObservableFunction PdfLogValue(int nd, int j = 0) { return {MCI_OBS_PDF_LOG, nd, {(double)j}}; }     // log of pdf j
inline ObservableFunction PdfProtoValues(int nd, int j = 0) { return {MCI_OBS_PDF_PROTO, nd, {(double)j}}; } // its proto values
inline ObservableFunction SquaredObservable(const ObservableFunction &dep, int k)
{
    return {MCI_OBS_DEP_SQUARE, dep.ndim, {(double)k}, dep.getNObs()};
}

// estimator functions on host data, run on the device (Estimators.hpp)
inline void estimate(EstimatorType est, int64_t n, int ndim, const double x[], double average[], double error[])
{
    check(mci_estimate(static_cast<int>(est), n, ndim, x, average, error));
}
inline void UncorrelatedEstimator(int64_t n, int ndim, const double x[], double avg[], double err[]) { estimate(EstimatorType::Uncorrelated, n, ndim, x, avg, err); }
inline void CorrelatedEstimator(int64_t n, int ndim, const double x[], double avg[], double err[]) { estimate(EstimatorType::Correlated, n, ndim, x, avg, err); }
inline void FCBlockerEstimator(int64_t n, int ndim, const double x[], double avg[], double err[]) { estimate(EstimatorType::FCBlocker, n, ndim, x, avg, err); }
inline void MJBlockerEstimator(int64_t n, int ndim, const double x[], double avg[], double err[]) { estimate(EstimatorType::MJBlocker, n, ndim, x, avg, err); }
inline void NoopEstimator(int64_t n, int ndim, const double x[], double avg[], double err[]) { estimate(EstimatorType::Noop, n, ndim, x, avg, err); }

class MCI
{
    mci_ctx *_ctx = nullptr;
    int _ndim;
    int64_t _nwalkers;
    std::vector<SamplingFunction> _pdfs;
    std::vector<ObservableFunction> _obs;
    std::function<void(const MCI &)> _cback;   // MCI::setCallback
    const mci_callback_view *_view = nullptr;  // the recorded state while a callback call runs
    std::exception_ptr _cback_error;

    static void trampoline(void *user, const mci_callback_view *v)
    {
        MCI *m = static_cast<MCI *>(user);
        if (m->_cback_error) { return; }
        m->_view = v;
        try {
            m->_cback(*m);
        }
        catch (...) {
            m->_cback_error = std::current_exception();
        }
        m->_view = nullptr;
    }

public:
    explicit MCI(int ndim, int64_t nwalkers = 1): _ndim(ndim), _nwalkers(nwalkers) { check(mci_create(ndim, nwalkers, &_ctx)); }
    ~MCI() { mci_destroy(_ctx); }
    MCI(const MCI &) = delete;
    MCI &operator=(const MCI &) = delete;

    // --- setters (MCIntegrator.hpp:93-123) ---
    void setSeed(uint_fast64_t seed) { check(mci_set_seed(_ctx, seed)); }
    // checkpoint and resume: every walker's position, step sizes, random stream and saved normals
    std::vector<char> getWalkerState() const
    {
        int64_t n = 0;
        check(mci_get_state_size(_ctx, &n));
        std::vector<char> buf((size_t)n);
        check(mci_get_state(_ctx, buf.data(), n));
        return buf;
    }
    void setWalkerState(const std::vector<char> &state) { check(mci_set_state(_ctx, state.data(), (int64_t)state.size())); }
    void setWalkerSeeds(const std::vector<uint64_t> &seeds) { check(mci_set_walker_seeds(_ctx, seeds.data(), (int64_t)seeds.size())); }
    void setRandomMode(RandomMode m) { check(mci_set_rng(_ctx, static_cast<int>(m))); }
    void setPrecision(Precision p) { check(mci_set_precision(_ctx, static_cast<int>(p))); }
    void setWalkerCoupling(WalkerCoupling c) { check(mci_set_coupling(_ctx, static_cast<int>(c))); }
    void setX(int i, double val)
    {
        std::vector<double> x = getXv(0);
        x[i] = val;
        setX(x.data());
    }
    void setX(const double x[]) { check(mci_set_x(_ctx, x, _ndim)); }
    void setWalkerX(const std::vector<double> &x) { check(mci_set_x(_ctx, x.data(), (int64_t)x.size())); }
    void moveX() { check(mci_move_x(_ctx)); }
    void newRandomX() { check(mci_new_random_x(_ctx)); }
    void centerX() { check(mci_center_x(_ctx)); }
    void setMRT2Step(double mrt2step) { check(mci_set_mrt2step(_ctx, -1, mrt2step)); }
    void setMRT2Step(int i, double mrt2step) { check(mci_set_mrt2step(_ctx, i, mrt2step)); }
    void setMRT2Step(const double mrt2step[])
    {
        int n = 0;
        check(mci_get_n_step_sizes(_ctx, &n));
        check(mci_set_mrt2steps(_ctx, mrt2step, n));
    }
    void setTargetAcceptanceRate(double r) { check(mci_set_target_acceptance_rate(_ctx, r)); }
    void setNfindMRT2Iterations(int n) { check(mci_set_nfind_mrt2_iterations(_ctx, n)); }
    void setNdecorrelationSteps(int64_t n) { check(mci_set_ndecorrelation_steps(_ctx, n)); }

    // --- domain (MCIntegrator.hpp:131-138) ---
    void resetDomain() { check(mci_reset_domain(_ctx)); }
    void setIRange(double lbound, double ubound)
    {
        std::vector<double> lb(_ndim, lbound), ub(_ndim, ubound);
        setIRange(lb.data(), ub.data());
    }
    void setIRange(const double lbounds[], const double ubounds[]) { check(mci_set_irange(_ctx, lbounds, ubounds)); }

    // --- trial moves (MCIntegrator.hpp:141-150) ---
    void setTrialMove(MoveType move) { check(mci_set_trial_move_type(_ctx, static_cast<int>(move))); }
    void setTrialMove(SRRDType srrd, int veclen = 0, int ntypes = 1, int typeEnds[] = nullptr)
    {
        check(mci_set_trial_move(_ctx, static_cast<int>(srrd), veclen, ntypes, typeEnds));
    }
    // --- file output (MCIntegrator.hpp:192-196) ---
    void storeObservablesOnFile(const std::string &filepath, int freq) { check(mci_store_observables_on_file(_ctx, filepath.c_str(), freq)); }
    void clearObservableFile() { check(mci_clear_observable_file(_ctx)); }
    void storeWalkerPositionsOnFile(const std::string &filepath, int freq) { check(mci_store_walker_positions_on_file(_ctx, filepath.c_str(), freq)); }
    void clearWalkerFile() { check(mci_clear_walker_file(_ctx)); }

    void setTrialMove(const MultiStepMove &m)
    {
        if (m.ndim != _ndim) { throw std::invalid_argument("[MCI::setTrialMove] Passed trial move's number of inputs is not equal to MCI's number of walkers."); }
        check(mci_set_trial_move_multistep(_ctx, m.nsteps, static_cast<int>(m.srrd), m.veclen, m.ntypes, m.typeEnds.empty() ? nullptr : m.typeEnds.data()));
        for (const auto &p : m.pdfs) { check(mci_add_multistep_sampling_function(_ctx, p.kind, p.ndim, p.params.data(), (int)p.params.size())); }
    }
    void setTrialMove(const LearnedMixtureMove &m)
    {
        if (m.ndim != _ndim) { throw std::invalid_argument("[MCI::setTrialMove] Passed trial move's number of inputs is not equal to MCI's number of walkers."); }
        check(mci_set_trial_move(_ctx, static_cast<int>(m.srrd), m.veclen, m.ntypes, m.typeEnds.empty() ? nullptr : const_cast<int *>(m.typeEnds.data())));
        check(mci_set_learned_move(_ctx, (int)m.weights.size(), m.weights.data(), m.means.data(), m.sigmas.data(), m.pmix));
    }
    void setTrialMove(const DeviceTrialMove &m)
    {
        check(mci_set_trial_move_device(_ctx, m.name.c_str(), m.source.c_str(), m.nproto, (int)m.steps.size(),
                                        m.typeEnds.empty() ? nullptr : m.typeEnds.data(), m.changeRate, m.steps.data(), m.params.data(),
                                        (int)m.params.size()));
    }

    // --- observables (MCIntegrator.hpp:154-169) ---
    void addObservable(const ObservableFunction &obs, int blocksize, int nskip, bool flag_equil, EstimatorType estimType)
    {
        if (obs.kind == MCI_OBS_DEVICE) {
            check(mci_add_observable_device(_ctx, obs.name.c_str(), obs.source.c_str(), obs.ndim, obs.nobs_dep, obs.params.data(),
                                            (int)obs.params.size(), blocksize, nskip, flag_equil ? 1 : 0, static_cast<int>(estimType)));
        }
        else {
            check(mci_add_observable(_ctx, obs.kind, obs.ndim, obs.params.data(), (int)obs.params.size(), blocksize, nskip,
                                     flag_equil ? 1 : 0, static_cast<int>(estimType)));
        }
        _obs.push_back(obs);
    }
    void addObservable(const ObservableFunction &obs, int blocksize, int nskip, bool flag_equil, bool flag_correlated)
    {
        addObservable(obs, blocksize, nskip, flag_equil, selectEstimatorType(flag_correlated, blocksize > 0));
    }
    void addObservable(const ObservableFunction &obs, int blocksize = 1, int nskip = 1)
    {
        addObservable(obs, blocksize, nskip, blocksize > 0, blocksize == 1);
    }
    ObservableFunction popObservable()
    {
        check(mci_pop_observable(_ctx));
        ObservableFunction o = _obs.back();
        _obs.pop_back();
        return o;
    }
    void clearObservables()
    {
        check(mci_clear_observables(_ctx));
        _obs.clear();
    }

    // --- sampling functions (MCIntegrator.hpp:172-175) ---
    void addSamplingFunction(const SamplingFunction &pdf)
    {
        if (pdf.kind == MCI_PDF_DEVICE) {
            check(mci_add_sampling_function_device(_ctx, pdf.name.c_str(), pdf.source.c_str(), pdf.ndim, pdf.nproto, pdf.params.data(),
                                                   (int)pdf.params.size()));
        }
        else { check(mci_add_sampling_function(_ctx, pdf.kind, pdf.ndim, pdf.params.data(), (int)pdf.params.size())); }
        _pdfs.push_back(pdf);
    }
    SamplingFunction popSamplingFunction()
    {
        check(mci_pop_sampling_function(_ctx));
        SamplingFunction f = _pdfs.back();
        _pdfs.pop_back();
        return f;
    }
    void clearSamplingFunctions()
    {
        check(mci_clear_sampling_functions(_ctx));
        _pdfs.clear();
    }

    // --- getters (MCIntegrator.hpp:199-219) ---
    int getNDim() const { return _ndim; }
    int64_t getNWalkers() const { return _nwalkers; }
    std::vector<double> getXv(int64_t walker) const
    {
        std::vector<double> all((size_t)(_ndim * _nwalkers));
        check(mci_get_x(_ctx, all.data(), (int64_t)all.size()));
        return std::vector<double>(all.begin() + walker * _ndim, all.begin() + (walker + 1) * _ndim);
    }
    double getX(int i) const { return _view ? _view->x[i] : getXv(0)[i]; }
    double getMRT2Step(int i) const
    {
        double s = 0.;
        check(mci_get_mrt2step(_ctx, i, &s));
        return s;
    }
    double getTargetAcceptanceRate() const
    {
        double t = 0.;
        check(mci_get_settings(_ctx, &t, nullptr, nullptr));
        return t;
    }
    int getNfindMRT2Iterations() const
    {
        int n = 0;
        check(mci_get_settings(_ctx, nullptr, &n, nullptr));
        return n;
    }
    int64_t getNdecorrelationSteps() const
    {
        int64_t n = 0;
        check(mci_get_settings(_ctx, nullptr, nullptr, &n));
        return n;
    }
    double getAcceptanceRate() const
    {
        if (_view) { return _view->rate; }
        int64_t a = 0, r = 0;
        double rate = 0.;
        check(mci_get_acceptance(_ctx, &a, &r, &rate));
        return rate;
    }
    double getChangeRate() const
    {
        double r = 0.;
        check(mci_get_change_rate(_ctx, &r));
        return r;
    }
    int getNPDF() const { return (int)_pdfs.size(); }
    int getNObs() const { return (int)_obs.size(); }
    int getNObsDim() const
    {
        int n = 0;
        check(mci_get_nobsdim(_ctx, &n));
        return n;
    }
    const SamplingFunction &getSamplingFunction(int i) const { return _pdfs[i]; }
    const ObservableFunction &getObservable(int i) const { return _obs[i]; }
    mci_ctx *handle() const { return _ctx; }

    // --- integrate (MCIntegrator.hpp:225) ---
    void integrate(int64_t Nmc, double average[], double error[], bool doFindMRT2step = true, bool doDecorrelation = true)
    {
        _cback_error = nullptr;
        const mci_status st = mci_integrate(_ctx, Nmc, average, error, doFindMRT2step ? 1 : 0, doDecorrelation ? 1 : 0);
        if (_cback_error) { std::rethrow_exception(_cback_error); }
        check(st);
    }

    // --- callback (MCIntegrator.hpp:189-190): called in order after the device work of each
    //     integrate, with getX and getAcceptanceRate returning what the reference's callback
    //     would see at that call (walker 0); capacity bounds the recorded calls per integrate ---
    void setCallback(const std::function<void(const MCI &)> &cback, int64_t capacity = (int64_t)1 << 20)
    {
        check(mci_set_callback(_ctx, &MCI::trampoline, this, capacity));
        _cback = cback;
    }
    void clearCallback()
    {
        check(mci_clear_callback(_ctx));
        _cback = nullptr;
    }
    // per-walker estimates of the last integrate, [walker][nobsdim]
    void getWalkerResults(std::vector<double> &average, std::vector<double> &error) const
    {
        const size_t n = (size_t)(getNObsDim() * _nwalkers);
        average.assign(n, 0.);
        error.assign(n, 0.);
        check(mci_get_walker_results(_ctx, average.data(), error.data(), (int64_t)n));
    }
};
} // namespace mci

#endif
