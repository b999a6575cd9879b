/*
 * mci_amd.h — C-ABI of the MI355X-native Metropolis integrator.
 *
 * This is the drop-in boundary for the Metropolis sampling path behind
 * mci::MCI::integrate (reference: include/mci/MCIntegrator.hpp:26-226,
 * src/MCIntegrator.cpp:43-649).  Every entry point below names the reference
 * interface it replaces.  The C++ facade in include/mci_amd/MCIntegrator.hpp
 * re-exposes the same calls with the reference's class, method names and
 * exception types; Python binds the same symbols through ctypes.
 *
 * Conventions
 *  - Plain pointers and sizes only; the caller owns every output array.
 *  - Every function returns an mci_status; on failure mci_last_error() holds a
 *    thread-local message.  Status codes map 1:1 onto the reference's exception
 *    types (std::invalid_argument, std::domain_error, std::runtime_error).
 *  - A context is one "MCI object" extended to W walkers.  Each walker is one
 *    Metropolis chain (one lane on the GPU).  With W == 1 a context behaves
 *    like one reference MCI; with W > 1 the walkers behave either like W
 *    independent MCI objects (MCI_COUPLING_INDEPENDENT, the USE_MPI=0 build of
 *    the reference run W times) or like W MPI ranks under MPIMCI
 *    (MCI_COUPLING_SHARED: shared step calibration and decorrelation,
 *    src/MCIntegrator.cpp:107,131-138,21-34; results combined as in
 *    src/MPIMCI.cpp:85-92).
 *  - Each context owns one HIP stream; calls on one context are serialised.
 */
#ifndef MCI_AMD_H
#define MCI_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference exception types) ---- */
typedef int mci_status;
#define MCI_OK                    0
#define MCI_E_INVALID_ARGUMENT    1  /* std::invalid_argument */
#define MCI_E_DOMAIN              2  /* std::domain_error     */
#define MCI_E_RUNTIME             3  /* std::runtime_error    */
#define MCI_E_DEVICE              4  /* HIP runtime failure (facade throws std::runtime_error) */

/* ---- enums that mirror include/mci/Factories.hpp ---- */
/* EstimatorType, Factories.hpp:52-59 (same order) */
#define MCI_ESTIM_NOOP            0
#define MCI_ESTIM_UNCORRELATED    1
#define MCI_ESTIM_CORRELATED      2  /* MJBlocker if n is a power of 2, else FCBlocker */
#define MCI_ESTIM_FCBLOCKER       3
#define MCI_ESTIM_MJBLOCKER       4

/* MoveType, Factories.hpp:108-113 */
#define MCI_MOVE_ALL              0
#define MCI_MOVE_VEC              1
#define MCI_MOVE_MULTISTEP        2  /* MultiStepMove(ndim): ndim single-coordinate uniform steps, no own sampling function */
#define MCI_MOVE_DEVICE           3  /* the caller's own trial move as device code (mci_set_trial_move_device) */

/* SRRDType, Factories.hpp:119-133 (same order), each with the default parameters of createSymRRD
 * (TrialMoveInterface.hpp:107-183): uniform on (-1, 1), normal(0, 1), Student-t(1), Cauchy(0, 1),
 * then +/- (fair sign) exponential(1), gamma(1, 1), Weibull(1, 1), lognormal(0, 1), chi-squared(1),
 * Fisher-F(1, 1).  The reference-stream mode draws them as libstdc++ (GCC 11) does. */
#define MCI_SRRD_UNIFORM          0
#define MCI_SRRD_GAUSSIAN         1
#define MCI_SRRD_STUDENT          2
#define MCI_SRRD_CAUCHY           3
#define MCI_SRRD_EXPONENTIAL      4
#define MCI_SRRD_GAMMA            5
#define MCI_SRRD_WEIBULL          6
#define MCI_SRRD_LOGNORMAL        7
#define MCI_SRRD_CHISQ            8
#define MCI_SRRD_FISHER           9

/* ---- device functor catalogue (sampling functions) ----
 * The reference's SamplingFunctionInterface subclasses become device functors
 * selected by kind + parameter blob (SamplingFunctionInterface.hpp:38-105).
 * The arithmetic of each kind is fixed below so that a reference subclass
 * written with the same expressions gives bit-identical acceptance ratios.  */
#define MCI_PDF_GAUSS_ISO   1 /* nproto 1: p = x0*x0 + x1*x1 + ... ; ratio exp(-pnew + pold)
                                 (test/common/TestMCIFunctions.hpp:123-148 ThreeDimGaussianPDF form). params: none */
#define MCI_PDF_GAUSS_DIAG  2 /* nproto ndim: p_i = (x_i*x_i)*a_i ; full ratio exp(sum pold - sum pnew),
                                 single-index ratio exp(-(sum_changed pnew_i - pold_i))
                                 (TestMCIFunctions.hpp:151-190 Gauss form with a_i = 1).
                                 params: a[ndim] (optional; default all 1) */
#define MCI_PDF_EXP_DIAG    3 /* nproto ndim: p_i = fabs(x_i)*a_i (TestMCIFunctions.hpp:219-256 ExpNDPDF form).
                                 params: a[ndim] (optional) */
#define MCI_PDF_GMM_ISO     4 /* isotropic Gaussian mixture, log-density via log-sum-exp.
                                 params: [K, w_0..w_{K-1}, s_0..s_{K-1}, mu_0[ndim] .. mu_{K-1}[ndim]] */
#define MCI_PDF_DEVICE    100 /* the caller's own sampling function as device code (mci_add_sampling_function_device) */

/* ---- device functor catalogue (observables), ObservableFunctionInterface.hpp:29-64 ---- */
#define MCI_OBS_X2          1 /* nobs ndim: out_i = x_i*x_i (X2 / XYZSquared) */
#define MCI_OBS_X2_SUM      2 /* nobs 1: out = (0 + x0*x0 + x1*x1 + ...)/c  (X2Sum, XSquared with c=3). params: [c] (default 1) */
#define MCI_OBS_X           3 /* nobs ndim: out_i = x_i (XND / X1D) */
#define MCI_OBS_X_SUM       4 /* nobs 1: out = 0 + x0 + x1 + ... (Polynom) */
#define MCI_OBS_CONST       5 /* nobs 1: out = c (Constval). params: [c] */
#define MCI_OBS_EXP_X2      6 /* nobs 1: out = exp(-((0 + x0*x0 + ...)/c)). params: [c] (default ndim) */
#define MCI_OBS_INDICATOR   7 /* nobs 1: out = x_j > 0 ? 1 : 0. params: [j] (default 0) */
#define MCI_OBS_GAUSS_X2    8 /* nobs 1: out = exp(-s)*s*c with s = x0*x0 + x1*x1 + ... (GaussXSquared,
                                 TestMCIFunctions.hpp:278-295). params: [c] (default 1./sqrt(pi*pi*pi)/3.) */
/* Dependent observables (DependentObservableInterface, include/mci/DependentObservableInterface.hpp:24-54):
 * they read what a sampling function computed or what an earlier observable last computed.  The
 * reference hands a PDF-dependent observable the sampling function's proto values of the last
 * accepted state through SamplingFunctionInterface::observationCallback (SamplingFunctionInterface.hpp:
 * 100-104), called by MCI::sample (MCIntegrator.cpp:294-305) whenever the state changed since the last
 * call and the step is observed (every gcd of the PDF-dependent observables' nskip steps,
 * ObservableContainer.cpp:16-35); at every step such an observable is evaluated it therefore sees the
 * current state's proto values, which is what the device reads directly. */
#define MCI_OBS_PDF_LOG     9 /* nobs 1: log of sampling function j's value from its proto values p of the
                                 accepted state: -(0 + p_0 + p_1 + ...) for GAUSS_ISO, GAUSS_DIAG and EXP_DIAG,
                                 log p for GMM_ISO.  params: [j] (default 0) */
#define MCI_OBS_PDF_PROTO  10 /* nobs ndim: the proto values p_i of sampling function j (GAUSS_DIAG or EXP_DIAG,
                                 nproto == ndim), e.g. a_i x_i^2.  params: [j] (default 0) */
#define MCI_OBS_DEP_SQUARE 11 /* nobs of observable k: the squares of observable k's last values.  k must be an
                                 earlier observable whose nskip divides this one's (the two rules of
                                 DependentObservableInterface::isObsDepValid, DependentObservableInterface.hpp:36-44);
                                 in the automatic decorrelation k must be an equilibration observable too.
                                 params: [k] */
#define MCI_OBS_DEVICE    100 /* the caller's own observable as device code (mci_add_observable_device) */

/* ---- random-number modes ---- */
#define MCI_RNG_MT19937_64  0 /* reference streams: one std::mt19937_64 per walker, libstdc++ distributions
                                 (MCIntegrator.hpp:32-34, TrialMoveInterface.hpp:120-124, SRRDVecMove.hpp:27) */
#define MCI_RNG_PHILOX      1 /* native: Philox4x32-10, key = seed, counter = (word block, walker, step) */

/* ---- arithmetic of the chain ---- */
#define MCI_PREC_FP64       0
#define MCI_PREC_FP32       1 /* FP32 state/proposal/exp, Kahan FP32 block sums, FP64 estimators */

/* ---- walker coupling ---- */
#define MCI_COUPLING_INDEPENDENT 0 /* W independent MCI objects (USE_MPI=0 semantics per walker) */
#define MCI_COUPLING_SHARED      1 /* W MPI ranks (USE_MPI=1 semantics, shared calibration/decorrelation) */

/* ---- kernel selection (no counterpart in the reference: a tuning knob) ---- */
#define MCI_KERNELS_AUTO     0 /* specialised kernels where the configuration allows, else generic */
#define MCI_KERNELS_GENERIC  1 /* always the generic kernels (reference-order accumulation) */

typedef struct mci_ctx mci_ctx;

/* thread-local message of the last failing call (never NULL) */
const char *mci_last_error(void);
/* library version string and the HIP device it runs on */
const char *mci_version(void);
mci_status mci_device_count(int *count);

/* ---- lifetime: MCI::MCI(int ndim) src/MCIntegrator.cpp:624-649 ---- */
mci_status mci_create(int ndim, int64_t nwalkers, mci_ctx **out);
mci_status mci_destroy(mci_ctx *ctx);

/* ---- settings ---- */
/* MCI::setSeed src/MCIntegrator.cpp:546-549.  Walker w gets seed_w = seed for w == 0 and a
 * splitmix64-derived seed otherwise (MT mode); Philox uses seed as the key. */
mci_status mci_set_seed(mci_ctx *ctx, uint64_t seed);
/* per-walker seeds for MT mode (the role of MPIMCI::setSeed src/MPIMCI.cpp:38-65) */
mci_status mci_set_walker_seeds(mci_ctx *ctx, const uint64_t *seeds, int64_t nseeds);
mci_status mci_set_rng(mci_ctx *ctx, int rng_kind);
mci_status mci_set_precision(mci_ctx *ctx, int precision);
mci_status mci_set_coupling(mci_ctx *ctx, int coupling);
/* MCI_KERNELS_AUTO / MCI_KERNELS_GENERIC.  Both give identical accept/reject decisions,
 * positions and results, except the block means of single-coordinate moves on the native
 * streams, which the specialised kernel sums by dwell time (within 1e-15 relative). */
mci_status mci_set_kernel_path(mci_ctx *ctx, int path);
/* MCI::setTargetAcceptanceRate src/MCIntegrator.cpp:551-554 */
mci_status mci_set_target_acceptance_rate(mci_ctx *ctx, double rate);
/* MCI::setNfindMRT2Iterations / setNdecorrelationSteps MCIntegrator.hpp:115-123 (sign convention kept) */
mci_status mci_set_nfind_mrt2_iterations(mci_ctx *ctx, int niterations);
mci_status mci_set_ndecorrelation_steps(mci_ctx *ctx, int64_t nsteps);
mci_status mci_get_settings(mci_ctx *ctx, double *target_rate, int *nfind, int64_t *ndecorr);

/* ---- walker positions: MCI::setX/getX/centerX src/MCIntegrator.cpp:594-604, MCIntegrator.hpp:108,202-203 ----
 * x has layout [nwalkers][ndim]; n == ndim broadcasts one position to all walkers. */
mci_status mci_set_x(mci_ctx *ctx, const double *x, int64_t n);
mci_status mci_get_x(mci_ctx *ctx, double *x, int64_t n);
mci_status mci_center_x(mci_ctx *ctx);
/* MCI::moveX src/MCIntegrator.cpp:606-612 and MCI::newRandomX :614-619 (draw from each walker's stream) */
mci_status mci_move_x(mci_ctx *ctx);
mci_status mci_new_random_x(mci_ctx *ctx);

/* ---- step sizes: MCI::setMRT2Step (3 overloads) / getMRT2Step src/MCIntegrator.cpp:556-585 ----
 * idx < 0 sets all step sizes; per_walker values have layout [nwalkers][nsteps]. */
mci_status mci_set_mrt2step(mci_ctx *ctx, int idx, double step);
mci_status mci_set_mrt2steps(mci_ctx *ctx, const double *steps, int nsteps);
mci_status mci_get_mrt2step(mci_ctx *ctx, int idx, double *step);            /* walker 0 */
mci_status mci_get_walker_mrt2steps(mci_ctx *ctx, double *steps, int64_t n); /* [nwalkers][nsteps] */
mci_status mci_get_n_step_sizes(mci_ctx *ctx, int *n);

/* ---- domain: MCI::setDomain/resetDomain/setIRange src/MCIntegrator.cpp:381-415;
 *      OrthoPeriodicDomain src/OrthoPeriodicDomain.cpp:38-61; UnboundDomain.hpp ---- */
mci_status mci_reset_domain(mci_ctx *ctx);
mci_status mci_set_irange(mci_ctx *ctx, const double *lbounds, const double *ubounds); /* ndim each */
mci_status mci_domain_is_finite(mci_ctx *ctx, int *finite);

/* ---- trial move: MCI::setTrialMove(SRRDType, veclen, ntypes, typeEnds) src/MCIntegrator.cpp:437-453,
 *      setTrialMove(MoveType) :429-435.  veclen 0 = all-move (SRRDAllMove.hpp:67-80),
 *      veclen > 0 = single-vector move (SRRDVecMove.hpp:75-96). typeEnds may be NULL if ntypes == 1. */
mci_status mci_set_trial_move(mci_ctx *ctx, int srrd, int veclen, int ntypes, const int *type_ends);
mci_status mci_set_trial_move_type(mci_ctx *ctx, int move_type);
/* MultiStepMove(ndim, nsteps) (MultiStepMove.hpp:40-46, MultiStepMove.cpp:6-47): every trial move
 * is nsteps Metropolis steps of the contained SRRD move (built from srrd, veclen, ntypes,
 * type_ends as mci_set_trial_move does) under the move's own sampling functions, with acceptance
 * factor oldPDF / newPDF of those; without own sampling functions every inner step is accepted.
 * The step sizes are the contained move's.  The generic kernels run it. */
mci_status mci_set_trial_move_multistep(mci_ctx *ctx, int nsteps, int srrd, int veclen, int ntypes, const int *type_ends);
/* MultiStepMove::addSamplingFunction / clearSamplingFunctions: the move's own sampling
 * functions, from the same catalogue as mci_add_sampling_function (MCI_E_RUNTIME if the trial
 * move is not a MultiStepMove) */
mci_status mci_add_multistep_sampling_function(mci_ctx *ctx, int kind, int ndim, const double *params, int nparams);
/* A mixed trial move with a learned half (the reference's non-symmetric trial-move hook: the move's
 * factor multiplies the acceptance ratio, TrialMoveInterface.hpp:63-69, MCIntegrator.cpp:343, as
 * MultiStepMove.cpp:46 does): at every step, with probability pmix, an independence move x' ~ q from
 * the diagonal Gaussian mixture q = sum_k pi_k N(mu_k, diag sigma_k^2) (K <= 8 components: weights
 * [K] normalised here, means [K][ndim], sigmas [K][ndim]) with factor q(x) / q(x'); otherwise the move
 * set by mci_set_trial_move / _type / _multistep before (a later call of those replaces the whole
 * mixed move, as setTrialMove replaces a move; set the learned half again after it).  The target stays exact for
 * any mixture; a mixture fitted to the target (e.g. the learned proposal's, mci_lp_get_mixture)
 * moves walkers between far-apart modes.  It composes with every sampling function, observable and
 * domain of the catalogue; the generic kernels run it. */
mci_status mci_set_learned_move(mci_ctx *ctx, int K, const double *weights, const double *means, const double *sigmas, double pmix);
mci_status mci_clear_learned_move(mci_ctx *ctx);
mci_status mci_clear_multistep_sampling_functions(mci_ctx *ctx);
/* MCI::storeObservablesOnFile / storeWalkerPositionsOnFile and their clear calls
 * (MCIntegrator.cpp:495-542): during the main sampling of each integrate, every freq steps (the
 * step index first, then the last observable values, or the position), in the reference's stream
 * format.  One file per walker: the path itself for one walker, path.<walker> for several.  The
 * values are recorded on the device (at most 16 GB) and written when the sampling ends; FP64 only. */
mci_status mci_store_observables_on_file(mci_ctx *ctx, const char *path, int freq);
/* MCI::setCallback / clearCallback (MCIntegrator.hpp:189-190; calls at MCIntegrator.cpp:265,
 * 346-347, 373): the callback sees, at the start of every sample() call of calibration,
 * decorrelation and sampling and at every step after the decision, walker 0's committed
 * position and its accepted / rejected counters.  The device records up to `capacity` such
 * states during integrate and the calls are made in order before mci_integrate returns; more
 * calls than the capacity give MCI_E_RUNTIME after the first `capacity`.  The generic kernels run. */
typedef struct mci_callback_view
{
    int64_t call;      /* 0, 1, ... within the integrate */
    int ndim;
    const double *x;   /* [ndim], valid during the call */
    int64_t acc, rej;
    double rate;       /* acc / (acc + rej), 0 while acc is 0 (MCI::getAcceptanceRate) */
} mci_callback_view;
typedef void (*mci_callback_fn)(void *user, const mci_callback_view *view);
mci_status mci_set_callback(mci_ctx *ctx, mci_callback_fn fn, void *user, int64_t capacity);
mci_status mci_clear_callback(mci_ctx *ctx);
mci_status mci_clear_observable_file(mci_ctx *ctx);
mci_status mci_store_walker_positions_on_file(mci_ctx *ctx, const char *path, int freq);
mci_status mci_clear_walker_file(mci_ctx *ctx);
mci_status mci_get_change_rate(mci_ctx *ctx, double *rate); /* TrialMoveInterface::getChangeRate */

/* ---- sampling functions: MCI::addSamplingFunction src/MCIntegrator.cpp:486-492,
 *      popSamplingFunction / clearSamplingFunctions MCIntegrator.hpp:174-175 ---- */
mci_status mci_add_sampling_function(mci_ctx *ctx, int kind, int ndim, const double *params, int nparams);
mci_status mci_pop_sampling_function(mci_ctx *ctx);
mci_status mci_clear_sampling_functions(mci_ctx *ctx);
mci_status mci_get_npdf(mci_ctx *ctx, int *npdf);

/* ---- observables: MCI::addObservable(obs, blocksize, nskip, flag_equil, EstimatorType)
 *      src/MCIntegrator.cpp:455-469; popObservable :481-484; clearObservables MCIntegrator.hpp:168 ---- */
mci_status mci_add_observable(mci_ctx *ctx, int kind, int ndim, const double *params, int nparams,
                              int blocksize, int nskip, int flag_equil, int estimator);
mci_status mci_pop_observable(mci_ctx *ctx);
mci_status mci_clear_observables(mci_ctx *ctx);
mci_status mci_get_nobs(mci_ctx *ctx, int *nobs);
mci_status mci_get_nobsdim(mci_ctx *ctx, int *nobsdim);

/* ---- the caller's own plugins as device code: a user subclass of SamplingFunctionInterface
 *      (SamplingFunctionInterface.hpp:74-97 with ProtoFunctionInterface.hpp:24-65) or of
 *      ObservableFunctionInterface (ObservableFunctionInterface.hpp:48-61) added with
 *      MCI::addSamplingFunction / MCI::addObservable (src/MCIntegrator.cpp:455-492).  `source` is
 *      HIP C++ defining, for the C identifier `name` (vectors as MciVec, mci_amd/device_functor.hpp),
 *        sampling function: NAME_proto (protoFunction), NAME_acceptance (acceptanceFunction),
 *                           NAME_sampling (samplingFunction), over nproto proto values;
 *        observable:        NAME_observable (observableFunction), nobs values;
 *      params are passed to every call.  The chain's generic kernels are compiled with the sources
 *      at run time (hiprtc) for the context's GPU, once per set of functors and process, at the
 *      first integrate or at mci_build_device_functors, whose error carries the compiler's messages.
 *      The full proto recomputation is the reference's default updatedAcceptance
 *      (SamplingFunctionInterface.hpp:85-90); observables are evaluated in full
 *      (flag_updateable false). ---- */
mci_status mci_add_sampling_function_device(mci_ctx *ctx, const char *name, const char *source, int ndim, int nproto,
                                            const double *params, int nparams);
mci_status mci_add_observable_device(mci_ctx *ctx, const char *name, const char *source, int ndim, int nobs,
                                     const double *params, int nparams, int blocksize, int nskip, int flag_equil, int estimator);
/* MCI::setTrialMove(std::unique_ptr<TrialMoveInterface>) (MCIntegrator.hpp:140, MCIntegrator.cpp:399-424)
 * with the caller's own TrialMoveInterface subclass (TrialMoveInterface.hpp:16-70) as device code:
 * `source` defines NAME_move (trialMove) and, with nproto > 0 proto values, NAME_proto
 * (protoFunction).  nsteps step sizes (getNStepSizes), typeEnds[nsteps] their coordinate ranges as
 * for mci_set_trial_move (getStepSizeIndex; may be NULL for one), change_rate its getChangeRate,
 * steps[nsteps] the initial step sizes (NULL: 0.05).  The move's draws come from the chain's stream
 * (mci_uniform); its proto values follow accept and reject (newToOld / oldToNew,
 * MCIntegrator.cpp:350-359).  Replaced by mci_set_trial_move and the other move setters. */
mci_status mci_set_trial_move_device(mci_ctx *ctx, const char *name, const char *source, int nproto, int nsteps, const int *typeEnds,
                                     double change_rate, const double *steps, const double *params, int nparams);
mci_status mci_build_device_functors(mci_ctx *ctx);
/* compiles one functor (what: 0 sampling function, 1 observable, 2 trial move) with the chain's
 * kernels for gfx950 without a GPU; log (capacity logcap) receives the compiler's messages.  No
 * counterpart in the reference: a check before a GPU run. */
mci_status mci_check_device_functor(const char *name, const char *source, int what, char *log, int64_t logcap);

/* ---- integrate: MCI::integrate src/MCIntegrator.cpp:43-82 (combine over walkers:
 *      src/MPIMCI.cpp:85-92).  avg/err have nobsdim entries. ---- */
mci_status mci_integrate(mci_ctx *ctx, int64_t nmc, double *average, double *error,
                         int do_find_mrt2step, int do_decorrelation);
/* per-walker estimates of the last integrate, layout [nwalkers][nobsdim] */
mci_status mci_get_walker_results(mci_ctx *ctx, double *average, double *error, int64_t n);
/* finalised accumulator data of observable `obs` after the last integrate, layout
 * [nwalkers][nstore][nobs] (AccumulatorInterface::getData AccumulatorInterface.hpp:81; the
 * device-side counterpart of MCI::storeObservablesOnFile MCIntegrator.cpp:495-509 is reading the
 * Full accumulator's trace).  data == NULL only reports nstore. */
mci_status mci_get_observable_data(mci_ctx *ctx, int obs, double *data, int64_t n, int64_t *nstore);
/* MCI::getAcceptanceRate src/MCIntegrator.cpp:587-592: counters of the last sample() call,
 * summed over walkers; rate = mean over walkers of acc/(acc+rej) */
mci_status mci_get_acceptance(mci_ctx *ctx, int64_t *accepted, int64_t *rejected, double *rate);
mci_status mci_get_walker_acceptance(mci_ctx *ctx, int64_t *accepted, int64_t *rejected, int64_t n);

/* ---- checkpoint and resume of the walkers ----
 * The reference keeps a chain's state in the MCI object between integrate calls (positions, step
 * sizes, the private engine _rgen, MCIntegrator.hpp:33, and the move's distribution objects); a
 * second integrate continues it (test/ut2/main.cpp:74-76).  These calls save and restore that state
 * of every walker, so that a run can continue in another context or process with the same
 * configuration: positions, step sizes, the random streams (mt19937_64 states, or the Philox key and
 * counters) and the saved normals of the move's distribution.  The blob is plain bytes (a header
 * with the shape, then the arrays); restoring needs the same ndim, walkers, step-size count and
 * random mode (MCI_E_INVALID_ARGUMENT otherwise). */
mci_status mci_get_state_size(mci_ctx *ctx, int64_t *nbytes);
mci_status mci_get_state(mci_ctx *ctx, void *buf, int64_t nbytes);
mci_status mci_set_state(mci_ctx *ctx, const void *buf, int64_t nbytes);

/* ---- timing of the last integrate (HIP events on the context's stream, milliseconds) ----
 * phases: [0] calibration, [1] decorrelation, [2] sampling kernel, [3] estimation+combine, [4] total */
mci_status mci_get_timings(mci_ctx *ctx, double *ms, int n);

/* ---- standalone estimators on host arrays (x has layout [n][ndim] like the reference),
 *      run on the device: Estimators.cpp:125-155,191-246,271-286, MJBlocker.cpp:127-154 ---- */
mci_status mci_estimate(int estimator, int64_t n, int ndim, const double *x, double *average, double *error);
/* The same estimators without any allocation.  No counterpart in the reference: this entry point is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  x [n][ndim], average [ndim] and error [ndim] are DEVICE pointers the caller owns; the work
 * is enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call returns without
 * waiting.  The only other memory is a fixed array of MCI_ESTIMATE_DEVICE_MAX doubles per device,
 * held until the process ends: n * ndim must not exceed it (MCI_E_INVALID_ARGUMENT otherwise).  All
 * calls on one device share that array, so calls that overlap on one device (different streams,
 * or different host threads) race; serialise them.  The first call on a device looks up the
 * array's address (hipGetSymbolAddress): make it outside a graph capture. */
#define MCI_ESTIMATE_DEVICE_MAX 2097152
mci_status mci_estimate_device(int estimator, int64_t n, int ndim, const double *x, double *average, double *error, void *stream);

/* ---- batched integrals ----
 * Many independent integrals in one submission.  Each integral i is what one reference MCI
 * object computes with an isotropic Gaussian-mixture sampling function
 * (weights w_k, widths s_k, means mu_k; MCI_PDF_GMM_ISO's density), the |x|^2 observable
 * (MCI_OBS_X2_SUM, c = 1), single-coordinate uniform moves (SRRDVecMove, veclen 1) and step
 * calibration (MCI::findMRT2Step src/MCIntegrator.cpp:87-168, with one wave of 64 pooled lanes
 * as 64 MPI ranks) — i.e. it replaces a loop of MCI::integrate calls (src/MCIntegrator.cpp:43-82),
 * one per integral, whose results MPIMCI would combine (src/MPIMCI.cpp:85-92).
 *
 * Work per integral is split over R_i walkers of T_i steps each (R_i * T_i >= nmc_i); each
 * walker starts from an exact draw of its mixture (so it is stationary from step 0), keeps FP32
 * coordinates with Kahan-compensated FP32 block sums flushed to FP64, and stores nblocks block
 * means.  Walker means give avg_i (FP64, fixed summation order); err_i is the standard error of
 * the R_i independent walker means (MCI_BATCH_ERR_SPREAD, the default) or the MPIMCI combine of
 * per-walker MJBlocker errors (MCI_BATCH_ERR_COMBINE).  R_i and T_i (hence the block length
 * T_i / nblocks) come from the autotuner unless set.  All walkers run in one persistent launch
 * that takes 64-walker work items from a device queue.
 *
 * Parameter arrays are concatenated over integrals in index order:
 *   ncomp[i] in 1..4, weights and widths [sum ncomp], means [sum ncomp_i * ndim_i]
 *   (component-major within an integral).  ndim[i] in 1..MCI_BATCH_MAX_DIM. */
#define MCI_BATCH_MAX_DIM      256
#define MCI_BATCH_MAX_COMP     4
#define MCI_BATCH_ERR_SPREAD   0
#define MCI_BATCH_ERR_COMBINE  1
typedef struct mci_batch mci_batch;
mci_status mci_batch_create(int nint, const int *ndim, const int64_t *nmc, const int *ncomp, const double *weights,
                            const double *widths, const double *means, mci_batch **out);
mci_status mci_batch_destroy(mci_batch *b);
mci_status mci_batch_set_seed(mci_batch *b, uint64_t seed);
/* blocks per walker: a power of two in 4..64 (default 16) */
mci_status mci_batch_set_blocks(mci_batch *b, int nblocks);
/* fewest walkers per integral the autotuner aims for (default 64; 1..65536) */
mci_status mci_batch_set_min_walkers(mci_batch *b, int rmin);
mci_status mci_batch_set_error_mode(mci_batch *b, int mode);
/* calibration: MCI::setNfindMRT2Iterations / setTargetAcceptanceRate semantics (defaults -50, 0.5);
 * nfind == 0 keeps the initial step 2 * sum_k w_k s_k */
mci_status mci_batch_set_calibration(mci_batch *b, int nfind, double target_rate);
/* decorrelation steps per walker before accumulation (default 0: walkers start stationary) */
mci_status mci_batch_set_ndecorrelation_steps(mci_batch *b, int64_t nsteps);
/* explicit plan: steps per walker T_i (a power of two times nblocks); NULL returns to autotuning */
mci_status mci_batch_set_plan(mci_batch *b, const int64_t *steps_per_walker);
/* autotune the plan: measure == 0 uses the cost model only (no device work), measure > 0 also
 * times that many candidate plans on the device and keeps the fastest */
mci_status mci_batch_autotune(mci_batch *b, int measure);
/* the plan in force: walkers R_i and steps per walker T_i, [nint] each; the walkers of integral i
 * are walker_offset[i] .. walker_offset[i] + R_i - 1 in the walker order of the per-walker getters */
mci_status mci_batch_get_plan(mci_batch *b, int64_t *walkers, int64_t *steps_per_walker, int64_t *walker_offset);
mci_status mci_batch_get_nwalkers(mci_batch *b, int64_t *nwalkers);
/* calibrate, sample, estimate: avg/err [nint] in integral order */
mci_status mci_batch_run(mci_batch *b, double *average, double *error);
/* capture calibration, sampling and estimation into one hipGraph and replay it on later runs
 * (default on) */
mci_status mci_batch_set_graph(mci_batch *b, int on);
/* results of the last run: calibrated step and its last pooled acceptance rate [nint];
 * acceptance rate of the sampling walkers [nint]; walker means/errors [nwalkers];
 * block means [nblocks][nwalkers] */
mci_status mci_batch_get_steps(mci_batch *b, double *steps, double *calib_rate);
mci_status mci_batch_get_acceptance(mci_batch *b, double *rate);
mci_status mci_batch_get_walker_results(mci_batch *b, double *average, double *error, int64_t n);
mci_status mci_batch_get_block_means(mci_batch *b, double *data, int64_t n);
/* [0] calibration, [1] sampling (persistent launch), [2] estimation+combine, [3] total (ms) */
mci_status mci_batch_get_timings(mci_batch *b, double *ms, int n);

/* ---- learned proposal ----
 * A trial move whose proposal is learned (no single reference counterpart; it plugs into the
 * reference's trial-move hook, whose return value multiplies the acceptance ratio,
 * TrialMoveInterface.hpp:63-69 and src/MCIntegrator.cpp:343, as MultiStepMove.cpp:46 does for a
 * non-symmetric move).  The network maps the walker's position x to a K-component diagonal
 * Gaussian mixture q(x' | x); a learned move draws x' ~ q(. | x) and is accepted with
 * min(1, p(x') q(x | x') / (p(x) q(x' | x))), so the target p stays exact whatever the network
 * computes.  The network, d -> H -> H -> K (1 + 2 d) with SiLU, runs with BF16 weights and
 * activations and FP32 accumulation on the matrix cores, forward and backward, and is trained
 * with Adam on pairs of samples of the plain chain (same walker at two times, or two walkers)
 * to maximise log q(x_b | x_a). */
typedef struct mci_mlp mci_mlp;
mci_status mci_mlp_create(int d, int hidden, int ncomp, uint64_t seed, mci_mlp **out);
mci_status mci_mlp_destroy(mci_mlp *m);
/* parameters, flat FP32: W1 [d][H], b1 [H], W2 [H][H], b2 [H], W3 [H][K(1+2d)], b3 [K(1+2d)] */
mci_status mci_mlp_nparams(mci_mlp *m, int64_t *n);
mci_status mci_mlp_get_params(mci_mlp *m, float *params, int64_t n);
mci_status mci_mlp_set_params(mci_mlp *m, const float *params, int64_t n);
/* raw outputs [n][K(1+2d)] for positions x [n][d]: for each component k, d means / 4 then d
 * log-scale pre-activations (log-scale = o clamped to [-3, 3]); then K logits.  The input is x / 4. */
mci_status mci_mlp_forward(mci_mlp *m, int64_t n, const float *x, float *out);
/* mean of -log q(y_r | x_r) over n pairs and its gradient with respect to the FP32 parameters */
mci_status mci_mlp_loss_grad(mci_mlp *m, int64_t n, const float *x, const float *y, double *loss, float *grad);

/* the sampler: target p(x) = sum_k w_k N(x; mu_k, Sigma_k), d <= 32, K <= 8 (means [K][d],
 * covariances [K][d][d]); nwalkers independent walkers start from exact draws of p.
 * Observables: |x|^2 and 1[x_0 > 0]. */
#define MCI_LP_PLAIN_FP64   0  /* random-walk Metropolis (single-coordinate uniform moves), FP64 */
#define MCI_LP_PLAIN_FP32   1  /* the same chain in FP32 with Kahan block sums */
#define MCI_LP_MIXED_FP32   2  /* cycles of 32 steps: round(32 (1 - pmix)) random-walk moves, then round(32 pmix)
                                  (at least 1) learned moves; pmix in (0, 1] */
#define MCI_LP_INDEP_FP32   3  /* the same cycles, the learned moves from the independence proposal q(x') = the
                                  network's mixture at one fixed input (the mean of the last training snapshot),
                                  evaluated once per run */
typedef struct mci_lp mci_lp;
mci_status mci_lp_create(int d, int ncomp, const double *weights, const double *means, const double *covs, int64_t nwalkers,
                         mci_lp **out);
mci_status mci_lp_destroy(mci_lp *lp);
mci_status mci_lp_set_seed(mci_lp *lp, uint64_t seed);
mci_status mci_lp_set_network(mci_lp *lp, int hidden, int ncomp); /* default 128, 8 (resets the network) */
/* exact start, step calibration (MCI::findMRT2Step rule, walkers as ranks), nsteps steps in
 * nblocks blocks (a power of two <= 64), MJBlocker per walker.  avg/err [2]: mean over walkers
 * and the standard error of the walker means. */
mci_status mci_lp_run(mci_lp *lp, int mode, int64_t nsteps, int nblocks, double pmix, double *avg, double *err);
/* training samples: nsnap snapshots of every walker of the plain FP32 chain, lag steps apart */
mci_status mci_lp_collect(mci_lp *lp, int nsnap, int64_t lag);
/* training pairs (x_a, x_b): x_b is the same walker at another snapshot with probability same_fraction
 * (default 0.5; teaches moves within a mode), else another walker (an independent sample of p) */
mci_status mci_lp_set_training_pairs(mci_lp *lp, double same_fraction);
/* on: every training input is one fixed position (the mean of the last snapshot), so the network's
 * mixture there is an independence proposal q(x'), fitted to the plain chain's samples; it is the
 * proposal of MCI_LP_INDEP_FP32 (default off) */
mci_status mci_lp_set_fixed_input(mci_lp *lp, int on);
/* before training: fit a diagonal Gaussian mixture (the network's K components) to the first
 * 16,384 walkers of the last snapshot (mci_gmm_fit_diag_device on the sampler's device, em_iters
 * iterations) and write it into the network's last layer (biases = the fit, weights scaled by
 * 0.02), so that training starts from that mixture */
mci_status mci_lp_init_network_from_samples(mci_lp *lp, int em_iters);
/* host only (no device): diagonal Gaussian mixture of K components fitted to x [n][d] by em_iters
 * EM iterations from greedy k-means++ seeds (16 candidates per seed, stream of `seed`); writes
 * mu [K][d], var [K][d], pi [K] and, if loglik is not null, the mean log-likelihood of x */
mci_status mci_gmm_fit_diag(const double *x, int64_t n, int d, int K, int em_iters, uint64_t seed, double *mu, double *var,
                            double *pi, double *loglik);
/* the same fit on the current device (d <= 128, K <= 16; x and the outputs in host memory): the
 * seeding draws the same candidates (their squared distances are bit-identical to the host fit's,
 * the potentials that pick among them are sums in another order), the EM iterations run in FP64
 * device kernels with every sum over samples in a fixed order, so the result is deterministic and
 * agrees with mci_gmm_fit_diag to rounding */
mci_status mci_gmm_fit_diag_device(const double *x, int64_t n, int d, int K, int em_iters, uint64_t seed, double *mu, double *var,
                                   double *pi, double *loglik);
/* Adam on minibatches of training pairs; losses of the first and last minibatch */
mci_status mci_lp_train(mci_lp *lp, int iters, int batch, double lr, double *loss_first, double *loss_last);
/* last run: [0..1] avg, [2..3] standard error of walker means, [4..5] MPIMCI combine of the
 * walkers' MJBlocker errors, [6] random-walk acceptance, [7] learned-move acceptance */
mci_status mci_lp_get_results(mci_lp *lp, double *results, int n);
/* [0] calibration, [1] sampling, [2] estimation, [3] training (ms, host clock around synchronised work) */
mci_status mci_lp_get_timings(mci_lp *lp, double *ms, int n);
mci_status mci_lp_get_walker_means(mci_lp *lp, double *means, int64_t n); /* [2][nwalkers] */
mci_status mci_lp_get_network(mci_lp *lp, mci_mlp **m);                   /* owned by the sampler */
/* The trained independence proposal (the network's mixture at the fixed input, as MCI_LP_INDEP_FP32
 * uses it: means and widths rounded to FP32) in the form mci_set_learned_move takes, so that a context
 * of the catalogue path can mix it into its own trial move: *K components, weights [K] (normalised),
 * means [K][d], sigmas [K][d]; the arrays hold kcap components. */
mci_status mci_lp_get_mixture(mci_lp *lp, int kcap, int *K, double *weights, double *means, double *sigmas);
/* The independence learned move alone (MCI_LP_INDEP_FP32's learned steps, TrialMoveInterface's
 * non-symmetric hook, TrialMoveInterface.hpp:63-69): n walkers at the FP32 positions x [d][n] take nl
 * consecutive moves of the current independence proposal, steps t0 .. t0 + nl - 1 of their Philox
 * streams (walker w uses stream w); x is updated in place, accepted [n] receives each walker's
 * accepted moves.  mix (optional, 3 * 8 * d + 16 floats) receives the proposal as the move uses it:
 * means [8][d], 1 / sigma [8][d], sigma [8][d], log normalisers [8] and cumulative weights [8]
 * (components past the network's K are inert).  For callers that compose the move with moves of
 * their own, and for checking it. */
mci_status mci_lp_indep_move(mci_lp *lp, int64_t n, float *x, int nl, uint64_t t0, int32_t *accepted, float *mix);
/* the network as the learned move's kernel evaluates it (BF16 matrix cores, transformed outputs), at
 * n positions x [d][n]: out [K(1+2d)][n] holds per component k the d means, then the d log-scales,
 * then the K logits.  Needs d in {8, 16, 32} and hidden in {64, 128}. */
mci_status mci_lp_eval_network(mci_lp *lp, int64_t n, const float *x, float *out);

#ifdef __cplusplus
}
#endif

#endif /* MCI_AMD_H */
