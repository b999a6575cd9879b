/*
 * mci_oracle.h — CPU oracle of the Metropolis sampling path (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of the reference's algorithm, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only as
 * the checker.  The product (mcintegratorplusplus_amd) never links or calls it.
 *
 * The API mirrors include/mci_amd.h one-to-one with an orc_ prefix, so a test
 * can build the same integrator on both sides and compare.  Every function
 * cites the reference file:line it restates; see mci_oracle.c.
 *
 * Arithmetic is FP64 with contraction disabled (-ffp-contract=off), matching
 * the reference built with -O2 -ffp-contract=off.
 */
#ifndef MCI_ORACLE_H
#define MCI_ORACLE_H

#include <stdint.h>
#include "../include/mci_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

const char *orc_last_error(void);

mci_status orc_create(int ndim, int64_t nwalkers, orc_ctx **out);
mci_status orc_destroy(orc_ctx *c);

mci_status orc_set_seed(orc_ctx *c, uint64_t seed);
mci_status orc_set_walker_seeds(orc_ctx *c, const uint64_t *seeds, int64_t nseeds);
mci_status orc_set_rng(orc_ctx *c, int rng_kind);
mci_status orc_set_precision(orc_ctx *c, int precision);
mci_status orc_set_coupling(orc_ctx *c, int coupling);
mci_status orc_set_target_acceptance_rate(orc_ctx *c, double rate);
mci_status orc_set_nfind_mrt2_iterations(orc_ctx *c, int niterations);
mci_status orc_set_ndecorrelation_steps(orc_ctx *c, int64_t nsteps);

mci_status orc_set_x(orc_ctx *c, const double *x, int64_t n);
mci_status orc_get_x(orc_ctx *c, double *x, int64_t n);
mci_status orc_center_x(orc_ctx *c);
mci_status orc_move_x(orc_ctx *c);
mci_status orc_new_random_x(orc_ctx *c);

mci_status orc_set_mrt2step(orc_ctx *c, int idx, double step);
mci_status orc_set_mrt2steps(orc_ctx *c, const double *steps, int nsteps);
mci_status orc_get_mrt2step(orc_ctx *c, int idx, double *step);
mci_status orc_get_walker_mrt2steps(orc_ctx *c, double *steps, int64_t n);
mci_status orc_get_n_step_sizes(orc_ctx *c, int *n);

mci_status orc_reset_domain(orc_ctx *c);
mci_status orc_set_irange(orc_ctx *c, const double *lbounds, const double *ubounds);
mci_status orc_domain_is_finite(orc_ctx *c, int *finite);

mci_status orc_set_trial_move(orc_ctx *c, int srrd, int veclen, int ntypes, const int *type_ends);
mci_status orc_set_trial_move_type(orc_ctx *c, int move_type);
mci_status orc_set_trial_move_multistep(orc_ctx *c, int nsteps, int srrd, int veclen, int ntypes, const int *type_ends);
mci_status orc_set_learned_move(orc_ctx *c, int K, const double *weights, const double *means, const double *sigmas, double pmix);
mci_status orc_clear_learned_move(orc_ctx *c);
mci_status orc_add_multistep_sampling_function(orc_ctx *c, int kind, int ndim, const double *params, int nparams);
mci_status orc_clear_multistep_sampling_functions(orc_ctx *c);
mci_status orc_get_change_rate(orc_ctx *c, double *rate);

mci_status orc_add_sampling_function(orc_ctx *c, int kind, int ndim, const double *params, int nparams);
mci_status orc_pop_sampling_function(orc_ctx *c);
mci_status orc_clear_sampling_functions(orc_ctx *c);
mci_status orc_get_npdf(orc_ctx *c, int *npdf);

mci_status orc_add_observable(orc_ctx *c, int kind, int ndim, const double *params, int nparams,
                              int blocksize, int nskip, int flag_equil, int estimator);
mci_status orc_pop_observable(orc_ctx *c);
mci_status orc_clear_observables(orc_ctx *c);
mci_status orc_get_nobs(orc_ctx *c, int *nobs);
mci_status orc_get_nobsdim(orc_ctx *c, int *nobsdim);

mci_status orc_integrate(orc_ctx *c, int64_t nmc, double *average, double *error,
                         int do_find_mrt2step, int do_decorrelation);
mci_status orc_get_walker_results(orc_ctx *c, double *average, double *error, int64_t n);
mci_status orc_get_acceptance(orc_ctx *c, int64_t *accepted, int64_t *rejected, double *rate);
mci_status orc_get_walker_acceptance(orc_ctx *c, int64_t *accepted, int64_t *rejected, int64_t n);
/* number of "max steps reached" warnings of the last integrate (MCIntegrator.cpp:220) */
mci_status orc_get_warnings(orc_ctx *c, int64_t *nwarn);

/* per-step trace of walker 0 during the next integrate's main sampling:
 * x after each step ([nsteps][ndim]) and the accepted flag; capacity in steps */
mci_status orc_trace_enable(orc_ctx *c, double *xbuf, uint8_t *accbuf, int64_t capacity);
/* FP32 chain (MCI_PREC_FP32): also accumulate the block means of the same FP32 trajectory step by
 * step in FP64 (the reference's BlockAccumulator on the FP32 chain), read by orc_get_shadow_data */
mci_status orc_set_fp32_shadow(orc_ctx *c, int on);
mci_status orc_get_shadow_data(orc_ctx *c, int obs, int64_t walker, double *data, int64_t n, int64_t *nstore);
/* the raw (not yet finalised) accumulator data of observable i, walker w after the last integrate */
mci_status orc_get_accu_data(orc_ctx *c, int obs, int64_t walker, double *data, int64_t n, int64_t *nstore);

/* standalone estimators (Estimators.cpp, MJBlocker.cpp) */
mci_status orc_estimate(int estimator, int64_t n, int ndim, const double *x, double *average, double *error);
mci_status orc_block_estimate(int64_t n, int ndim, const double *x, int64_t nblocks, double *average, double *error);

/* RNG primitives (exposed for known-answer tests) */
uint64_t orc_mt64_nth(uint64_t seed, int64_t n);              /* n-th output (1-based) of mt19937_64(seed) */
void orc_mt64_words(uint64_t seed, int64_t n, uint64_t *out); /* first n outputs */
double orc_canonical53(uint64_t word);                        /* generate_canonical<double,53> of one word */
uint64_t orc_lemire64(const uint64_t *words, int64_t nwords, uint64_t range, int64_t *used); /* uniform_int */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* deterministic reduction used for every cross-walker sum (chunks of 256, tree inside, chunks in order) */
double orc_det_sum(const double *v, int64_t n);

/* batched Gaussian-mixture integrals (mci_batch_oracle.c): FP32 exp/log of the batched chain,
 * the walker order of a plan, and one complete run of a given plan (T = steps per walker).
 * Output arrays may be NULL except avg/err; walker arrays have the plan's walker count,
 * bmeans / bmeans_fp64 are [nblocks][nwalkers] (bmeans_fp64: the same samples summed in FP64). */
float orc_bexp(float x);
float orc_blog(float x);
mci_status orc_batch_walker_order(int nint, const int *ndim, const int *ncomp, const int64_t *nmc, const int64_t *T,
                                  int64_t *walkers, int64_t *woff, int64_t *nwalkers);
mci_status orc_batch_run(int nint, const int *ndim, const int64_t *nmc, const int *ncomp, const double *weights,
                         const double *widths, const double *means, uint64_t seed, int nblocks, int nfind, double target,
                         int64_t ndecorr, int mode, const int64_t *T, double *avg, double *err, double *steps,
                         double *calib_rate, double *accrate, double *wavg, double *werr, double *bmeans,
                         double *bmeans_fp64);

/* the learned independence move (mci_lp_oracle.c): the library's sine/cosine of 2 pi u, its
 * Box-Muller normals of one Philox block, and nl moves of n walkers (positions x [d][n], FP32) with
 * the proposal mix in the layout mci_lp_indep_move returns and the target given by FP64 means
 * [KT][d], precision matrices [KT][d][d] and log normalisers c_k = log w_k - log det(2 pi Sigma_k)/2;
 * decisions [n][nl] and MH margins [n][nl] may be NULL */
void orc_bsincos2pi(float u, float *s, float *c);
void orc_normal4(uint32_t c0, uint32_t id, uint32_t tl, uint32_t th, uint32_t k0, uint32_t k1, float z[4]);
mci_status orc_lp_indep_move(int d, int K, const float *mix, int KT, const double *tmu, const double *tprec, const double *tc,
                             uint64_t seed, int64_t n, float *x, int nl, uint64_t t0, int32_t *accepted, uint8_t *decisions,
                             double *margins);

#ifdef __cplusplus
}
#endif

#endif
