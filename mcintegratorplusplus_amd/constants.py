"""Constants of include/mci_amd.h (enums mirror include/mci/Factories.hpp of the reference)."""
OK, E_INVALID_ARGUMENT, E_DOMAIN, E_RUNTIME, E_DEVICE = 0, 1, 2, 3, 4

# EstimatorType (Factories.hpp:52-59)
ESTIM_NOOP, ESTIM_UNCORRELATED, ESTIM_CORRELATED, ESTIM_FCBLOCKER, ESTIM_MJBLOCKER = 0, 1, 2, 3, 4
# MoveType (Factories.hpp:108-113) and SRRDType (Factories.hpp:119-133)
MOVE_ALL, MOVE_VEC, MOVE_MULTISTEP, MOVE_DEVICE = 0, 1, 2, 3
(SRRD_UNIFORM, SRRD_GAUSSIAN, SRRD_STUDENT, SRRD_CAUCHY, SRRD_EXPONENTIAL, SRRD_GAMMA, SRRD_WEIBULL, SRRD_LOGNORMAL,
 SRRD_CHISQ, SRRD_FISHER) = range(10)
# device functor catalogue
PDF_GAUSS_ISO, PDF_GAUSS_DIAG, PDF_EXP_DIAG, PDF_GMM_ISO = 1, 2, 3, 4
PDF_DEVICE = 100  # the caller's own sampling function as device code
OBS_X2, OBS_X2_SUM, OBS_X, OBS_X_SUM, OBS_CONST, OBS_EXP_X2, OBS_INDICATOR, OBS_GAUSS_X2 = 1, 2, 3, 4, 5, 6, 7, 8
OBS_PDF_LOG, OBS_PDF_PROTO, OBS_DEP_SQUARE = 9, 10, 11  # dependent observables
OBS_DEVICE = 100  # the caller's own observable as device code
RNG_MT19937_64, RNG_PHILOX = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
COUPLING_INDEPENDENT, COUPLING_SHARED = 0, 1
KERNELS_AUTO, KERNELS_GENERIC = 0, 1

# batched integrals: error of each integral
BATCH_ERR_SPREAD, BATCH_ERR_COMBINE = 0, 1
