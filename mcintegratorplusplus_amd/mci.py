"""Python mirror of the reference's mci::MCI (include/mci/MCIntegrator.hpp:26-226) over the
C-ABI of the HIP library.  Method names, argument meaning and defaults follow the
reference; failures raise MciError carrying the reference's exception category."""
import ctypes as C

import numpy as np

from . import constants as K
from ._lib import load

_P, _I, _I64, _U64, _D = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_DP, _IP, _I64P, _U64P = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_uint64)

_SIGS = {
    "mci_create": [_I, _I64, C.POINTER(_P)], "mci_destroy": [_P], "mci_set_seed": [_P, _U64],
    "mci_set_walker_seeds": [_P, _U64P, _I64], "mci_set_rng": [_P, _I], "mci_set_precision": [_P, _I],
    "mci_set_coupling": [_P, _I], "mci_set_kernel_path": [_P, _I], "mci_set_target_acceptance_rate": [_P, _D],
    "mci_set_nfind_mrt2_iterations": [_P, _I], "mci_set_ndecorrelation_steps": [_P, _I64],
    "mci_get_settings": [_P, _DP, _IP, _I64P],
    "mci_set_x": [_P, _DP, _I64], "mci_get_x": [_P, _DP, _I64], "mci_center_x": [_P], "mci_move_x": [_P],
    "mci_new_random_x": [_P], "mci_set_mrt2step": [_P, _I, _D], "mci_set_mrt2steps": [_P, _DP, _I],
    "mci_get_mrt2step": [_P, _I, _DP], "mci_get_walker_mrt2steps": [_P, _DP, _I64], "mci_get_n_step_sizes": [_P, _IP],
    "mci_reset_domain": [_P], "mci_set_irange": [_P, _DP, _DP], "mci_domain_is_finite": [_P, _IP],
    "mci_set_trial_move": [_P, _I, _I, _I, _IP], "mci_set_trial_move_type": [_P, _I], "mci_get_change_rate": [_P, _DP],
    "mci_set_learned_move": [_P, _I, _DP, _DP, _DP, _D], "mci_clear_learned_move": [_P],
    "mci_add_sampling_function": [_P, _I, _I, _DP, _I], "mci_pop_sampling_function": [_P],
    "mci_clear_sampling_functions": [_P], "mci_get_npdf": [_P, _IP],
    "mci_add_observable": [_P, _I, _I, _DP, _I, _I, _I, _I, _I], "mci_pop_observable": [_P],
    "mci_clear_observables": [_P], "mci_get_nobs": [_P, _IP], "mci_get_nobsdim": [_P, _IP],
    "mci_integrate": [_P, _I64, _DP, _DP, _I, _I], "mci_get_walker_results": [_P, _DP, _DP, _I64],
    "mci_get_observable_data": [_P, _I, _DP, _I64, _I64P],
    "mci_get_acceptance": [_P, _I64P, _I64P, _DP], "mci_get_walker_acceptance": [_P, _I64P, _I64P, _I64],
    "mci_get_timings": [_P, _DP, _I], "mci_estimate": [_I, _I64, _I, _DP, _DP, _DP], "mci_device_count": [_IP],
    "mci_get_state_size": [_P, _I64P], "mci_get_state": [_P, C.c_void_p, _I64], "mci_set_state": [_P, C.c_void_p, _I64],
    "mci_add_sampling_function_device": [_P, C.c_char_p, C.c_char_p, _I, _I, _DP, _I],
    "mci_add_observable_device": [_P, C.c_char_p, C.c_char_p, _I, _I, _DP, _I, _I, _I, _I, _I],
    "mci_build_device_functors": [_P], "mci_check_device_functor": [C.c_char_p, C.c_char_p, _I, C.c_char_p, _I64],
    "mci_set_trial_move_device": [_P, C.c_char_p, C.c_char_p, _I, _I, _IP, _D, _DP, _DP, _I],
}
_bound = None


def _lib():
    global _bound
    if _bound is None:
        lib = load()
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _I
        lib.mci_last_error.restype = C.c_char_p
        lib.mci_version.restype = C.c_char_p
        _bound = lib
    return _bound


class MciError(RuntimeError):
    """A failing call; .code is the reference's exception category (constants.E_*)."""

    KIND = {K.E_INVALID_ARGUMENT: "invalid_argument", K.E_DOMAIN: "domain_error", K.E_RUNTIME: "runtime_error",
            K.E_DEVICE: "device_error"}

    def __init__(self, code, msg):
        super().__init__(f"{self.KIND.get(code, code)}: {msg}")
        self.code = code


def _check(st):
    if st != K.OK:
        raise MciError(st, _lib().mci_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


def estimate(estimator, x):
    """Run one of the reference's estimators on the GPU over data x [n, ndim] (Estimators.hpp)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    a, e = np.zeros(x.shape[1]), np.zeros(x.shape[1])
    _check(_lib().mci_estimate(estimator, x.shape[0], x.shape[1], _dp(x), _dp(a), _dp(e)))
    return a, e


def check_device_functor(name, source, what="sampling"):
    """Compile a device functor ("sampling", "observable" or "move") with the chain's kernels for
    gfx950 (no GPU needed); raises MciError with the compiler's messages if it does not build."""
    log = C.create_string_buffer(1 << 16)
    _check(_lib().mci_check_device_functor(name.encode(), source.encode(), {"sampling": 0, "observable": 1, "move": 2}[what], log, len(log)))
    return log.value.decode()


class MCI:
    """mci::MCI with nwalkers independent (or MPI-coupled) chains on one GPU."""

    def __init__(self, ndim, nwalkers=1):
        self._lib = _lib()
        h = _P()
        _check(self._lib.mci_create(ndim, nwalkers, C.byref(h)))
        self._h = h
        self.ndim = ndim
        self.nwalkers = nwalkers

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.mci_destroy(h)
            self._h = None

    def _c(self, name, *args):
        _check(getattr(self._lib, name)(self._h, *args))

    # --- setters (MCIntegrator.hpp:93-183) ---
    def setSeed(self, seed): self._c("mci_set_seed", seed)

    def setWalkerSeeds(self, seeds):
        a = np.ascontiguousarray(seeds, dtype=np.uint64)
        self._c("mci_set_walker_seeds", a.ctypes.data_as(_U64P), a.size)

    def setRandomMode(self, mode): self._c("mci_set_rng", mode)
    def setPrecision(self, prec): self._c("mci_set_precision", prec)
    def setWalkerCoupling(self, mode): self._c("mci_set_coupling", mode)
    def setKernelPath(self, path): self._c("mci_set_kernel_path", path)
    def setTargetAcceptanceRate(self, r): self._c("mci_set_target_acceptance_rate", r)
    def setNfindMRT2Iterations(self, n): self._c("mci_set_nfind_mrt2_iterations", n)
    def setNdecorrelationSteps(self, n): self._c("mci_set_ndecorrelation_steps", n)

    def setX(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64).ravel()
        self._c("mci_set_x", _dp(a), a.size)

    def centerX(self): self._c("mci_center_x")
    def moveX(self): self._c("mci_move_x")
    def newRandomX(self): self._c("mci_new_random_x")

    def setMRT2Step(self, step, i=None):
        if np.ndim(step) > 0:
            a = np.ascontiguousarray(step, dtype=np.float64)
            self._c("mci_set_mrt2steps", _dp(a), a.size)
        else:
            self._c("mci_set_mrt2step", -1 if i is None else i, float(step))

    def setIRange(self, lbound, ubound):
        lo = np.ascontiguousarray(np.broadcast_to(lbound, (self.ndim,)), dtype=np.float64)
        hi = np.ascontiguousarray(np.broadcast_to(ubound, (self.ndim,)), dtype=np.float64)
        self._c("mci_set_irange", _dp(lo), _dp(hi))

    def resetDomain(self): self._c("mci_reset_domain")

    def setTrialMove(self, move=K.SRRD_UNIFORM, veclen=0, ntypes=1, typeEnds=None, movetype=None):
        if movetype is not None:
            return self._c("mci_set_trial_move_type", movetype)
        te = None
        if typeEnds is not None:
            te = C.cast((C.c_int * len(typeEnds))(*typeEnds), _IP)
        self._c("mci_set_trial_move", move, veclen, ntypes, te)

    def setLearnedMove(self, weights, means, sigmas, pmix):
        """Mix an independence move from the diagonal Gaussian mixture (weights [K], means [K][ndim],
        sigmas [K][ndim]) into the trial move with probability pmix, e.g. the trained mixture of
        LearnedProposalSampler.mixture(); a later setTrialMove replaces the mixed move."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        m = np.ascontiguousarray(means, dtype=np.float64)
        s = np.ascontiguousarray(sigmas, dtype=np.float64)
        if m.shape != (w.size, self.ndim) or s.shape != m.shape:
            raise ValueError("means and sigmas must be [K][ndim] with K = len(weights)")
        self._c("mci_set_learned_move", w.size, _dp(w), _dp(m), _dp(s), float(pmix))

    def clearLearnedMove(self): self._c("mci_clear_learned_move")

    def addSamplingFunction(self, kind, params=()):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._c("mci_add_sampling_function", kind, self.ndim, _dp(p) if p.size else None, p.size)

    def addSamplingFunctionDevice(self, name, source, nproto, params=()):
        """A sampling function of the caller's as device code (SamplingFunctionInterface's protoFunction,
        acceptanceFunction and samplingFunction as NAME_proto, NAME_acceptance, NAME_sampling over MciVec
        views, include/mci_amd/device_functor.hpp), compiled with the chain's kernels at the first integrate."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._c("mci_add_sampling_function_device", name.encode(), source.encode(), self.ndim, nproto, _dp(p) if p.size else None, p.size)

    def setTrialMoveDevice(self, name, source, nproto=0, steps=(0.05,), typeEnds=None, change_rate=1.0, params=()):
        """MCI::setTrialMove with a TrialMoveInterface subclass of the caller's as device code
        (NAME_move, and NAME_proto with nproto > 0); len(steps) step sizes over typeEnds."""
        st = np.ascontiguousarray(steps, dtype=np.float64)
        te = None
        if typeEnds is not None:
            te = (C.c_int * len(typeEnds))(*typeEnds)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._c("mci_set_trial_move_device", name.encode(), source.encode(), nproto, st.size, te, float(change_rate), _dp(st),
                _dp(p) if p.size else None, p.size)

    def buildDeviceFunctors(self): self._c("mci_build_device_functors")

    def popSamplingFunction(self): self._c("mci_pop_sampling_function")
    def clearSamplingFunctions(self): self._c("mci_clear_sampling_functions")

    def addObservable(self, kind, params=(), blocksize=1, nskip=1, flag_equil=None, flag_correlated=None, estimator=None):
        """MCI::addObservable overloads (MCIntegrator.hpp:144-170): estimator (EstimatorType) wins;
        else flag_correlated picks Correlated/Uncorrelated (selectEstimatorType, Factories.hpp:61-71)."""
        if flag_equil is None:
            flag_equil = blocksize > 0
        if estimator is None:
            if flag_correlated is None:
                flag_correlated = blocksize == 1
            flag_error = blocksize > 0
            if flag_correlated and not flag_error:
                raise MciError(K.E_INVALID_ARGUMENT, "[selectEstimatorType] Error calculation is set off, but correlated error estimation is set on.")
            estimator = K.ESTIM_CORRELATED if flag_correlated else (K.ESTIM_UNCORRELATED if flag_error else K.ESTIM_NOOP)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._c("mci_add_observable", kind, self.ndim, _dp(p) if p.size else None, p.size, blocksize, nskip,
                int(bool(flag_equil)), estimator)

    def addObservableDevice(self, name, source, nobs, params=(), blocksize=1, nskip=1, flag_equil=None, flag_correlated=None, estimator=None):
        """An observable of the caller's as device code (observableFunction as NAME_observable), with
        addObservable's accumulator and estimator arguments."""
        if flag_equil is None:
            flag_equil = blocksize > 0
        if estimator is None:
            if flag_correlated is None:
                flag_correlated = blocksize == 1
            flag_error = blocksize > 0
            if flag_correlated and not flag_error:
                raise MciError(K.E_INVALID_ARGUMENT, "[selectEstimatorType] Error calculation is set off, but correlated error estimation is set on.")
            estimator = K.ESTIM_CORRELATED if flag_correlated else (K.ESTIM_UNCORRELATED if flag_error else K.ESTIM_NOOP)
        p = np.ascontiguousarray(params, dtype=np.float64)
        self._c("mci_add_observable_device", name.encode(), source.encode(), self.ndim, nobs, _dp(p) if p.size else None, p.size,
                blocksize, nskip, int(bool(flag_equil)), estimator)

    def popObservable(self): self._c("mci_pop_observable")
    def clearObservables(self): self._c("mci_clear_observables")

    # --- getters ---
    def getNDim(self): return self.ndim

    def getX(self):
        a = np.zeros(self.nwalkers * self.ndim)
        self._c("mci_get_x", _dp(a), a.size)
        return a.reshape(self.nwalkers, self.ndim)

    def getMRT2Step(self, i=0):
        v = _D()
        self._c("mci_get_mrt2step", i, C.byref(v))
        return v.value

    def getNObsDim(self):
        v = _I()
        self._c("mci_get_nobsdim", C.byref(v))
        return v.value

    def getAcceptanceRate(self):
        a, r, rate = _I64(), _I64(), _D()
        self._c("mci_get_acceptance", C.byref(a), C.byref(r), C.byref(rate))
        return rate.value

    def getAcceptedRejected(self):
        a, r, rate = _I64(), _I64(), _D()
        self._c("mci_get_acceptance", C.byref(a), C.byref(r), C.byref(rate))
        return a.value, r.value

    # --- checkpoint and resume of the walkers ---
    def getState(self):
        """positions, step sizes, random streams and saved normals of every walker, as bytes"""
        n = _I64()
        self._c("mci_get_state_size", C.byref(n))
        buf = (C.c_char * n.value)()
        self._c("mci_get_state", buf, n.value)
        return bytes(buf)

    def setState(self, state):
        buf = C.create_string_buffer(bytes(state), len(state))
        self._c("mci_set_state", buf, len(state))

    def getTimings(self):
        """[calibration, decorrelation, sampling, estimation, total] of the last integrate, ms"""
        t = np.zeros(5)
        self._c("mci_get_timings", _dp(t), 5)
        return t

    # --- integrate (MCIntegrator.cpp:43-82) ---
    def integrate(self, Nmc, doFindMRT2step=True, doDecorrelation=True):
        nd = max(self.getNObsDim(), 1)
        avg, err = np.zeros(nd), np.zeros(nd)
        self._c("mci_integrate", int(Nmc), _dp(avg), _dp(err), int(doFindMRT2step), int(doDecorrelation))
        return avg, err

    def walkerResults(self):
        nd = self.getNObsDim()
        a, e = np.zeros(self.nwalkers * nd), np.zeros(self.nwalkers * nd)
        self._c("mci_get_walker_results", _dp(a), _dp(e), a.size)
        return a.reshape(self.nwalkers, nd), e.reshape(self.nwalkers, nd)
