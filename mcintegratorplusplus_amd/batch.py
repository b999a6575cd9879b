"""Batched Gaussian-mixture integrals (mci_batch_* of include/mci_amd.h) from Python.

One BatchIntegrals object is many independent integrals, each what one reference MCI object
with a mixture sampling function and the |x|^2 observable computes (src/MCIntegrator.cpp:43-82),
run as one submission on the GPU.
"""
import ctypes as C

import numpy as np

from .mci import _check, _lib as _mci_lib

_P, _I, _I64, _U64, _D = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_DP, _IP, _I64P = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)

_SIGS = {
    "mci_batch_create": [_I, _IP, _I64P, _IP, _DP, _DP, _DP, C.POINTER(_P)],
    "mci_batch_destroy": [_P], "mci_batch_set_seed": [_P, _U64], "mci_batch_set_blocks": [_P, _I],
    "mci_batch_set_min_walkers": [_P, _I], "mci_batch_set_error_mode": [_P, _I],
    "mci_batch_set_calibration": [_P, _I, _D], "mci_batch_set_ndecorrelation_steps": [_P, _I64],
    "mci_batch_set_plan": [_P, _I64P], "mci_batch_autotune": [_P, _I],
    "mci_batch_get_plan": [_P, _I64P, _I64P, _I64P], "mci_batch_get_nwalkers": [_P, _I64P],
    "mci_batch_run": [_P, _DP, _DP], "mci_batch_set_graph": [_P, _I],
    "mci_batch_get_steps": [_P, _DP, _DP], "mci_batch_get_acceptance": [_P, _DP],
    "mci_batch_get_walker_results": [_P, _DP, _DP, _I64], "mci_batch_get_block_means": [_P, _DP, _I64],
    "mci_batch_get_timings": [_P, _DP, _I],
}
_bound = None


def _lib():
    global _bound
    if _bound is None:
        lib = _mci_lib()
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _I
        _bound = lib
    return _bound


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class BatchIntegrals:
    """Integrals i = 0..n-1: dimension ndim[i], nmc[i] Metropolis steps of work, and an isotropic
    Gaussian mixture with weights[i] (K_i values), widths[i] (K_i) and means[i] (K_i x ndim[i])."""

    def __init__(self, ndim, nmc, weights, widths, means):
        self.ndim = np.ascontiguousarray(ndim, dtype=np.int32)
        self.nmc = np.ascontiguousarray(nmc, dtype=np.int64)
        self.n = len(self.ndim)
        self.ncomp = np.array([len(w) for w in weights], dtype=np.int32)
        self._w = np.ascontiguousarray(np.concatenate([np.asarray(w, float) for w in weights]))
        self._s = np.ascontiguousarray(np.concatenate([np.asarray(s, float) for s in widths]))
        self._mu = np.ascontiguousarray(np.concatenate([np.asarray(m, float).ravel() for m in means]))
        h = _P()
        _check(_lib().mci_batch_create(self.n, _ptr(self.ndim, C.c_int), _ptr(self.nmc, C.c_int64),
                                       _ptr(self.ncomp, C.c_int), _ptr(self._w, C.c_double),
                                       _ptr(self._s, C.c_double), _ptr(self._mu, C.c_double), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib().mci_batch_destroy(h)
            self._h = None

    def setSeed(self, seed):
        _check(_lib().mci_batch_set_seed(self._h, seed))

    def setBlocks(self, nblocks):
        _check(_lib().mci_batch_set_blocks(self._h, nblocks))

    def setMinWalkers(self, rmin):
        _check(_lib().mci_batch_set_min_walkers(self._h, rmin))

    def setErrorMode(self, mode):
        _check(_lib().mci_batch_set_error_mode(self._h, mode))

    def setCalibration(self, nfind=-50, target=0.5):
        _check(_lib().mci_batch_set_calibration(self._h, nfind, target))

    def setNdecorrelationSteps(self, n):
        _check(_lib().mci_batch_set_ndecorrelation_steps(self._h, n))

    def setPlan(self, steps_per_walker=None):
        if steps_per_walker is None:
            _check(_lib().mci_batch_set_plan(self._h, None))
        else:
            t = np.ascontiguousarray(steps_per_walker, dtype=np.int64)
            _check(_lib().mci_batch_set_plan(self._h, _ptr(t, C.c_int64)))

    def setGraph(self, on):
        _check(_lib().mci_batch_set_graph(self._h, 1 if on else 0))

    def autotune(self, measure=0):
        _check(_lib().mci_batch_autotune(self._h, measure))

    def plan(self):
        """(walkers R_i, steps per walker T_i, first walker of each integral)"""
        R, T, off = (np.zeros(self.n, np.int64) for _ in range(3))
        _check(_lib().mci_batch_get_plan(self._h, _ptr(R, C.c_int64), _ptr(T, C.c_int64), _ptr(off, C.c_int64)))
        return R, T, off

    def nwalkers(self):
        n = C.c_int64()
        _check(_lib().mci_batch_get_nwalkers(self._h, C.byref(n)))
        return n.value

    def run(self):
        a, e = np.zeros(self.n), np.zeros(self.n)
        _check(_lib().mci_batch_run(self._h, _ptr(a, C.c_double), _ptr(e, C.c_double)))
        return a, e

    def steps(self):
        s, r = np.zeros(self.n), np.zeros(self.n)
        _check(_lib().mci_batch_get_steps(self._h, _ptr(s, C.c_double), _ptr(r, C.c_double)))
        return s, r

    def acceptance(self):
        r = np.zeros(self.n)
        _check(_lib().mci_batch_get_acceptance(self._h, _ptr(r, C.c_double)))
        return r

    def walkerResults(self):
        W = self.nwalkers()
        a, e = np.zeros(W), np.zeros(W)
        _check(_lib().mci_batch_get_walker_results(self._h, _ptr(a, C.c_double), _ptr(e, C.c_double), W))
        return a, e

    def blockMeans(self, nblocks):
        W = self.nwalkers()
        d = np.zeros(nblocks * W)
        _check(_lib().mci_batch_get_block_means(self._h, _ptr(d, C.c_double), d.size))
        return d.reshape(nblocks, W)

    def timings(self):
        ms = np.zeros(4)
        _check(_lib().mci_batch_get_timings(self._h, _ptr(ms, C.c_double), 4))
        return ms


def mixture_workload(n=4096, seed=4096, dmax=128, nmc_range=(1e3, 1e5), kmax=4):
    """The batched benchmark set: n integrals with d = round(exp(U(0, ln dmax))), nmc =
    round(exp(U(ln a, ln b))), K ~ U{1..kmax} isotropic components with means N(0, 2^2 I),
    widths U(0.3, 1.5) and Dirichlet(1) weights.  Returns (ndim, nmc, weights, widths, means,
    exact) with exact[i] = sum_k w_k (|mu_k|^2 + d s_k^2) = E|x|^2."""
    rng = np.random.default_rng(seed)
    ndim = np.clip(np.rint(np.exp(rng.uniform(0.0, np.log(dmax), n))), 1, dmax).astype(np.int32)
    nmc = np.rint(np.exp(rng.uniform(np.log(nmc_range[0]), np.log(nmc_range[1]), n))).astype(np.int64)
    weights, widths, means, exact = [], [], [], np.zeros(n)
    for i in range(n):
        K = int(rng.integers(1, kmax + 1))
        w = rng.dirichlet(np.ones(K))
        s = rng.uniform(0.3, 1.5, K)
        mu = rng.normal(0.0, 2.0, (K, ndim[i]))
        weights.append(w)
        widths.append(s)
        means.append(mu)
        exact[i] = float(np.sum(w * (np.sum(mu * mu, axis=1) + ndim[i] * s * s)))
    return ndim, nmc, weights, widths, means, exact
