"""Test helpers for the batched Gaussian-mixture integrals: the oracle's batch runner
(oracle/mci_batch_oracle.c) with the same calling convention as the product's
BatchIntegrals, and small workloads."""
import ctypes as C

import numpy as np

from mcitest import load

_I, _I64, _U64, _D = C.c_int, C.c_int64, C.c_uint64, C.c_double
_DP, _IP, _I64P = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)


def _orc():
    lib = load("orc")
    if not getattr(lib, "_batch_bound", False):
        lib.orc_bexp.argtypes = [C.c_float]
        lib.orc_bexp.restype = C.c_float
        lib.orc_blog.argtypes = [C.c_float]
        lib.orc_blog.restype = C.c_float
        lib.orc_batch_walker_order.argtypes = [_I, _IP, _IP, _I64P, _I64P, _I64P, _I64P, _I64P]
        lib.orc_batch_walker_order.restype = _I
        lib.orc_batch_run.argtypes = [_I, _IP, _I64P, _IP, _DP, _DP, _DP, _U64, _I, _I, _D, _I64, _I, _I64P] + [_DP] * 9
        lib.orc_batch_run.restype = _I
        lib._batch_bound = True
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def flat(weights, widths, means):
    w = np.ascontiguousarray(np.concatenate([np.asarray(v, float) for v in weights]))
    s = np.ascontiguousarray(np.concatenate([np.asarray(v, float) for v in widths]))
    m = np.ascontiguousarray(np.concatenate([np.asarray(v, float).ravel() for v in means]))
    return w, s, m


def orc_walker_order(ndim, ncomp, nmc, T):
    n = len(ndim)
    nd, nc = np.ascontiguousarray(ndim, np.int32), np.ascontiguousarray(ncomp, np.int32)
    nm, t = np.ascontiguousarray(nmc, np.int64), np.ascontiguousarray(T, np.int64)
    R, off, W = np.zeros(n, np.int64), np.zeros(n, np.int64), C.c_int64()
    _orc().orc_batch_walker_order(n, _p(nd, C.c_int), _p(nc, C.c_int), _p(nm, C.c_int64), _p(t, C.c_int64),
                                  _p(R, C.c_int64), _p(off, C.c_int64), C.byref(W))
    return R, off, W.value


def orc_batch(ndim, nmc, weights, widths, means, T, seed=1, nblocks=16, nfind=-50, target=0.5, ndecorr=0, mode=0,
              fp64_shadow=False):
    """The oracle's run of plan T; returns a dict of every output."""
    n = len(ndim)
    nd = np.ascontiguousarray(ndim, np.int32)
    nm = np.ascontiguousarray(nmc, np.int64)
    nc = np.array([len(w) for w in weights], np.int32)
    t = np.ascontiguousarray(T, np.int64)
    w, s, m = flat(weights, widths, means)
    R, off, W = orc_walker_order(nd, nc, nm, t)
    out = dict(avg=np.zeros(n), err=np.zeros(n), steps=np.zeros(n), calib_rate=np.zeros(n), accrate=np.zeros(n),
               wavg=np.zeros(W), werr=np.zeros(W), bmeans=np.zeros(nblocks * W),
               bmeans_fp64=np.zeros(nblocks * W) if fp64_shadow else None)
    st = _orc().orc_batch_run(n, _p(nd, C.c_int), _p(nm, C.c_int64), _p(nc, C.c_int), _p(w, C.c_double),
                              _p(s, C.c_double), _p(m, C.c_double), seed, nblocks, nfind, target, ndecorr, mode,
                              _p(t, C.c_int64), *[_p(out[k], C.c_double) for k in
                                                  ("avg", "err", "steps", "calib_rate", "accrate", "wavg", "werr",
                                                   "bmeans", "bmeans_fp64")])
    assert st == 0
    out["bmeans"] = out["bmeans"].reshape(nblocks, W)
    if fp64_shadow:
        out["bmeans_fp64"] = out["bmeans_fp64"].reshape(nblocks, W)
    out.update(R=R, off=off, W=W)
    return out


def small_mixtures(n=24, seed=7, dmax=48, nmc=(1000, 12000)):
    """A small batch of mixtures for parity tests (every K from 1 to 4, d from 1 to dmax)."""
    rng = np.random.default_rng(seed)
    ndim = np.concatenate([[1, 2, 3, dmax], rng.integers(1, dmax + 1, n - 4)]).astype(np.int32)
    nmcs = np.rint(np.exp(rng.uniform(np.log(nmc[0]), np.log(nmc[1]), n))).astype(np.int64)
    weights, widths, means, exact = [], [], [], np.zeros(n)
    for i in range(n):
        K = 1 + i % 4
        w = rng.dirichlet(np.ones(K))
        s = rng.uniform(0.3, 1.5, K)
        mu = rng.normal(0.0, 2.0, (K, ndim[i]))
        weights.append(w)
        widths.append(s)
        means.append(mu)
        exact[i] = float(np.sum(w * (np.sum(mu * mu, axis=1) + ndim[i] * s * s)))
    return ndim, nmcs, weights, widths, means, exact
