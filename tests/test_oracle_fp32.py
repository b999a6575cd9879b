"""The oracle's FP32 precision chain (MCI_PREC_FP32; no reference counterpart: the reference is FP64
only) pinned by what the mathematics fixes:

- Coupling: the FP32 chain draws every proposal and uniform from the same Philox words as the FP64
  chain, on a 24-bit instead of a 32/53-bit grid, so the two make the same decisions except where u
  lies within about 1e-6 of the acceptance ratio.  Walker by walker, their acceptance counts agree
  and their positions differ by FP32 rounding only.  A wrong operand in the FP32 step breaks this.
- Accumulation: the FP32 dwell-time Kahan block means against the same trajectory's block means
  summed step by step in FP64 (the oracle's shadow, i.e. the reference's BlockAccumulator applied
  to the FP32 chain): within 1e-6 relative, the precision bar; a dwell weight off by one step, or a
  plain FP32 sum, is far outside it.
- Physics: the analytic moments sigma_i^2 and prod (1 + 2 sigma_i^2 / c)^(-1/2) within 3 error bars.
"""
import ctypes as C

import numpy as np
import pytest

from mcitest import (COUPLING_SHARED, ESTIM_MJ, ESTIM_UNCORR, MciError, OBS_EXP_X2, OBS_X2, OBS_X2_SUM, PDF_GAUSS_DIAG,
                     PREC_FP32, PREC_FP64, RNG_MT, RNG_PHILOX, chain_setup, load)

_DP = C.POINTER(C.c_double)


def diag_cfg(sig, obs, seed=99, ndecorr=0, step=None):
    a = 0.5 / np.asarray(sig) ** 2
    cfg = dict(ndim=len(sig), seed=seed, ndecorr=ndecorr, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, a)], obs=obs)
    if step is not None:
        cfg["step"] = step
    return cfg


def oracle_chain(cfg, W, prec, shadow=False):
    c = chain_setup("orc", cfg, nwalkers=W)
    c.setRng(RNG_PHILOX)
    c.setPrecision(prec)
    if shadow:
        load("orc").orc_set_fp32_shadow.argtypes = [C.c_void_p, C.c_int]
        c.lib.orc_set_fp32_shadow(c.ctx, 1)
    return c


def shadow_data(c, i, nobs):
    lib = load("orc")
    lib.orc_get_shadow_data.argtypes = [C.c_void_p, C.c_int, C.c_int64, _DP, C.c_int64, C.POINTER(C.c_int64)]
    lib.orc_get_accu_data.argtypes = [C.c_void_p, C.c_int, C.c_int64, _DP, C.c_int64, C.POINTER(C.c_int64)]
    out, sh = [], []
    for w in range(c.W):
        ns = C.c_int64()
        buf = np.zeros(1 << 16)
        assert lib.orc_get_accu_data(c.ctx, i, w, buf.ctypes.data_as(_DP), buf.size, C.byref(ns)) == 0
        out.append(buf[:ns.value * nobs].copy())
        assert lib.orc_get_shadow_data(c.ctx, i, w, buf.ctypes.data_as(_DP), buf.size, C.byref(ns)) == 0
        sh.append(buf[:ns.value * nobs].copy())
    return np.array(out), np.array(sh)


@pytest.mark.parametrize("ndim", [5, 8])
def test_fp32_chain_is_coupled_to_the_fp64_chain(ndim):
    sig = np.linspace(0.5, 2.0, ndim)
    cfg = diag_cfg(sig, [(OBS_X2, [], 50, 1, 1, ESTIM_UNCORR)], step=2.2, seed=5)
    W, nmc = 64, 4000
    res = {}
    for prec in (PREC_FP64, PREC_FP32):
        c = oracle_chain(cfg, W, prec)
        c.setX(np.zeros(ndim))
        c.integrate(nmc, False, False)
        res[prec] = (c.getX(), c.walkerAcceptance()[0])
    same = res[PREC_FP64][1] == res[PREC_FP32][1]
    assert same.sum() >= W - 3, same.sum()
    assert np.all(res[PREC_FP32][1] > 0.2 * nmc)
    dx = np.abs(res[PREC_FP64][0] - res[PREC_FP32][0])[same]
    assert dx.max() < 1e-4, dx.max()
    # FP32-representable positions (the chain keeps its state in FP32)
    x32 = res[PREC_FP32][0]
    assert np.array_equal(x32, x32.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("obs_set,blocksize,nmc", [
    ("x2_exp", 160, 10240),    # the precision-form configuration's blocks
    ("x2_sum", 0, 1 << 17),    # Simple accumulator: one block of 131,072 steps (dwell weights up to 2^17)
    ("all", 1024, 8192),
])
def test_fp32_block_means_within_1e6_of_fp64_sums(obs_set, blocksize, nmc):
    d = 64 if obs_set == "x2_exp" else 16
    sig = np.random.default_rng(7).uniform(0.5, 2.0, d)
    est = ESTIM_MJ if blocksize else 0
    obs = {"x2_exp": [(OBS_X2, [], blocksize, 1, 1, est), (OBS_EXP_X2, [float(d)], blocksize, 1, 1, est)],
           "x2_sum": [(OBS_X2, [], blocksize, 1, 0, est), (OBS_X2_SUM, [], blocksize, 1, 0, est)],
           "all": [(OBS_X2, [], blocksize, 1, 1, est), (OBS_X2_SUM, [3.0], blocksize, 1, 1, est),
                   (OBS_EXP_X2, [], blocksize, 1, 1, est)]}[obs_set]
    W = 8
    c = oracle_chain(diag_cfg(sig, obs, ndecorr=500), W, PREC_FP32, shadow=True)
    c.setNfindMRT2Iterations(-50)
    c.setX(sig)  # a typical point; calibration and decorrelation run first
    c.integrate(nmc)
    worst = 0.0
    for i, (kind, params, *_r) in enumerate(obs):
        nobs = d if kind == OBS_X2 else 1
        got, ref = shadow_data(c, i, nobs)
        got, ref = got.reshape(W, -1, nobs), ref.reshape(W, -1, nobs)
        # each walker's estimate of each component (the mean of its block means): the 1e-6 bar
        rel = np.abs(got.mean(axis=1) - ref.mean(axis=1)) / np.abs(ref.mean(axis=1))
        worst = max(worst, rel.max())
        assert rel.max() <= 1e-6, (kind, rel.max())
        # each block mean, relative to the component's scale (a block mean near 0 has no relative
        # precision: its FP32 terms B x^2 cancel): 1e-5, where a dwell weight off by one step
        # gives about 1/B
        scale = np.abs(ref).mean(axis=1, keepdims=True)
        assert np.max(np.abs(got - ref) / scale) <= 1e-5, (kind, np.max(np.abs(got - ref) / scale))
    assert worst > 0  # the FP32 sums do round


def test_fp32_simple_block_of_2_to_20_steps():
    """One Simple block of 2^20 steps: dwell weights far above 2^16, still within the bar."""
    d = 4
    sig = np.array([0.5, 1.0, 1.5, 2.0])
    c = oracle_chain(diag_cfg(sig, [(OBS_X2, [], 0, 1, 0, 0)], step=2.0), 1, PREC_FP32, shadow=True)
    c.setX(sig)
    c.integrate(1 << 20, False, False)
    got, ref = shadow_data(c, 0, d)
    assert np.all(np.abs(got - ref) / ref <= 1e-6), np.abs(got - ref) / ref
    assert not np.array_equal(got, ref)


def test_fp32_physics_within_three_error_bars():
    sig = np.array([0.5, 0.9, 1.3, 2.0])
    d = 4
    obs = [(OBS_X2, [], 1024, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [4.0], 1024, 1, 1, ESTIM_MJ)]
    c = oracle_chain(diag_cfg(sig, obs, ndecorr=2000), 32, PREC_FP32)
    c.setCoupling(COUPLING_SHARED)
    c.setX(np.zeros(d))
    a, e = c.integrate(1024 * 64)
    exact = np.concatenate([sig ** 2, [np.prod((1 + 2 * sig ** 2 / 4.0) ** -0.5)]])
    z = (a - exact) / e
    assert np.all(np.abs(z) < 3), z


def test_fp32_refuses_what_the_library_refuses():
    sig = np.ones(4)
    for rng, obs in ((RNG_MT, [(OBS_X2, [], 16, 1, 1, ESTIM_MJ)]),
                     (RNG_PHILOX, [(OBS_X2, [], 1, 1, 1, 2)]),            # Full accumulator
                     (RNG_PHILOX, [(OBS_X2, [], 16, 2, 1, ESTIM_MJ)])):   # nskip 2
        c = chain_setup("orc", diag_cfg(sig, obs, ndecorr=0), nwalkers=2)
        c.setRng(rng)
        c.setPrecision(PREC_FP32)
        with pytest.raises(MciError):
            c.integrate(64, False, False)
