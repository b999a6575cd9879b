"""The FP32 precision form (MCI_PREC_FP32) on the GPU against the oracle's FP32 chain: the same
decisions, positions, counters and block means bit for bit (the kernel uses only correctly rounded
FP32 operations and the library's own exp), at ragged walker counts, with Block and Simple
accumulators, long blocks (|x|^2 refreshed every 1,024 steps), a non-power-of-two coordinate count
(Lemire index with rejections) and both couplings.  Then the precision-form configuration at full
size: 262,144 walkers x 10,240 steps at d = 64, within 3 error bars of the analytic values, its
error bars calibrated, and each estimate's error against FP64 sums of the same chain."""
import ctypes as C

import numpy as np
import pytest

import golden_io as G
from mcitest import (COUPLING_INDEPENDENT, COUPLING_SHARED, ESTIM_MJ, ESTIM_NOOP, ESTIM_UNCORR, OBS_EXP_X2, OBS_X2, OBS_X2_SUM,
                     PDF_GAUSS_DIAG, PREC_FP32, PREC_FP64, RNG_PHILOX, chain_setup, load)

pytestmark = pytest.mark.gpu
_DP = C.POINTER(C.c_double)


def sigmas(d):
    if d in (16, 64):
        return G.hexv(G.load("sigmas.json")["sigma16" if d == 16 else "sigma64"])
    return np.random.default_rng(d).uniform(0.5, 2.0, d)


def fp32_cfg(d, obs, seed=4242, ndecorr=1000):
    sig = sigmas(d)
    return dict(ndim=d, seed=seed, ndecorr=ndecorr, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, 0.5 / sig ** 2)], obs=obs)


def oracle_data(c, i, nobs, shadow=False):
    lib = load("orc")
    fn = lib.orc_get_shadow_data if shadow else lib.orc_get_accu_data
    fn.argtypes = [C.c_void_p, C.c_int, C.c_int64, _DP, C.c_int64, C.POINTER(C.c_int64)]
    rows = []
    for w in range(c.W):
        ns = C.c_int64()
        buf = np.zeros(1 << 20)
        assert fn(c.ctx, i, w, buf.ctypes.data_as(_DP), buf.size, C.byref(ns)) == 0
        rows.append(buf[:ns.value * nobs].reshape(ns.value, nobs).copy())
    return np.array(rows)


def run_both(cfg, W, nmc, coupling, do_find=True):
    out = {}
    for prefix in ("orc", "mci"):
        c = chain_setup(prefix, cfg, nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(coupling)
        c.setPrecision(PREC_FP32)
        a, e = c.integrate(nmc, do_find, True)
        out[prefix] = (c, a, e, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance())
    return out


@pytest.mark.parametrize("case", ["cfg3_subset", "simple_all", "lemire_long_blocks", "sum_only"])
def test_fp32_chain_bit_exact_with_oracle(case):
    print("fp32 case", case, flush=True)
    if case == "cfg3_subset":   # the precision-form configuration on 200 walkers (last wave: 8 lanes)
        d, W, nmc, coupling = 64, 200, 10240, COUPLING_SHARED
        obs = [(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [64.0], 160, 1, 1, ESTIM_MJ)]
    elif case == "simple_all":  # Simple accumulators over 5,000 steps: four |x|^2 refreshes inside the block
        d, W, nmc, coupling = 16, 130, 5000, COUPLING_INDEPENDENT
        obs = [(OBS_X2, [], 0, 1, 0, ESTIM_NOOP), (OBS_X2_SUM, [3.0], 0, 1, 0, ESTIM_NOOP), (OBS_EXP_X2, [], 0, 1, 0, ESTIM_NOOP)]
    elif case == "lemire_long_blocks":  # d = 5: Lemire index with rejections; blocks of 2,048 steps
        d, W, nmc, coupling = 5, 96, 2048 * 8, COUPLING_SHARED
        obs = [(OBS_X2, [], 2048, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 2048, 1, 1, ESTIM_UNCORR)]
    else:                       # |x|^2 and exp without the per-coordinate records
        d, W, nmc, coupling = 8, 64, 4096, COUPLING_INDEPENDENT
        obs = [(OBS_X2_SUM, [], 256, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [8.0], 256, 1, 1, ESTIM_MJ)]
    out = run_both(fp32_cfg(d, obs), W, nmc, coupling)
    o, g = out["orc"], out["mci"]
    for k in range(1, len(o)):
        assert np.array_equal(o[k], g[k]), k
    for i, (kind, *_r) in enumerate(obs):
        nobs = d if kind == OBS_X2 else 1
        assert np.array_equal(oracle_data(o[0], i, nobs), g[0].observableData(i, nobs)), (case, kind)


def test_fp32_refuses_a_simple_block_beyond_2_to_31():
    """The host refuses before any kernel runs (the FP32 kernel's block counters are 32-bit)."""
    c = chain_setup("mci", fp32_cfg(4, [(OBS_X2, [], 0, 1, 0, ESTIM_NOOP)], ndecorr=0), nwalkers=64)
    c.setRng(RNG_PHILOX)
    c.setPrecision(PREC_FP32)
    with pytest.raises(Exception) as ei:
        c.integrate(2 ** 31, False, False)
    assert "2^31" in str(ei.value)


def test_fp32_error_against_fp64_sums_of_the_same_chain():
    """Walkers 0..255 of the precision-form configuration (fixed step, no calibration, so that a
    256-walker oracle reproduces them): the device's FP32 block means against the oracle's FP64
    step-by-step sums of the same FP32 trajectory.  Each walker's estimate, and the 256 walkers'
    combined estimate, stay within 1e-6 relative."""
    d, W, nmc = 64, 256, 10240
    obs = [(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [64.0], 160, 1, 1, ESTIM_MJ)]
    cfg = fp32_cfg(d, obs)
    cfg["step"] = 2.05
    res = {}
    for prefix in ("orc", "mci"):
        c = chain_setup(prefix, cfg, nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setPrecision(PREC_FP32)
        if prefix == "orc":
            load("orc").orc_set_fp32_shadow.argtypes = [C.c_void_p, C.c_int]
            c.lib.orc_set_fp32_shadow(c.ctx, 1)
        a, e = c.integrate(nmc, False, True)
        res[prefix] = c
    worst_walker, worst_comb = 0.0, 0.0
    for i, (kind, *_r) in enumerate(obs):
        nobs = d if kind == OBS_X2 else 1
        dev = res["mci"].observableData(i, nobs)
        assert np.array_equal(dev, oracle_data(res["orc"], i, nobs))
        ref = oracle_data(res["orc"], i, nobs, shadow=True)
        rel_w = np.abs(dev.mean(axis=1) - ref.mean(axis=1)) / np.abs(ref.mean(axis=1))
        rel_c = np.abs(dev.mean(axis=(0, 1)) - ref.mean(axis=(0, 1))) / np.abs(ref.mean(axis=(0, 1)))
        worst_walker, worst_comb = max(worst_walker, rel_w.max()), max(worst_comb, rel_c.max())
    print(f"FP32 vs FP64 sums of the same chain: walker estimates max rel {worst_walker:.3g}, combined max rel {worst_comb:.3g}",
          flush=True)
    assert worst_walker <= 1e-6 and worst_comb <= 1e-6


def cfg3():
    return fp32_cfg(64, [(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [64.0], 160, 1, 1, ESTIM_MJ)])


def test_fp32_precision_form_full_size():
    """The precision-form configuration: d = 64, 262,144 walkers x 10,240 steps, shared calibration,
    FP32 chain with Kahan block sums, FP64 estimators and combine; the FP64 chain driven by the
    same Philox words beside it.  In both precisions at least 64 of the 65 estimates lie within 3
    error bars of their analytic values (sigma_i^2 and prod (1 + 2 sigma_i^2 / 64)^(-1/2); with 65
    independent components one beyond 3 sigma has probability 0.16) and none beyond 3.5, the error
    bars are calibrated (mean z^2 of the 65 components near 1), and FP32 and FP64 agree to a small
    fraction of an error bar.  The error bar is the standard error of the 262,144 walker means (the
    walkers sample independently once the shared calibration is done); the MPIMCI combine of the
    walkers' MJBlocker errors, which the integrate returns, runs low here (64 blocks of 160 steps
    per walker are too few for the blocking to reach the chain's correlation time at d = 64): its
    ratio to the spread is printed and bounded."""
    W, nmc = 262144, 10240
    sig = sigmas(64)
    exact = np.concatenate([sig ** 2, [np.prod((1 + 2 * sig ** 2 / 64) ** -0.5)]])
    res = {}
    for prec in (PREC_FP64, PREC_FP32):
        c = chain_setup("mci", cfg3(), nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(COUPLING_SHARED)
        c.setPrecision(prec)
        a, e = c.integrate(nmc)
        wa, _ = c.walkerResults()
        res[prec] = (a, e, wa.std(axis=0, ddof=1) / np.sqrt(W))
        assert np.allclose(wa.mean(axis=0), a, rtol=1e-12, atol=0)
    for prec, (a, e, es) in res.items():
        z = (a - exact) / es
        print(f"prec {prec}: max |z| {np.max(np.abs(z)):.2f}, mean z^2 {np.mean(z * z):.2f}; "
              f"MJBlocker combine / walker spread: median {np.median(e / es):.3f}", flush=True)
        assert np.all(np.abs(z) < 3.5), np.max(np.abs(z))
        assert np.mean(np.abs(z) < 3) >= 64 / 65
        assert 0.5 < np.mean(z * z) < 1.6
        assert 0.5 < np.median(e / es) < 1.1
    a64, _, e64 = res[PREC_FP64]
    a32 = res[PREC_FP32][0]
    print("FP32 vs FP64 (chains coupled through the same words): max |diff|/err",
          np.max(np.abs(a32 - a64) / e64), "max rel", np.max(np.abs(a32 - a64) / np.abs(a64)), flush=True)
    assert np.all(np.abs(a32 - a64) < 0.25 * e64)
