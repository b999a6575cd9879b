"""Closed forms and invariants that pin the oracle's estimators and chain independently of the
reference binary (they hold for any correct implementation)."""
import numpy as np
import pytest
from scipy import stats

from mcitest import (Chain, MciError, E_INVALID, E_DOMAIN, ESTIM_CORR, ESTIM_FC, ESTIM_MJ, ESTIM_NOOP,
                     ESTIM_UNCORR, OBS_GAUSS_X2, OBS_X, OBS_X2, OBS_X2_SUM, PDF_GAUSS_DIAG, PDF_GAUSS_ISO, chain_setup, estimate)


def test_chi2_table_is_the_095_quantile():
    # MJBlocker.cpp:6-13 holds chi^2_{0.95}(df) for df = 1..64 to 6 decimals
    import re
    src = open(__import__("os").path.join(__import__("mcitest").ROOT, "oracle", "mci_oracle.c")).read()
    body = src[src.index("chi2_q95[64]"):]
    body = body[body.index("{") + 1:body.index("};")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    vals = [float(v) for v in body.split(",") if v.strip()]
    assert len(vals) == 64
    assert np.allclose(vals, stats.chi2.ppf(0.95, np.arange(1, 65)), atol=6e-7, rtol=0)


def test_uncorrelated_estimator_textbook():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(257, 4)) * [1, 2, 3, 4] + [0, 1, -1, 5]
    a, e = estimate("orc", ESTIM_UNCORR, x)
    assert np.allclose(a, x.mean(axis=0), rtol=1e-14)
    # sqrt((<x^2> - <x>^2)/(n-1)) = std(ddof=0)/sqrt(n-1)
    assert np.allclose(e, x.std(axis=0) / np.sqrt(len(x) - 1), rtol=1e-10)
    a1, e1 = estimate("orc", ESTIM_UNCORR, x[:, 2])
    assert np.isclose(a1[0], a[2], rtol=1e-14) and np.isclose(e1[0], e[2], rtol=1e-10)


def test_noop_copies_first_row():
    x = np.arange(12.0).reshape(4, 3)
    a, e = estimate("orc", ESTIM_NOOP, x)
    assert np.array_equal(a, x[0]) and np.array_equal(e, np.zeros(3))


def test_mjblocker_invariants():
    rng = np.random.default_rng(5)
    # n = 2: one level, err^2 = var/2 with var = mean of squared deviations
    x = np.array([[1.0], [3.0]])
    a, e = estimate("orc", ESTIM_MJ, x)
    assert a[0] == 2.0 and np.isclose(e[0], np.sqrt(1.0 / 2.0))
    # constant data: the reference reads out of bounds here; the defined result is err = 0
    a, e = estimate("orc", ESTIM_MJ, np.full((64, 2), 1.25))
    assert np.array_equal(a, [1.25, 1.25]) and np.array_equal(e, [0.0, 0.0])
    # i.i.d. data: blocking finds no correlation, error ~ uncorrelated error
    x = rng.normal(size=(2 ** 14, 3))
    a, e = estimate("orc", ESTIM_MJ, x)
    au, eu = estimate("orc", ESTIM_UNCORR, x)
    assert np.allclose(a, au, rtol=1e-12, atol=1e-15)
    assert np.all(np.abs(e / eu - 1) < 0.25)
    # AR(1) with phi = 0.9: integrated autocorrelation time (1+phi)/(1-phi) = 19
    n = 2 ** 16
    y = np.zeros(n)
    eps = rng.normal(size=n)
    for i in range(1, n):
        y[i] = 0.9 * y[i - 1] + eps[i]
    _, e = estimate("orc", ESTIM_MJ, y)
    _, eu = estimate("orc", ESTIM_UNCORR, y)
    assert 0.6 < (e[0] / eu[0]) ** 2 / 19.0 < 1.5


def test_mjblocker_needs_power_of_two():
    with pytest.raises(MciError) as ex:
        estimate("orc", ESTIM_MJ, np.zeros((48, 1)))
    assert ex.value.code == E_INVALID


def test_fcblocker_needs_50_and_correlated_dispatch():
    with pytest.raises(MciError) as ex:
        estimate("orc", ESTIM_FC, np.random.default_rng(0).normal(size=(49, 2)))
    assert ex.value.code == E_INVALID
    x = np.random.default_rng(1).normal(size=(100, 2))
    assert np.array_equal(np.array(estimate("orc", ESTIM_CORR, x)), np.array(estimate("orc", ESTIM_FC, x)))
    x = np.random.default_rng(1).normal(size=(128, 2))
    assert np.array_equal(np.array(estimate("orc", ESTIM_CORR, x)), np.array(estimate("orc", ESTIM_MJ, x)))


def test_block_means_average_to_trace_mean():
    # mean of block means == mean of the full trace (test/ut1/main.cpp:341-345)
    cfg = dict(ndim=3, seed=11, pdfs=[(PDF_GAUSS_ISO, [])],
               obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR), (OBS_X2, [], 1, 1, 0, ESTIM_UNCORR), (OBS_X2, [], 0, 1, 0, ESTIM_NOOP)])
    c = chain_setup("orc", cfg)
    a, _ = c.integrate(4096, True, False)
    assert np.allclose(a[0:3], a[3:6], rtol=1e-12) and np.allclose(a[3:6], a[6:9], rtol=1e-12)


def test_gaussian_physics_within_error_bars():
    # x_i^2 under exp(-sum a_i x_i^2) has mean 1/(2 a_i)
    a_i = np.array([0.5, 1.0, 2.0, 0.25])
    cfg = dict(ndim=4, seed=2, move=(0, 1, 1, None), ndecorr=1000, pdfs=[(PDF_GAUSS_DIAG, a_i)],
               obs=[(OBS_X2, [], 64, 1, 1, ESTIM_MJ)])
    c = chain_setup("orc", cfg, nwalkers=16)
    avg, err = c.integrate(64 * 256)
    assert np.all(np.abs(avg - 0.5 / a_i) < 3 * err)


def test_uniform_sampling_volume_integral():
    # no sampling function: integrate returns <f> * V over the box (MCIntegrator.cpp:190-198).
    # int exp(-r^2) r^2 d^3x / (3 pi^1.5) = 1/2 over R^3; [-5, 5]^3 misses < 1e-9 of it
    c = chain_setup("orc", dict(ndim=3, seed=17, irange=([-5.] * 3, [5.] * 3),
                                obs=[(OBS_GAUSS_X2, [], 1, 1, 0, ESTIM_UNCORR)]), nwalkers=16)
    avg, err = c.integrate(20000, False, False)
    assert abs(avg[0] - 0.5) < 3 * err[0] and err[0] < 0.01


def test_errors_map_to_reference_exception_types():
    c = Chain("orc", 3)
    with pytest.raises(MciError) as ex:  # infinite domain without sampling function
        c.addObservable(OBS_X2_SUM, [3.0], 1, 1, False, ESTIM_UNCORR)
        c.integrate(100)
    assert ex.value.code == E_DOMAIN
    with pytest.raises(MciError) as ex:  # dimension mismatch
        c.call("add_sampling_function", PDF_GAUSS_ISO, 2, None, 0)
    assert ex.value.code == E_INVALID
    with pytest.raises(MciError) as ex:  # flag_equil with Noop
        c.addObservable(OBS_X, [], 0, 1, True, ESTIM_NOOP)
    assert ex.value.code == E_INVALID
    c.clearObservables()
    c.addSamplingFunction(PDF_GAUSS_ISO)
    c.addObservable(OBS_X, [], 7, 1, False, ESTIM_UNCORR)
    with pytest.raises(MciError) as ex:  # block size does not divide
        c.integrate(100, False, False)
    assert ex.value.code == E_INVALID
    with pytest.raises(MciError) as ex:  # veclen does not divide ndim
        c.setTrialMove(0, 2, 1, None)
    assert ex.value.code == E_INVALID
