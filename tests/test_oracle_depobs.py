"""Dependent observables (DependentObservableInterface.hpp) in the oracle: the reference's own chains
with them are in tests/golden/chains.json (depobs_*, run by test_oracle_golden.py); here the physics
they must give and the rules the C-ABI enforces.

For p(x) = exp(-sum a_i x_i^2) the proto values a_i x_i^2 have mean 1/2 each, the log value
-sum a_i x_i^2 has mean -d/2, and the square of the observable x_i^2 has mean E[x_i^4] = 3 sigma_i^4
with sigma_i^2 = 1/(2 a_i)."""
import numpy as np
import pytest

from mcitest import (ESTIM_MJ, ESTIM_UNCORR, MciError, OBS_DEP_SQUARE, OBS_PDF_LOG, OBS_PDF_PROTO, OBS_X, OBS_X2, PDF_GAUSS_DIAG,
                     PDF_GAUSS_ISO, E_INVALID, Chain, chain_setup)


def test_dependent_observables_physics():
    a = np.array([0.5, 1.0, 2.0])
    sig2 = 0.5 / a
    cfg = dict(ndim=3, seed=77, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, a)],
               obs=[(OBS_X2, [], 256, 1, 1, ESTIM_MJ), (OBS_PDF_PROTO, [0], 256, 2, 1, ESTIM_MJ),
                    (OBS_PDF_LOG, [0], 256, 2, 1, ESTIM_MJ), (OBS_DEP_SQUARE, [0], 256, 2, 1, ESTIM_MJ)])
    c = chain_setup("orc", cfg, nwalkers=16)
    avg, err = c.integrate(256 * 512)
    exact = np.concatenate([sig2, [0.5] * 3, [-1.5], 3 * sig2 ** 2])
    z = (avg - exact) / err
    assert np.all(np.abs(z) < 3.5), z


def test_dependency_rules():
    c = Chain("orc", 3)
    c.addSamplingFunction(PDF_GAUSS_ISO)
    c.addObservable(OBS_X, [], 4, 2, 0, ESTIM_UNCORR)
    for params, nskip in (([1], 2), ([3], 2), ([0], 3), ([], 2), ([0.5], 2)):  # not earlier, not synced, no index
        with pytest.raises(MciError) as ei:
            c.addObservable(OBS_DEP_SQUARE, params, 4, nskip, 0, ESTIM_UNCORR)
        assert ei.value.code == E_INVALID
    c.addObservable(OBS_DEP_SQUARE, [0], 4, 4, 0, ESTIM_UNCORR)   # nskip 4 is a multiple of 2
    c.addObservable(OBS_PDF_LOG, [1], 4, 1, 0, ESTIM_UNCORR)      # a second sampling function, not there yet
    with pytest.raises(MciError):
        c.integrate(64, False, False)
    c.addSamplingFunction(PDF_GAUSS_ISO)
    a, e = c.integrate(64, False, False)
    assert len(a) == 3 + 3 + 1
    c.addObservable(OBS_PDF_PROTO, [0], 4, 1, 0, ESTIM_UNCORR)    # GAUSS_ISO has one proto value, not ndim
    with pytest.raises(MciError):
        c.integrate(64, False, False)


def test_equilibration_needs_the_dependency_too():
    c = Chain("orc", 2)
    c.addSamplingFunction(PDF_GAUSS_ISO)
    c.addObservable(OBS_X2, [], 1, 1, 0, 2)         # not an equilibration observable
    c.addObservable(OBS_DEP_SQUARE, [0], 1, 1, 1, 2)  # an equilibration observable that depends on it
    with pytest.raises(MciError):
        c.integrate(128, True, True)
    c.setNdecorrelationSteps(100)                   # fixed decorrelation: no equilibration container
    c.integrate(128, True, True)
