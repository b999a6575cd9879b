"""Checkpoint and resume of the walkers (mci_get_state / mci_set_state): a run continued from a saved
state, in a fresh context or in the same one after it moved on, is bit-identical to the run that
continued without interruption, in both random-stream modes, with a move distribution whose saved
normals are part of the state (Gaussian steps) and with the specialised kernel's Philox counters."""
import numpy as np
import pytest

from mcitest import (COUPLING_SHARED, ESTIM_MJ, MciError, OBS_X2, OBS_X2_SUM, PDF_GAUSS_DIAG, RNG_MT, RNG_PHILOX, SRRD_UNIFORM,
                     chain_setup)

pytestmark = pytest.mark.gpu


def cfg(srrd):
    a = np.array([0.5, 1.0, 2.0, 0.25, 1.5, 0.8])
    return dict(ndim=6, seed=123, ndecorr=200, move=(srrd, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, a)],
                obs=[(OBS_X2, [], 16, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 16, 1, 1, ESTIM_MJ)])


def snapshot(c):
    a, e = c.integrate(16 * 64)
    return (a, e, *c.walkerResults(), c.getX(), c.getWalkerSteps(), *c.walkerAcceptance())


@pytest.mark.parametrize("rng,srrd", [(RNG_MT, 1), (RNG_PHILOX, 1), (RNG_PHILOX, SRRD_UNIFORM)])
def test_resume_from_a_checkpoint(rng, srrd):
    print("checkpoint", rng, srrd, flush=True)
    W = 100
    c = chain_setup("mci", cfg(srrd), nwalkers=W)
    c.setRng(rng)
    c.setCoupling(COUPLING_SHARED)
    c.integrate(16 * 64)
    state = c.getState()
    cont = snapshot(c)                       # the uninterrupted continuation
    fresh = chain_setup("mci", cfg(srrd), nwalkers=W)
    fresh.setRng(rng)
    fresh.setCoupling(COUPLING_SHARED)
    fresh.setState(state)
    again = snapshot(fresh)                  # resumed in a new context
    for k in range(len(cont)):
        assert np.array_equal(cont[k], again[k]), k
    c.setState(state)                        # and in the old one, which had moved on
    back = snapshot(c)
    for k in range(len(cont)):
        assert np.array_equal(cont[k], back[k]), k


def test_state_must_match_the_context():
    c = chain_setup("mci", cfg(SRRD_UNIFORM), nwalkers=64)
    state = c.getState()
    other = chain_setup("mci", cfg(SRRD_UNIFORM), nwalkers=65)
    with pytest.raises(MciError):
        other.setState(state)
    other2 = chain_setup("mci", cfg(SRRD_UNIFORM), nwalkers=64)
    other2.setRng(RNG_PHILOX)
    with pytest.raises(MciError):
        other2.setState(state)
    with pytest.raises(MciError):
        c.setState(state[:-8])
