"""The mixed trial move of the oracle (a learned independence draw with probability pmix, else the
move set before; the factor q(x)/q(x') enters the acceptance through the trial-move hook,
TrialMoveInterface.hpp:63-69, MCIntegrator.cpp:343), pinned by what the mathematics fixes:

- a proposal equal to the target (an isotropic mixture with different widths per component) makes
  every independence move accepted: the acceptance factor p(x')q(x)/(p(x)q(x')) is 1 only if the
  proposal's normalisation (log pi_k + sum log(1/sigma) - d/2 log 2 pi) and its draws match its
  density, so a dropped term or a wrong width fails;
- a bimodal target whose modes plain moves do not cross is sampled with the right mode weights
  from walkers that all start in one mode, with a proposal whose mode weights are reversed (0.25 on
  the heavy mode): leaving out or inverting the Hastings factor would give about 0.44 or less
  instead of 0.7;
- without the learned move the same walkers stay in their mode;
- pmix = 0, a cleared learned move and a trial move set afterwards give the plain chain bit for bit."""
import numpy as np
import pytest
from scipy import stats

from mcitest import ESTIM_UNCORR, OBS_INDICATOR, OBS_X2_SUM, PDF_GMM_ISO, RNG_MT, RNG_PHILOX, Chain, MciError, chain_setup

D = 4
W_T = np.array([0.3, 0.7])
S_T = np.array([0.8, 1.3])
MU_T = np.array([[-8.0, 0, 0, 0], [8.0, 0, 0, 0]])


def gmm_params():
    return [2, *W_T, *S_T, *MU_T.ravel()]


def exact():
    """P(x_0 > 0) and E|x|^2 of the target"""
    p = float(np.sum(W_T * stats.norm.sf(-MU_T[:, 0] / S_T)))
    r2 = float(np.sum(W_T * (np.sum(MU_T ** 2, axis=1) + D * S_T ** 2)))
    return np.array([p, r2])


def cfg(seed, learned=None):
    c = dict(ndim=D, seed=seed, ndecorr=200, pdfs=[(PDF_GMM_ISO, gmm_params())],
             obs=[(OBS_INDICATOR, [0], 64, 1, 0, ESTIM_UNCORR), (OBS_X2_SUM, [], 64, 1, 0, ESTIM_UNCORR)])
    if learned is not None:
        c["learned"] = learned
    return c


def run(c, W, nmc, rng, x0_mode=0, find=True):
    ch = chain_setup("orc", c, nwalkers=W)
    ch.setRng(rng)
    x = np.zeros((W, D))
    x[:, 0] = MU_T[x0_mode, 0]
    ch.setX(x)
    a, e = ch.integrate(nmc, find, True)
    wa, _ = ch.walkerResults()
    return ch, a, wa


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_MT])
def test_proposal_equal_to_the_target_is_always_accepted(rng):
    sig = np.repeat(S_T[:, None], D, axis=1)
    ch, a, wa = run(cfg(11, (W_T, MU_T, sig, 1.0)), W=16, nmc=4096, rng=rng, find=False)
    acc, rej, _ = ch.acceptance()
    assert rej <= 1e-6 * (acc + rej), (acc, rej)
    # the walkers then draw exactly from the target
    n = wa.shape[0]
    z = (wa.mean(axis=0) - exact()) / (wa.std(axis=0, ddof=1) / np.sqrt(n))
    assert np.all(np.abs(z) < 3), z


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_MT])
def test_bimodal_target_with_reversed_proposal_weights(rng):
    pw = np.array([0.75, 0.25])
    pmu = MU_T + np.array([[0.3, -0.2, 0.1, 0.0], [-0.4, 0.1, 0.0, 0.2]])
    psig = np.array([[1.2, 1.0, 0.9, 1.1], [1.5, 1.4, 1.6, 1.2]])
    ch, a, wa = run(cfg(12, (pw, pmu, psig, 0.3)), W=64, nmc=8192, rng=rng, x0_mode=1)
    z = (wa.mean(axis=0) - exact()) / (wa.std(axis=0, ddof=1) / np.sqrt(len(wa)))
    assert np.all(np.abs(z) < 3), (z, wa.mean(axis=0), exact())
    assert abs(a[0] - exact()[0]) < 0.03


def test_without_the_learned_move_walkers_stay_in_their_mode():
    ch, a, wa = run(cfg(13), W=16, nmc=4096, rng=RNG_PHILOX, x0_mode=1)
    assert a[0] > 0.99, a


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_MT])
def test_pmix_zero_clear_and_reset_give_the_plain_chain(rng):
    learned = (W_T, MU_T, np.ones((2, D)), 0.4)
    base = run(cfg(14), W=8, nmc=512, rng=rng)
    outs = [run(cfg(14, (W_T, MU_T, np.ones((2, D)), 0.0)), W=8, nmc=512, rng=rng)]
    c = cfg(14, learned)
    ch = chain_setup("orc", c, nwalkers=8)
    ch.clearLearnedMove()
    outs.append(ch)
    c2 = cfg(14, learned)
    c2["move"] = (0, 0, 1, None)  # setTrialMove after the learned move: chain_setup sets the move first
    ch2 = Chain("orc", D, 8)
    ch2.setSeed(14)
    ch2.setNdecorrelationSteps(200)
    ch2.setLearnedMove(*learned)
    ch2.setTrialMove(0, 0, 1, None)
    for kind, params in c2["pdfs"]:
        ch2.addSamplingFunction(kind, params)
    for o in c2["obs"]:
        ch2.addObservable(o[0], o[1], *o[2:])
    outs.append(ch2)
    for k, o in enumerate(outs):
        if isinstance(o, Chain):
            o.setRng(rng)
            x = np.zeros((8, D))
            x[:, 0] = MU_T[0, 0]
            o.setX(x)
            a, _ = o.integrate(512, True, True)
            wa, _ = o.walkerResults()
            o = (o, a, wa)
        assert np.array_equal(o[1], base[1]) and np.array_equal(o[2], base[2]), k
        assert np.array_equal(o[0].getX(), base[0].getX()), k


def test_learned_move_arguments():
    ch = Chain("orc", D, 4)
    ok = (W_T, MU_T, np.ones((2, D)))
    for bad in ((np.ones(9), np.zeros((9, D)), np.ones((9, D)), 0.5), (W_T, MU_T, -np.ones((2, D)), 0.5),
                (np.array([0.5, 0.0]), MU_T, np.ones((2, D)), 0.5), (*ok, 1.5), (*ok, -0.1),
                (W_T, np.full((2, D), np.nan), np.ones((2, D)), 0.5)):
        with pytest.raises(MciError):
            ch.setLearnedMove(*bad)
    ch.setLearnedMove(*ok, 0.5)
