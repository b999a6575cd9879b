"""The oracle's shared-coupling branch (W walkers as W MPI ranks of one MPIMCI run) pinned by what
the mathematics fixes.  MPI is absent here, so the reference's USE_MPI branches
(src/MCIntegrator.cpp:21-34, 103-138, 184-229; src/MPIMCI.cpp:85-92) cannot be run; instead:

- W walkers given one repeated seed in MT mode are W copies of one reference chain.  Each rank's
  rate, estimate and error are then that chain's, so the rank average of the rate is the chain's
  rate, the reduced average is the chain's average and the reduced error is its error / sqrt(W).
  The shared run must therefore follow one single walker that is driven by the same rule with
  those numbers, and those numbers come from separate single-walker integrates.
- The step counts of one calibration iteration and one decorrelation round are
  max(100, sqrt(40000 d) / W) truncated (MCIntegrator.cpp:107,194), read from the counters of
  the last sample() call.
- With W = 1 the shared branch is the reference chain itself (the golden chains).
"""
import ctypes as C
import math

import numpy as np
import pytest

import golden_io as G
from mcitest import (COUPLING_INDEPENDENT, COUPLING_SHARED, ESTIM_MJ, OBS_X2, PDF_GAUSS_ISO, RNG_MT, Chain,
                     chain_setup)

FLT_MAX = float(np.finfo(np.float32).max)
FLT_MIN = float(np.finfo(np.float32).tiny)

# static_cast<int64_t>(std::max(100., sqrt(40000. * d) / W)): sqrt(120000) = 346.41, sqrt(640000) = 800
MIN_STAT = {(3, 1): 346, (3, 2): 173, (3, 4): 100, (3, 9): 100, (3, 16): 100, (3, 64): 100,
            (16, 1): 800, (16, 2): 400, (16, 9): 100, (16, 64): 100}


def gauss_cfg(ndim, x0, step=0.05):
    return dict(ndim=ndim, seed=4242, move=(0, 0, 1, None), pdfs=[(PDF_GAUSS_ISO, [])],
                obs=[(OBS_X2, [], 1, 1, 1, 2)], x0=x0, step=step)


def shared_copies(cfg, W, seed=4242):
    """W walkers that all run the stream of one reference MCI seeded with `seed`"""
    c = chain_setup("orc", cfg, nwalkers=W)
    c.setRng(RNG_MT)
    c.setCoupling(COUPLING_SHARED)
    c.setWalkerSeeds([seed] * W)
    return c


def single(cfg, seed=4242):
    c = chain_setup("orc", cfg, nwalkers=1)
    c.setRng(RNG_MT)
    c.setWalkerSeeds([seed])
    return c


def last_sample_steps(c):
    a, r = c.walkerAcceptance()
    return a + r


@pytest.mark.parametrize("ndim", [3, 16])
@pytest.mark.parametrize("W", [1, 2, 9, 64])
def test_steps_per_calibration_iteration_and_decorrelation_round(ndim, W):
    cfg = gauss_cfg(ndim, [0.25] * ndim)
    c = chain_setup("orc", cfg, nwalkers=W)
    c.setCoupling(COUPLING_SHARED)
    c.setNfindMRT2Iterations(1)
    c.integrate(0, True, False)
    assert np.all(last_sample_steps(c) == MIN_STAT[(ndim, W)])
    c.setNdecorrelationSteps(-4 * MIN_STAT[(ndim, W)])  # auto; the last round's sample() has MIN_NMC steps
    c.integrate(0, False, True)
    assert np.all(last_sample_steps(c) == MIN_STAT[(ndim, W)])


@pytest.mark.parametrize("W", [9, 64])
def test_shared_rate_is_the_mean_of_the_walker_rates(W):
    """One calibration iteration of W distinct walkers: every step is scaled by
    clamp(mean_w rate_w / target, 0.5, 2) (MCIntegrator.cpp:131-138,148)."""
    cfg = gauss_cfg(3, [0.25] * 3, step=0.7)
    c = chain_setup("orc", cfg, nwalkers=W)
    c.setCoupling(COUPLING_SHARED)
    c.setNfindMRT2Iterations(1)
    c.integrate(0, True, False)
    a, r = c.walkerAcceptance()
    rates = a / (a + r)
    assert len(set(rates.tolist())) > 1  # distinct walkers
    fact = min(2.0, max(0.5, math.fsum(rates) / W / 0.5))
    steps = c.getWalkerSteps()
    assert np.all(steps == steps[0])
    assert steps[0, 0] == pytest.approx(0.7 * fact, rel=4e-16, abs=0)
    # a rank sum divided twice (or not at all) would give a factor of 0.5 or 2
    assert 0.6 < fact < 1.9


def drive_single_calibration(cfg, W, nfind=-50, target=0.5):
    """One walker driven through findMRT2Step's rule, each iteration a fixed run of the shared
    MIN_STAT steps; returns (walker, iterations, step)."""
    s1 = single(cfg)
    n = MIN_STAT[(cfg["ndim"], W)]
    s1.setNdecorrelationSteps(n)
    step = cfg["step"]
    cons, counter = 0, 0
    while (nfind < 0 and cons < 5) or counter < nfind:
        s1.setMRT2Step(step)
        s1.integrate(0, False, True)
        a, r = s1.walkerAcceptance()
        assert a[0] + r[0] == n
        rate = a[0] / (a[0] + r[0]) if a[0] > 0 else 0.0
        cons = cons + 1 if abs(rate - target) < 0.05 else 0
        step = step * min(2.0, max(0.5, rate / target))
        step = min(step, 0.5 * (2 * FLT_MAX))  # unbound domain: size 2 FLT_MAX
        step = max(step, FLT_MIN)
        counter += 1
        if nfind < 0 and counter >= abs(nfind):
            break
    s1.setMRT2Step(step)
    return s1, counter, step


@pytest.mark.parametrize("ndim,W,step0", [(3, 4, 0.05), (3, 64, 3.0), (16, 2, 0.05)])
def test_shared_calibration_follows_one_chain(ndim, W, step0):
    cfg = gauss_cfg(ndim, [0.5] * ndim, step=step0)
    s1, iters, step = drive_single_calibration(cfg, W)
    assert 5 <= iters <= 50  # 50: the default cap of -50 iterations
    c = shared_copies(cfg, W)
    c.integrate(0, True, False)
    assert np.all(c.getWalkerSteps() == step)
    xs = c.getX()
    assert np.array_equal(xs, np.repeat(s1.getX(), W, axis=0))
    # the same rule driven for one iteration more or less ends elsewhere
    _, _, step_more = drive_single_calibration(cfg, W, nfind=iters + 1)
    assert step_more != step


def single_round_estimates(cfg, n, rounds):
    """avg/err of the Correlated estimator over consecutive runs of n steps of one walker, and the
    position after each run (the rounds of MCI::initialDecorrelation, MCIntegrator.cpp:171-236)."""
    s1 = single(cfg)
    out = []
    for _ in range(rounds):
        a, e = s1.integrate(n, False, False)
        out.append((a.copy(), e.copy(), s1.getX()[0].copy()))
    return out


def predicted_rounds(est, scale, cap, n):
    """rounds sampled before the loop ends, for rank errors scaled by `scale`"""
    oa, oe = est[0][0], est[0][1] * scale
    count, k = 0, 1
    while True:
        count += n
        k += 1
        if count >= cap:
            return k, True
        na, ne = est[k - 1][0], est[k - 1][1] * scale
        moved = any(abs(oa[i] - na[i]) > 2 * math.sqrt(oe[i] * oe[i] + ne[i] * ne[i]) for i in range(len(na)))
        oa, oe = na, ne
        if not moved:
            return k, False


@pytest.mark.parametrize("W,x0,step", [(4, [2.5, 2.5, -2.5], 0.45), (4, [0.3, 0.2, 0.1], 0.9),
                                       (16, [3.0, -2.5, 2.0], 0.9), (16, [0.3, 0.2, 0.1], 0.9)])
def test_shared_decorrelation_combines_rank_estimates(W, x0, step):
    """Identical ranks: MPIReduceAvgErr gives avg = the chain's avg and err = its err / sqrt(W),
    so the 2-sigma test (MCIntegrator.cpp:225-231) runs on errors sqrt(W) smaller."""
    cfg = gauss_cfg(3, x0, step=step)
    n = MIN_STAT[(3, W)]
    cap = 60 * n
    est = single_round_estimates(cfg, n, 62)
    k_shared, capped = predicted_rounds(est, 1.0 / math.sqrt(W), cap, n)
    k_unreduced, _ = predicted_rounds(est, 1.0, cap, n)
    assert k_shared != k_unreduced and not capped  # the case tells a missing reduce apart
    c = shared_copies(cfg, W)
    c.setNdecorrelationSteps(-cap)
    c.integrate(0, False, True)
    assert np.array_equal(c.getX(), np.repeat(est[k_shared - 1][2][None, :], W, axis=0))
    nw = C.c_int64()
    c.call("get_warnings", C.byref(nw))
    assert nw.value == 0


def test_shared_decorrelation_cap_warns():
    W = 4
    cfg = gauss_cfg(3, [2.5, 2.5, -2.5], step=0.45)
    n = MIN_STAT[(3, W)]
    est = single_round_estimates(cfg, n, 4)
    k, capped = predicted_rounds(est, 0.5, 3 * n, n)
    assert capped and k == 4
    c = shared_copies(cfg, W)
    c.setNdecorrelationSteps(-3 * n)
    c.integrate(0, False, True)
    assert np.array_equal(c.getX()[0], est[3][2])
    nw = C.c_int64()
    c.call("get_warnings", C.byref(nw))
    assert nw.value == 1


@pytest.mark.parametrize("W", [4, 64])
def test_identical_ranks_combine_to_the_chain(W):
    """A whole integrate (auto calibration, auto decorrelation, MJBlocker blocks) of W identical
    ranks: every rank holds the chain, the combined average is its average and the combined
    error its error / sqrt(W) (MPIMCI.cpp:85-92)."""
    cfg = dict(ndim=3, seed=77, pdfs=[(PDF_GAUSS_ISO, [])], obs=[(OBS_X2, [], 64, 1, 1, ESTIM_MJ)],
               x0=[1.0, 0.5, -1.0])
    c = shared_copies(cfg, W, seed=77)
    a, e = c.integrate(64 * 128)
    wa, we = c.walkerResults()
    assert np.all(wa == wa[0]) and np.all(we == we[0])
    assert np.array_equal(a, wa[0])
    assert np.array_equal(e, we[0] / math.sqrt(W))  # sqrt(W) is 2 or 8: exact
    x = c.getX()
    assert np.all(x == x[0])
    acc = c.walkerAcceptance()
    assert np.all(acc[0] == acc[0][0])


CHAINS = [s for s in G.load("chains.json")["scenarios"] if not s["name"].startswith("multistep")]


@pytest.mark.parametrize("sc", CHAINS[:12], ids=[s["name"] for s in CHAINS[:12]])
def test_one_rank_is_the_reference_chain(sc):
    c = chain_setup("orc", G.cfg_of(sc["cfg"]))
    c.setCoupling(COUPLING_SHARED)
    for call, r in zip(sc["calls"], sc["results"]):
        if "irange" in call:
            c.setIRange(*call["irange"])
        if "setx" in call:
            c.setX(call["setx"])
        a, e = c.integrate(call["nmc"], call["find"], call["decorr"])
        nd = c.getNObsDim()
        assert np.array_equal(a[:nd], G.hexv(r["avg"])) and np.array_equal(e[:nd], G.hexv(r["err"]))
        assert np.array_equal(c.getWalkerSteps()[0], G.hexv(r["steps"]))
        assert np.array_equal(c.getX()[0], G.hexv(r["x"]))


def test_independent_and_shared_differ_only_in_coupling():
    """Sanity: with distinct walkers the shared calibration gives all walkers one step while the
    independent one gives each its own."""
    cfg = gauss_cfg(3, [0.25] * 3, step=0.7)
    out = {}
    for mode in (COUPLING_INDEPENDENT, COUPLING_SHARED):
        c = chain_setup("orc", cfg, nwalkers=8)
        c.setCoupling(mode)
        c.integrate(0, True, False)
        out[mode] = c.getWalkerSteps()[:, 0]
    assert len(set(out[COUPLING_INDEPENDENT].tolist())) > 1
    assert len(set(out[COUPLING_SHARED].tolist())) == 1
