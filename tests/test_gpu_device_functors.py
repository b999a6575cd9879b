"""The caller's own plugins as device code (mci_add_sampling_function_device,
mci_add_observable_device): the generic kernels compiled at run time with the functor sources.

- reference chains of plugins outside the library's catalogue (tests/golden/device_functors.json,
  written by the reference from oracle/ref_harness.cpp QuarticPDF and MomentsObs): rates,
  calibrated steps, averages, errors and positions bit-identical, through calibration, fixed and
  automatic decorrelation, Block and Full accumulators, all-coordinate and single-coordinate moves;
- restatements of the catalogue's diagonal Gaussian and x_i^2 give the catalogue's chains bit for
  bit on many walkers, both random modes and both couplings;
- a functor that does not compile fails integrate with the compiler's message."""
import numpy as np
import pytest

import device_functors as DF
import golden_io as G
from mcitest import (COUPLING_INDEPENDENT, COUPLING_SHARED, ESTIM_MJ, KERNELS_GENERIC, OBS_X2, PDF_GAUSS_DIAG, RNG_MT, RNG_PHILOX, Chain,
                     MciError, chain_setup)

pytestmark = pytest.mark.gpu

SCEN = G.load("device_functors.json")["scenarios"]


@pytest.mark.parametrize("sc", SCEN, ids=[s["name"] for s in SCEN])
def test_device_functors_reproduce_reference_chains(sc):
    print("device functor scenario", sc["name"], flush=True)
    cfg = G.cfg_of(sc["cfg"])
    c = chain_setup("mci", cfg)
    for call, r in zip(sc["calls"], sc["results"]):
        if "setx" in call:
            c.setX(call["setx"])
        a, e = c.integrate(call["nmc"], call["find"], call["decorr"])
        nd = c.getNObsDim()
        assert c.acceptance()[2] == float.fromhex(r["rate"])
        assert np.array_equal(c.getWalkerSteps()[0], G.hexv(r["steps"]))
        assert np.array_equal(a[:nd], G.hexv(r["avg"]))
        assert np.array_equal(e[:nd], G.hexv(r["err"]))
        assert np.array_equal(c.getX()[0], G.hexv(r["x"]))


@pytest.mark.parametrize("rng", [RNG_MT, RNG_PHILOX])
@pytest.mark.parametrize("coupling", [COUPLING_INDEPENDENT, COUPLING_SHARED])
def test_device_functor_restating_the_catalogue_is_bit_identical(rng, coupling):
    d, W = 6, 130
    print("device functor vs catalogue", rng, coupling, flush=True)
    a = 0.5 / np.random.default_rng(3).uniform(0.5, 2.0, d) ** 2
    out = {}
    for name in ("catalogue", "device"):
        c = Chain("mci", d, W)
        c.setSeed(4242)
        c.setRng(rng)
        c.setCoupling(coupling)
        c.setNdecorrelationSteps(200)
        c.setKernelPath(KERNELS_GENERIC)
        if name == "catalogue":
            c.addSamplingFunction(PDF_GAUSS_DIAG, a)
            c.addObservable(OBS_X2, (), 20, 1, True, ESTIM_MJ)
        else:
            c.addSamplingFunctionDevice("diag", DF.DIAG, d, a)
            c.addObservableDevice("xsq", DF.X2, d, (), 20, 1, True, ESTIM_MJ)
        av, er = c.integrate(20 * 64)
        av2, er2 = c.integrate(20 * 16, False, False)
        out[name] = (av, er, av2, er2, c.getX(), c.getWalkerSteps(), *c.walkerAcceptance(), *c.walkerResults(), c.observableData(0, d))
    f, g = out["device"], out["catalogue"]
    for k in range(len(f)):
        assert np.array_equal(f[k], g[k]), k


def test_device_functor_errors():
    c = Chain("mci", 2, 64)
    c.addSamplingFunctionDevice("broken", "__device__ void broken_proto(int n, MciVec x, const double *p, MciVec q) { q[0] = y; }", 1)
    c.addObservable(OBS_X2, (), 1, 1, False)
    with pytest.raises(MciError) as e:
        c.integrate(64)
    assert "undeclared identifier 'y'" in str(e.value)
    c2 = Chain("mci", 2, 64)
    c2.addSamplingFunctionDevice("quartic", DF.QUARTIC, 2, [1.0, 1.0])
    c2.addObservableDevice("quartic", DF.MOMENTS, 2)
    with pytest.raises(MciError) as e:
        c2.integrate(64)
    assert "different sources" in str(e.value)


def test_updated_acceptance_override_is_used_for_partial_moves_only():
    """A functor whose updatedAcceptance accepts everything: single-coordinate moves (the override)
    are all accepted, all-coordinate moves (protoFunction and acceptanceFunction) are not."""
    src = DF.QUARTIC.replace("quartic_", "always_") + r"""
__device__ double always_updated_acceptance(int ndim, int nproto, MciVec xnew, MciChanged changed, MciVec po, MciVec pn,
                                            const double *par)
{
    for (int i = 0; i < changed.n; ++i) { const int j = changed[i]; const double q = xnew[j] * xnew[j]; pn[j] = (q * q) * par[j]; }
    return 1.0;
}
"""
    rates = {}
    for veclen in (1, 0):
        c = Chain("mci", 3, 64)
        c.setSeed(7)
        c.setTrialMove(0, veclen, 1, None)
        c.setMRT2Step(2.0)
        c.addSamplingFunctionDevice("always", src, 3, [1.0, 1.0, 1.0])
        c.addObservableDevice("moments", DF.MOMENTS, 2, (), 16, 1, False, ESTIM_MJ)
        c.integrate(16 * 16, False, False)
        rates[veclen] = c.acceptance()[2]
    assert rates[1] == 1.0 and rates[0] < 0.9, rates
