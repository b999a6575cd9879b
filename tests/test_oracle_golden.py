"""The oracle against golden vectors written by the reference itself (oracle/_ref/mci_ref,
built from the reference sources; tests/golden/make_golden.py).  Bit-exact equality is
required everywhere: same flags (-O2 -ffp-contract=off), same libm."""
import numpy as np
import pytest

import golden_io as G
from mcitest import Chain, chain_setup, estimate, load, OBS_X, OBS_X2, ESTIM_UNCORR

CHAINS = G.load("chains.json")["scenarios"]


def run_calls(prefix, sc):
    c = chain_setup(prefix, G.cfg_of(sc["cfg"]))
    out = []
    for call in sc["calls"]:
        if "irange" in call:
            c.setIRange(*call["irange"])
        if "setx" in call:
            c.setX(call["setx"])
        a, e = c.integrate(call["nmc"], call["find"], call["decorr"])
        nd = c.getNObsDim()
        out.append(dict(avg=a[:nd], err=e[:nd], rate=c.acceptance()[2], steps=c.getWalkerSteps()[0], x=c.getX()[0]))
    return out


@pytest.mark.parametrize("sc", CHAINS, ids=[s["name"] for s in CHAINS])
def test_integrate_bit_exact_vs_reference(sc):
    got = run_calls("orc", sc)
    for g, r in zip(got, sc["results"]):
        assert r["cmd"] == "integrate"
        assert np.array_equal(g["avg"], G.hexv(r["avg"])), (g["avg"], G.hexv(r["avg"]))
        assert np.array_equal(g["err"], G.hexv(r["err"])), (g["err"], G.hexv(r["err"]))
        assert g["rate"] == float.fromhex(r["rate"])
        assert np.array_equal(g["steps"], G.hexv(r["steps"]))
        assert np.array_equal(g["x"], G.hexv(r["x"]))


TRACES = G.load("traces.json")["traces"]


@pytest.mark.parametrize("tr", TRACES, ids=[t["name"] for t in TRACES])
def test_step_trace_bit_exact(tr):
    import ctypes as C
    cfg = G.cfg_of(tr["cfg"])
    nd, n = cfg["ndim"], tr["nmc"]
    c = chain_setup("orc", cfg)
    xb = np.zeros(n * nd)
    ab = np.zeros(n, dtype=np.uint8)
    c.call("trace_enable", xb.ctypes.data_as(C.POINTER(C.c_double)), ab.ctypes.data_as(C.POINTER(C.c_uint8)), n)
    c.integrate(n, False, False)
    r = tr["result"]
    # the callback runs at initialisation and inside every step before the new state is
    # committed (MCIntegrator.cpp:265,347), so callback j >= 1 sees the position after step j-1
    ref_x = G.hexv(r["trace_x"]).reshape(-1, nd)
    assert ref_x.shape[0] == n + 1
    xs = xb.reshape(n, nd)
    assert np.array_equal(xs[:-1], ref_x[2:])
    assert np.array_equal(xs[-1], G.hexv(r["x"]))
    rates = G.hexv(r["trace_rate"])[1:]
    acc = np.rint(rates * np.arange(1, n + 1)).astype(np.int64)
    decided = np.diff(np.concatenate([[0], acc]))
    assert np.array_equal(decided, ab.astype(np.int64))
    assert 0 < ab.sum() < n


def test_independent_walkers_match_separate_reference_runs():
    d = G.load("walkers.json")
    cfg = G.cfg_of(d["cfg"])
    W = len(d["seeds"])
    c = chain_setup("orc", cfg, nwalkers=W)
    c.setWalkerSeeds(d["seeds"])
    a, e = c.integrate(d["nmc"])
    wa, we = c.walkerResults()
    for k, r in enumerate(d["results"]):
        assert np.array_equal(wa[k], G.hexv(r["avg"]))
        assert np.array_equal(we[k], G.hexv(r["err"]))
    steps = c.getWalkerSteps()
    for k, r in enumerate(d["results"]):
        assert np.array_equal(steps[k], G.hexv(r["steps"]))
    # combined as in MPIMCI::integrate (src/MPIMCI.cpp:85-92)
    assert np.allclose(a, wa.mean(axis=0), rtol=1e-14, atol=0)
    assert np.allclose(e, np.sqrt((we ** 2).sum(axis=0)) / W, rtol=1e-14, atol=0)


EST = G.load("estimators.json")["sets"]


@pytest.mark.parametrize("s", EST, ids=[f"n{s['n']}_d{s['ndim']}" for s in EST])
def test_estimators_bit_exact(s):
    x = G.b64(s["x_b64"]).reshape(s["n"], s["ndim"])
    for est, r in s["results"].items():
        a, e = estimate("orc", int(est), x)
        assert np.array_equal(a, G.hexv(r["avg"])), (est, a, r)
        assert np.array_equal(e, G.hexv(r["err"])), (est, e, r)


def test_walkdata_estimators_and_accumulators():
    """The checks of test/ut1/main.cpp on a rand()-driven walk, against the reference's outputs."""
    import ctypes as C
    w = G.load("walkdata.json")["walk"]
    x = G.b64(w["x_b64"]).reshape(-1, 2)
    n = x.shape[0]
    for est, key in [(1, "uncorr_nd"), (2, "corr_nd"), (3, "fc_nd"), (4, "mj_nd")]:
        a, e = estimate("orc", est, x)
        assert np.array_equal(a, G.hexv(w[key]["avg"])) and np.array_equal(e, G.hexv(w[key]["err"]))
    lib = load("orc")
    for j in range(2):
        col = np.ascontiguousarray(x[:, j])
        for est, key in [(1, "uncorr_1d"), (3, "fc_1d"), (4, "mj_1d")]:
            a, e = estimate("orc", est, col)
            assert (a[0], e[0]) == tuple(float.fromhex(v) for v in w[f"{key}_{j}"])
        a1, e1 = np.zeros(1), np.zeros(1)
        lib.orc_block_estimate(n, 1, col.ctypes.data_as(C.POINTER(C.c_double)), 2048,
                               a1.ctypes.data_as(C.POINTER(C.c_double)), e1.ctypes.data_as(C.POINTER(C.c_double)))
        assert (a1[0], e1[0]) == tuple(float.fromhex(v) for v in w[f"block_1d_{j}"])
    a2, e2 = np.zeros(2), np.zeros(2)
    lib.orc_block_estimate(n, 2, x.ctypes.data_as(C.POINTER(C.c_double)), 2048,
                           a2.ctypes.data_as(C.POINTER(C.c_double)), e2.ctypes.data_as(C.POINTER(C.c_double)))
    assert np.array_equal(a2, G.hexv(w["block_nd"]["avg"])) and np.array_equal(e2, G.hexv(w["block_nd"]["err"]))
