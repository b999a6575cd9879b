"""Parity of the HIP path (through the C-ABI) with the reference and the oracle.

- reference streams (MT19937_64 mode): golden vectors written by the reference binary must be
  reproduced bit for bit — averages, MJBlocker/FCBlocker errors, calibrated step sizes, final
  positions, acceptance rates and every step of the traced chains;
- native streams (Philox): the oracle restates the same draws; walkers and combined results
  must match it bit for bit (FP64), at full configuration size through a walker subset;
- physics: estimates within 3 sigma of the analytic values.
"""
import numpy as np
import pytest

import golden_io as G
from mcitest import (COUPLING_INDEPENDENT, COUPLING_SHARED, ESTIM_CORR, ESTIM_FC, ESTIM_MJ, ESTIM_NOOP, ESTIM_UNCORR, KERNELS_AUTO, KERNELS_GENERIC,
                     OBS_EXP_X2, OBS_X, OBS_X2, OBS_X2_SUM, PDF_GAUSS_DIAG, PREC_FP32, RNG_MT, RNG_PHILOX, Chain,
                     chain_setup, estimate)

pytestmark = pytest.mark.gpu


def exact_components(cfg):
    """Observable components that must be bit-identical: all but those whose value passes through
    exp or log (device ocml exp/log and glibc's may differ by one ulp on a sample; the decisions,
    the positions and every other component stay bit-exact): EXP_X2, the log value of a mixture
    density, and the squares of such observables."""
    from mcitest import OBS_DEP_SQUARE, OBS_PDF_LOG, OBS_PDF_PROTO, PDF_GMM_ISO
    masks = []
    for kind, params, *_ in cfg.get("obs", []):
        if kind == OBS_DEP_SQUARE:
            masks.append(masks[int(params[0])])
            continue
        n = cfg["ndim"] if kind in (OBS_X2, OBS_X, OBS_PDF_PROTO) else 1
        exact = kind != OBS_EXP_X2
        if kind == OBS_PDF_LOG:
            exact = cfg["pdfs"][int(params[0]) if params else 0][0] != PDF_GMM_ISO
        masks.append([exact] * n)
    return np.array([m for ms in masks for m in ms], dtype=bool)


def check_components(got, ref, exact):
    """bit-exact where `exact`, else within 1e-12 relative (the FP64 bar for transcendental observables)"""
    assert np.array_equal(got[exact], ref[exact]), (np.nonzero(got != ref), got, ref)
    if (~exact).any():
        np.testing.assert_allclose(got[~exact], ref[~exact], rtol=1e-12, atol=0)

CHAINS = G.load("chains.json")["scenarios"]


def move_srrd(cfg):
    """the move distribution (the contained one of a multi-step move)"""
    if "multistep" in cfg:
        return cfg["multistep"][1]
    return cfg["move"][0] if "move" in cfg else 0


def check_close(got, ref, rtol):
    """within rtol of the reference, relative to the largest magnitude of the reference vector"""
    np.testing.assert_allclose(got, ref, rtol=rtol, atol=rtol * float(np.max(np.abs(ref))))


@pytest.mark.parametrize("sc", CHAINS, ids=[s["name"] for s in CHAINS])
def test_reference_chains_bit_exact(sc):
    """Every reference scenario on the device.  Moves drawn from distributions other than the
    uniform one pass through the device's log, sqrt, tan, pow and exp, which can differ from
    glibc's in the last place: their decisions, acceptance rates and calibrated steps must still
    be identical, positions and estimates within 1e-12 (errors 1e-10) relative."""
    print("scenario", sc["name"], flush=True)
    cfg = G.cfg_of(sc["cfg"])
    c = chain_setup("mci", cfg)
    for call, r in zip(sc["calls"], sc["results"]):
        if "irange" in call:
            c.setIRange(*call["irange"])
        if "setx" in call:
            c.setX(call["setx"])
        a, e = c.integrate(call["nmc"], call["find"], call["decorr"])
        nd = c.getNObsDim()
        assert c.acceptance()[2] == float.fromhex(r["rate"])
        assert np.array_equal(c.getWalkerSteps()[0], G.hexv(r["steps"]))
        if move_srrd(cfg) != 0:
            check_close(a[:nd], G.hexv(r["avg"]), 1e-12)
            check_close(e[:nd], G.hexv(r["err"]), 1e-10)
            check_close(c.getX()[0], G.hexv(r["x"]), 1e-12)
            continue
        exact = exact_components(cfg)
        check_components(a[:nd], G.hexv(r["avg"]), exact)
        check_components(e[:nd], G.hexv(r["err"]), exact)
        assert np.array_equal(c.getX()[0], G.hexv(r["x"]))


TRACES = G.load("traces.json")["traces"]


@pytest.mark.parametrize("tr", TRACES, ids=[t["name"] for t in TRACES])
def test_reference_step_traces_bit_exact(tr):
    """Every position of the chain: a Full accumulator of x holds the trace (FullAccumulator.cpp:14-18)."""
    cfg = G.cfg_of(tr["cfg"])
    nd, n = cfg["ndim"], tr["nmc"]
    cfg["obs"] = [(OBS_X, [], 1, 1, 0, 1)]
    c = chain_setup("mci", cfg)
    c.integrate(n, False, False)
    xs = c.observableData(0, nd)[0]
    ref_x = G.hexv(tr["result"]["trace_x"]).reshape(-1, nd)
    assert np.array_equal(xs[:-1], ref_x[2:])
    assert np.array_equal(xs[-1], G.hexv(tr["result"]["x"]))


PHASES = G.load("traces.json")["phase_traces"]


@pytest.mark.parametrize("tr", PHASES, ids=[t["name"] for t in PHASES])
def test_reference_callback_through_all_phases(tr):
    """setCallback: the calls and what each sees (position, acceptance rate) through calibration,
    decorrelation and sampling equal the reference's, bit for bit, in the same order."""
    print("callback", tr["name"], flush=True)
    cfg = G.cfg_of(tr["cfg"])
    nd = cfg["ndim"]
    seen_x, seen_rate = [], []
    c = chain_setup("mci", cfg)
    c.setCallback(lambda x, acc, rej, rate: (seen_x.append(x), seen_rate.append(rate)))
    c.integrate(tr["nmc"], bool(tr["find"]), bool(tr["decorr"]))
    ref_x = G.hexv(tr["result"]["trace_x"]).reshape(-1, nd)
    ref_rate = G.hexv(tr["result"]["trace_rate"])
    assert len(seen_x) == len(ref_x)
    assert np.array_equal(np.array(seen_x), ref_x)
    assert np.array_equal(np.array(seen_rate), ref_rate)


FILES = G.load("files.json")["files"]


@pytest.mark.parametrize("fs", FILES, ids=[f["name"] for f in FILES])
def test_reference_observable_and_walker_files(fs, tmp_path):
    """storeObservablesOnFile / storeWalkerPositionsOnFile: the files equal the reference's text;
    with three walkers, walker 0 (the context seed) writes the reference's files and the others
    files of the same shape."""
    print("files", fs["name"], flush=True)
    cfg = G.cfg_of(fs["cfg"])
    for W in (1, 3):
        c = chain_setup("mci", cfg, nwalkers=W)
        c.storeObservablesOnFile(tmp_path / "obs.txt", fs["freq_obs"])
        c.storeWalkerPositionsOnFile(tmp_path / "wlk.txt", fs["freq_wlk"])
        c.integrate(fs["nmc"], True, True)
        if W == 1:
            assert (tmp_path / "obs.txt").read_text() == fs["obs"]
            assert (tmp_path / "wlk.txt").read_text() == fs["wlk"]
            continue
        assert (tmp_path / "obs.txt.0").read_text() == fs["obs"]
        assert (tmp_path / "wlk.txt.0").read_text() == fs["wlk"]
        for w in (1, 2):
            for nm, ref in (("obs.txt", fs["obs"]), ("wlk.txt", fs["wlk"])):
                got = (tmp_path / f"{nm}.{w}").read_text().splitlines()
                want = ref.splitlines()
                assert len(got) == len(want)
                assert [g.split()[0] for g in got] == [r.split()[0] for r in want]
                assert [len(g.split()) for g in got] == [len(r.split()) for r in want]


def test_reference_walkers_bit_exact():
    d = G.load("walkers.json")
    W = len(d["seeds"])
    c = chain_setup("mci", G.cfg_of(d["cfg"]), nwalkers=W)
    c.setWalkerSeeds(d["seeds"])
    c.integrate(d["nmc"])
    wa, we = c.walkerResults()
    steps = c.getWalkerSteps()
    for k, r in enumerate(d["results"]):
        assert np.array_equal(wa[k], G.hexv(r["avg"])), k
        assert np.array_equal(we[k], G.hexv(r["err"])), k
        assert np.array_equal(steps[k], G.hexv(r["steps"])), k


EST = G.load("estimators.json")["sets"]


@pytest.mark.parametrize("s", EST, ids=[f"n{s['n']}_d{s['ndim']}" for s in EST])
def test_device_estimators_bit_exact(s):
    x = G.b64(s["x_b64"]).reshape(s["n"], s["ndim"])
    for est, r in s["results"].items():
        a, e = estimate("mci", int(est), x)
        assert np.array_equal(a, G.hexv(r["avg"])), (est, a)
        assert np.array_equal(e, G.hexv(r["err"])), (est, e)


@pytest.mark.parametrize("n", [2, 8, 32, 64, 128, 256, 2048, 16384])
def test_device_mjblocker_power_of_two_sizes_match_oracle(n):
    """MJBlocker, and Correlated (MJBlocker at these sizes), on 2^k values: the device's
    register-resident form (n <= 64) and its streaming form with the first four levels computed
    on blocks of 16 in registers (n >= 128) equal the oracle bit for bit, on an AR(1) series in 5
    components.  (n = 1 is left out: the reference's error formula then reads before its level
    arrays.)"""
    rng = np.random.default_rng(100 + n)
    x = np.zeros((n, 5))
    e = rng.normal(size=(n, 5))
    for i in range(1, n):
        x[i] = 0.8 * x[i - 1] + e[i]
    for est in (ESTIM_MJ, ESTIM_CORR):
        a, er = estimate("mci", est, x)
        b, eb = estimate("orc", est, x)
        assert np.array_equal(a, b) and np.array_equal(er, eb), (n, est)


@pytest.mark.parametrize("n", [50, 51, 63, 99, 100, 101, 128, 129, 137, 160, 161, 250, 800, 1000, 2047, 2049, 4097])
def test_device_fcblocker_matches_oracle(n):
    """FCBlocker (and Correlated, FCBlocker off the powers of 2) against the oracle bit for bit over
    the block lengths the 45 block counts take and the device's forms: n = 50 (every count its own
    length, down to 1), n = 100 (in registers), n <= 128 (staged in LDS, counts sharing a length),
    129..2048 (streamed past running block sums; 2047 with more than 64 KB of LDS), longer series
    (swept from memory), in one and in five components (the two error normalisations), with a
    constant component (zero variance) among AR(1) series."""
    rng = np.random.default_rng(700 + n)
    x = np.zeros((n, 5))
    e = rng.normal(size=(n, 5))
    for i in range(1, n):
        x[i] = np.array([0.3, 0.8, 0.95, 0.0, -0.5]) * x[i - 1] + e[i]
    x[:, 3] = 1.75
    for data in (x, x[:, 1:2]):
        for est in (ESTIM_FC, ESTIM_CORR):
            a, er = estimate("mci", est, data)
            b, eb = estimate("orc", est, data)
            assert np.array_equal(a, b) and np.array_equal(er, eb), (n, data.shape, est, a - b, er - eb)


def sigma_a(n):
    sig = G.hexv(G.load("sigmas.json")["sigma16" if n == 16 else "sigma64"])
    return sig, 0.5 / (sig * sig)


def cfg2(nmc_blocks=64, blocksize=160, ndecorr=1000):
    sig, a = sigma_a(16)
    return dict(ndim=16, seed=1337, ndecorr=ndecorr, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, a)],
                obs=[(OBS_X2, [], blocksize, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], blocksize, 1, 1, ESTIM_MJ)])


@pytest.mark.parametrize("rng", [RNG_MT, RNG_PHILOX])
@pytest.mark.parametrize("coupling", [COUPLING_INDEPENDENT, COUPLING_SHARED])
def test_oracle_parity_many_walkers(rng, coupling):
    W, nmc = 200, 64 * 20
    res = {}
    for prefix in ("orc", "mci"):
        c = chain_setup(prefix, cfg2(blocksize=20), nwalkers=W)
        c.setRng(rng)
        c.setCoupling(coupling)
        if prefix == "mci":
            c.setKernelPath(KERNELS_GENERIC)
        a, e = c.integrate(nmc)
        res[prefix] = (a, e, *c.walkerResults(), c.getWalkerSteps(), c.getX(), c.walkerAcceptance())
    o, g = res["orc"], res["mci"]
    for k in range(6):
        assert np.array_equal(o[k], g[k]), k
    assert np.array_equal(o[6][0], g[6][0]) and np.array_equal(o[6][1], g[6][1])


@pytest.mark.parametrize("srrd", range(1, 10))
def test_move_distributions_many_walkers(srrd):
    """Each move distribution on 128 walkers in both random-stream modes, the generic kernels
    against the oracle: identical decisions and calibrated steps, positions and estimates within
    1e-12 relative (errors 1e-10; the device's transcendental functions may differ in the last place)."""
    print("srrd", srrd, flush=True)
    W, nmc = 128, 16 * 64
    for rng in (RNG_MT, RNG_PHILOX):
        res = {}
        for prefix in ("orc", "mci"):
            cfg = dict(ndim=4, seed=500 + srrd, ndecorr=100, move=(srrd, 1 if srrd % 2 else 0, 1, None),
                       pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0, 0.25])], obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR)])
            c = chain_setup(prefix, cfg, nwalkers=W)
            c.setRng(rng)
            a, e = c.integrate(nmc)
            a2, e2 = c.integrate(nmc, False, False)
            res[prefix] = (a, e, a2, e2, *c.walkerResults(), c.getX(), c.getWalkerSteps(), c.walkerAcceptance())
        o, g = res["orc"], res["mci"]
        assert np.array_equal(o[7], g[7]), (rng, "steps")
        assert np.array_equal(o[8][0], g[8][0]) and np.array_equal(o[8][1], g[8][1]), (rng, "acceptance")
        for k in (0, 2, 4, 6):
            check_close(g[k], o[k], 1e-12)
        for k in (1, 3, 5):
            check_close(g[k], o[k], 1e-10)


@pytest.mark.parametrize("srrd", [0, 1])
def test_multistep_move_many_walkers(srrd):
    """MultiStepMove (four inner single-coordinate steps under an own Gaussian, around a
    Gaussian-mixture target) on 96 walkers in both random-stream modes, the generic kernels against
    the oracle; with the uniform inner move bit-exact, with the Gaussian one within 1e-12."""
    print("multistep srrd", srrd, flush=True)
    W, nmc = 96, 16 * 32
    for rng in (RNG_MT, RNG_PHILOX):
        res = {}
        for prefix in ("orc", "mci"):
            cfg = dict(ndim=3, seed=31 + srrd, ndecorr=50, step=0.7, multistep=(4, srrd, 1, 1, None, [(PDF_GAUSS_DIAG, [0.2, 0.3, 0.4])]),
                       pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0])], obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR)])
            c = chain_setup(prefix, cfg, nwalkers=W)
            c.setRng(rng)
            a, e = c.integrate(nmc, False, True)
            res[prefix] = (a, e, *c.walkerResults(), c.getX(), c.walkerAcceptance())
        o, g = res["orc"], res["mci"]
        assert np.array_equal(o[5][0], g[5][0]) and np.array_equal(o[5][1], g[5][1]), (rng, "acceptance")
        for k in range(5):
            if srrd == 0:
                assert np.array_equal(o[k], g[k]), (rng, k)
            else:
                check_close(g[k], o[k], 1e-12 if k % 2 == 0 else 1e-10)


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_MT])
def test_auto_decorrelation_and_calibration_parity(rng):
    """Default settings (auto calibration, auto decorrelation with FCBlocker) for many walkers, on
    the specialised kernels: independent walkers finish their decorrelation rounds at different
    counts (the later rounds skip the finished ones), so on the reference streams their generator
    positions diverge; the generic kernel and the oracle agree bit for bit."""
    for coupling in (COUPLING_INDEPENDENT, COUPLING_SHARED):
        res = {}
        for name in ("orc", "mci", "generic"):
            cfg = dict(ndim=3, seed=77, pdfs=[(1, [])], obs=[(OBS_X2, [], 1, 1, 1, 2)], x0=[2.0, -1.0, 3.0])
            c = chain_setup("orc" if name == "orc" else "mci", cfg, nwalkers=96)
            c.setRng(rng)
            c.setCoupling(coupling)
            if name == "generic":
                c.setKernelPath(KERNELS_GENERIC)
            a, e = c.integrate(4096)
            res[name] = (a, e, *c.walkerResults(), c.getWalkerSteps(), c.getX())
        for k in range(6):
            assert np.array_equal(res["orc"][k], res["mci"][k]), (coupling, k)
            assert np.array_equal(res["generic"][k], res["mci"][k]), (coupling, k)


@pytest.mark.parametrize("variant", ["all16", "single16", "all20"])
@pytest.mark.parametrize("rng", [RNG_MT, RNG_PHILOX])
@pytest.mark.parametrize("coupling", [COUPLING_INDEPENDENT, COUPLING_SHARED])
def test_reference_defaults_on_the_specialised_kernels(coupling, rng, variant):
    """The reference's defaults at d = 16 on 130 walkers (two full waves and two lanes): the
    all-coordinate move, mt19937_64 streams, automatic calibration (independent: 800-step rounds,
    the walkers still calibrating compacted for each launch), automatic decorrelation (independent:
    800-value series, the streaming FCBlocker, compacted rounds, then the generators realigned; shared:
    100-value series, FCBlocker in registers) and addObservable's Full accumulator with the
    Correlated estimator on 1,024 steps (MJBlocker with its first levels on register blocks).
    Averages, errors, per-walker results, steps and positions equal the generic kernel's and the
    oracle's bit for bit.  On the native streams as well (there a compacted lane draws from its
    walker's Philox stream through the lane-to-walker map), and on the other two specialised
    kernels: single-coordinate moves at d = 16 and all-coordinate moves at d = 20 (held in LDS)."""
    d = 20 if variant == "all20" else 16
    sig, a = sigma_a(16 if d == 16 else 64)
    a = a[:d]
    res = {}
    for name in ("orc", "mci", "generic"):
        cfg = dict(ndim=d, seed=2024, pdfs=[(PDF_GAUSS_DIAG, a)], obs=[(OBS_X2, [], 1, 1, 1, ESTIM_CORR)])
        if variant == "single16":
            cfg["move"] = (0, 1, 1, None)
        c = chain_setup("orc" if name == "orc" else "mci", cfg, nwalkers=130)
        c.setRng(rng)
        c.setCoupling(coupling)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a_, e_ = c.integrate(1024)
        res[name] = (a_, e_, *c.walkerResults(), c.getWalkerSteps(), c.getX())
    for k in range(6):
        assert np.array_equal(res["orc"][k], res["mci"][k]), (coupling, rng, variant, k)
        assert np.array_equal(res["generic"][k], res["mci"][k]), (coupling, rng, variant, k)


RTOL_BLOCKS = 1e-12  # dwell-time block sums vs step-by-step sums (same FP64 values, other rounding)


@pytest.mark.parametrize("blocksize,coupling,W", [(20, COUPLING_INDEPENDENT, 256), (20, COUPLING_SHARED, 256),
                                                 (0, COUPLING_INDEPENDENT, 256), (20, COUPLING_INDEPENDENT, 200)])
def test_specialised_kernel_matches_generic_and_oracle(blocksize, coupling, W):
    """The LDS/dwell-time kernel makes the same decisions as the generic kernel and the oracle
    (positions, step sizes and acceptance counters bit-identical); block means and estimates
    agree to 1e-12 relative.  (200 walkers: the last wave has 8 live lanes.)"""
    nmc = 64 * 20
    cfg = cfg2(blocksize=blocksize)
    if blocksize == 0:
        cfg["obs"] = [(OBS_X2, [], 0, 1, 0, ESTIM_NOOP), (OBS_X2_SUM, [], 0, 1, 0, ESTIM_NOOP), (OBS_EXP_X2, [16.0], 0, 1, 0, ESTIM_NOOP)]
    out = {}
    for name in ("fast", "generic", "orc"):
        c = chain_setup("orc" if name == "orc" else "mci", cfg, nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(coupling)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a, e = c.integrate(nmc)
        out[name] = (a, e, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance())
    f, g, o = out["fast"], out["generic"], out["orc"]
    for k in (4, 5, 6, 7):
        assert np.array_equal(f[k], g[k]) and np.array_equal(g[k], o[k]), k
    exact = exact_components(cfg)
    for k in (0, 1):
        check_components(g[k], o[k], exact)
    for k in (2, 3):
        for wk in range(W):
            check_components(g[k][wk], o[k][wk], exact)
    for k in (0, 1, 2, 3):
        np.testing.assert_allclose(f[k], g[k], rtol=RTOL_BLOCKS, atol=1e-300)


def test_specialised_shared_calibration_ragged_chunks_and_repeat():
    """Shared calibration on the specialised path with 1,000 walkers (four chunks of the
    deterministic walker sum, the last one partial: the fused rates/sums/control launch and its
    last-arriving workgroup), then a second integrate on the same context (its first calibration
    batch is sized from the first call): positions, step sizes and acceptance counters equal the
    oracle's bit for bit after each call, block means and estimates to 1e-12."""
    W, nmc = 1000, 64 * 20
    out = {}
    for name in ("fast", "orc"):
        c = chain_setup("orc" if name == "orc" else "mci", cfg2(blocksize=20), nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(COUPLING_SHARED)
        calls = []
        for _ in range(2):
            a, e = c.integrate(nmc)
            calls.append((a, e, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance()))
        out[name] = calls
    for f, o in zip(out["fast"], out["orc"]):
        for k in (4, 5, 6, 7):
            assert np.array_equal(f[k], o[k]), k
        for k in (0, 1, 2, 3):
            np.testing.assert_allclose(f[k], o[k], rtol=RTOL_BLOCKS, atol=1e-300)


@pytest.mark.parametrize("nfind", [1, 3, 6])
def test_specialised_shared_calibration_fixed_iterations(nfind):
    """A fixed number of shared calibration iterations on the specialised path (the loop ends on
    the counter, its last factor applied either by a following sampling launch or by the step
    scaling after the loop, depending on where the batch ends): three integrates in a row equal
    the oracle's positions, step sizes and counters bit for bit."""
    W, nmc = 320, 64 * 10
    out = {}
    for name in ("fast", "orc"):
        c = chain_setup("orc" if name == "orc" else "mci", cfg2(blocksize=10), nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(COUPLING_SHARED)
        c.setNfindMRT2Iterations(nfind)
        calls = []
        for _ in range(3):
            c.integrate(nmc)
            calls.append((c.getWalkerSteps(), c.getX(), *c.walkerAcceptance()))
        out[name] = calls
    for f, o in zip(out["fast"], out["orc"]):
        for k in range(4):
            assert np.array_equal(f[k], o[k]), (nfind, k)


def test_full_size_configuration2_walker_subset():
    """65,536 walkers x 10,240 steps (d=16): independent walkers 0..63 equal an oracle run of 64
    walkers (walker w's Philox stream depends only on (seed, w)): positions and counters bit for
    bit, estimates to 1e-12; the combined estimate agrees with sigma_i^2 and sum sigma_i^2
    within 3 error bars."""
    W, nmc = 65536, 10240
    g = chain_setup("mci", cfg2(), nwalkers=W)
    g.setRng(RNG_PHILOX)
    a, e = g.integrate(nmc)
    wa, we = g.walkerResults()
    o = chain_setup("orc", cfg2(), nwalkers=64)
    o.setRng(RNG_PHILOX)
    o.integrate(nmc)
    oa, oe = o.walkerResults()
    assert np.array_equal(g.getX()[:64], o.getX())
    assert np.array_equal(g.walkerAcceptance()[0][:64], o.walkerAcceptance()[0])
    np.testing.assert_allclose(wa[:64], oa, rtol=RTOL_BLOCKS)
    np.testing.assert_allclose(we[:64], oe, rtol=1e-10)
    sig, _ = sigma_a(16)
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    print("z", (a - exact) / e, flush=True)
    assert np.all(np.abs(a - exact) < 3 * e), (a - exact) / e
    # MPIMCI combine (MPIMCI.cpp:85-92) restated on the host
    assert np.allclose(a, wa.mean(axis=0), rtol=1e-13)
    assert np.allclose(e, np.sqrt((we ** 2).sum(axis=0)) / W, rtol=1e-13)


def test_full_size_shared_coupling_physics():
    W, nmc = 65536, 10240
    g = chain_setup("mci", cfg2(), nwalkers=W)
    g.setRng(RNG_PHILOX)
    g.setCoupling(COUPLING_SHARED)
    a, e = g.integrate(nmc)
    steps = g.getWalkerSteps()
    assert np.all(steps == steps[0, 0])  # one shared calibrated step (MCIntegrator.cpp:131-138)
    acc, rej, rate = g.acceptance()
    assert acc + rej == W * nmc and abs(rate - 0.5) < 0.06
    sig, _ = sigma_a(16)
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    print("z", (a - exact) / e, flush=True)
    assert np.all(np.abs(a - exact) < 3 * e), (a - exact) / e


@pytest.mark.parametrize("rng", [RNG_MT, RNG_PHILOX])
def test_dependent_observables_many_walkers(rng):
    """Dependent observables (a sampling function's log value and proto values, the squares of an
    earlier observable) on 100 walkers with automatic calibration and decorrelation: bit-exact with
    the oracle, which matches the reference's own chains with them (depobs_* in chains.json)."""
    from mcitest import OBS_DEP_SQUARE, OBS_PDF_LOG, OBS_PDF_PROTO
    a = np.array([0.5, 1.0, 2.0, 0.25])
    cfg = dict(ndim=4, seed=2718, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, a)],
               obs=[(OBS_X2, [], 16, 2, 1, ESTIM_MJ), (OBS_PDF_PROTO, [0], 16, 4, 1, ESTIM_MJ),
                    (OBS_PDF_LOG, [0], 1, 6, 0, ESTIM_CORR), (OBS_DEP_SQUARE, [0], 8, 4, 1, ESTIM_UNCORR)])
    W, nmc = 100, 2 ** 12
    res = {}
    for prefix in ("orc", "mci"):
        c = chain_setup(prefix, cfg, nwalkers=W)
        c.setRng(rng)
        a_, e_ = c.integrate(nmc)
        res[prefix] = (a_, e_, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance())
    for k in range(len(res["orc"])):
        assert np.array_equal(res["orc"][k], res["mci"][k]), k


def test_dependent_observable_rules_on_the_device():
    from mcitest import E_INVALID, MciError, OBS_DEP_SQUARE, OBS_PDF_LOG, OBS_PDF_PROTO, PDF_GAUSS_ISO
    c = Chain("mci", 3, 64)
    c.addSamplingFunction(PDF_GAUSS_ISO)
    c.addObservable(OBS_X, [], 4, 2, 0, ESTIM_UNCORR)
    for params, nskip in (([1], 2), ([0], 3), ([], 2)):
        with pytest.raises(MciError) as ei:
            c.addObservable(OBS_DEP_SQUARE, params, 4, nskip, 0, ESTIM_UNCORR)
        assert ei.value.code == E_INVALID
    c.addObservable(OBS_PDF_LOG, [1], 4, 1, 0, ESTIM_UNCORR)
    with pytest.raises(MciError):
        c.integrate(64, False, False)
    c.addSamplingFunction(PDF_GAUSS_ISO)
    c.integrate(64, False, False)
    c.addObservable(OBS_PDF_PROTO, [0], 4, 1, 0, ESTIM_UNCORR)
    with pytest.raises(MciError):
        c.integrate(64, False, False)


@pytest.mark.parametrize("d,W,coupling", [(16, 200, COUPLING_SHARED), (5, 130, COUPLING_INDEPENDENT)])
def test_full_traces_on_the_specialised_kernel(d, W, coupling):
    """FullAccumulator traces (the reference's default addObservable(obs): blocksize 1, Correlated)
    written by the specialised kernel: every step's x_i^2, |x|^2 and exp(-|x|^2/d) of every walker
    bit-identical to the generic kernel's traces, and to the oracle's (exp to 1e-12); estimates
    equal the generic kernel's bit for bit.  d = 5 takes the Lemire index with rejections."""
    print("full traces", d, W, coupling, flush=True)
    sig = np.random.default_rng(d).uniform(0.5, 2.0, d) if d != 16 else sigma_a(16)[0]
    cfg = dict(ndim=d, seed=99, ndecorr=500, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, 0.5 / sig ** 2)],
               obs=[(OBS_X2, [], 1, 1, 1, ESTIM_CORR), (OBS_X2_SUM, [], 1, 1, 1, ESTIM_CORR), (OBS_EXP_X2, [float(d)], 1, 1, 1, ESTIM_CORR)])
    nmc = 1024
    out = {}
    for name in ("fast", "generic", "orc"):
        c = chain_setup("orc" if name == "orc" else "mci", cfg, nwalkers=W)
        c.setRng(RNG_PHILOX)
        c.setCoupling(coupling)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a, e = c.integrate(nmc)
        out[name] = (c, a, e, c.getX(), *c.walkerAcceptance())
    f, g, o = out["fast"], out["generic"], out["orc"]
    for k in (1, 2, 3, 4, 5):
        assert np.array_equal(f[k], g[k]), k
    for k in (3, 4, 5):
        assert np.array_equal(f[k], o[k]), k
    for i, nobs in enumerate((d, 1, 1)):
        tf, tg = f[0].observableData(i, nobs), g[0].observableData(i, nobs)
        assert tf.shape == (W, nmc, nobs)
        assert np.array_equal(tf, tg), i
    # the oracle's traces of the first walkers
    import ctypes as C
    from mcitest import load
    lib = load("orc")
    lib.orc_get_accu_data.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]
    for i, nobs in enumerate((d, 1, 1)):
        tf = f[0].observableData(i, nobs)
        for w in (0, 1, W - 1):
            buf = np.zeros(nmc * nobs)
            ns = C.c_int64()
            assert lib.orc_get_accu_data(o[0].ctx, i, w, buf.ctypes.data_as(C.POINTER(C.c_double)), buf.size, C.byref(ns)) == 0
            ref = buf.reshape(nmc, nobs)
            if i < 2:
                assert np.array_equal(tf[w], ref), (i, w)
            else:
                np.testing.assert_allclose(tf[w], ref, rtol=1e-12, atol=0)


def test_error_bars_calibrated_on_the_benchmark_workload():
    """The benchmark workload (65,536 walkers x 10,240 steps, d = 16, 64 MJBlocker blocks per walker,
    MPIMCI combine) integrated 12 times in a row; the 12 x 17 deviations from the analytic values in
    units of an error bar.  With the standard error of the 65,536 independent walker means the
    deviations are distributed as N(0, 1) (mean z^2 near 1, Kolmogorov-Smirnov test).  The MPIMCI
    combine of the walkers' MJBlocker errors is the reference's estimator on 64 blocks of 160 steps
    per walker, whose blocking test has little power on so few blocks: it is reported beside the
    walker spread, and must stay within 40 % of it.  (The |x|^2 component is the sum of the
    others, so the components of one run are not independent; the runs are.)"""
    from scipy import stats
    sig, a = sigma_a(16)
    cfg = cfg2()
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    W = 65536
    g = chain_setup("mci", cfg, nwalkers=W)
    g.setRng(RNG_PHILOX)
    g.setCoupling(COUPLING_SHARED)
    z_mj, z_sp, ratio = [], [], []
    for _ in range(12):
        avg, err = g.integrate(64 * 160)
        wa, _ = g.walkerResults()
        err_sp = wa.std(axis=0, ddof=1) / np.sqrt(W)
        z_mj.append((avg - exact) / err)
        z_sp.append((avg - exact) / err_sp)
        ratio.append(err_sp / err)
    z_mj, z_sp, ratio = np.concatenate(z_mj), np.concatenate(z_sp), np.concatenate(ratio)
    ks = stats.kstest(z_sp, "norm")
    print(f"benchmark workload: walker-spread error: mean z^2 {np.mean(z_sp * z_sp):.3f}, max |z| {np.max(np.abs(z_sp)):.2f}, "
          f"KS p {ks.pvalue:.3f}; MJBlocker combine: mean z^2 {np.mean(z_mj * z_mj):.3f}, max |z| {np.max(np.abs(z_mj)):.2f}; "
          f"spread / MJBlocker error {np.mean(ratio):.3f} (min {ratio.min():.3f}, max {ratio.max():.3f})", flush=True)
    assert 0.7 < np.mean(z_sp * z_sp) < 1.4
    assert ks.pvalue > 0.01
    assert np.all((0.6 < ratio) & (ratio < 1.4))


@pytest.mark.parametrize("rng", [RNG_MT, RNG_PHILOX])
def test_learned_mixed_move_many_walkers(rng):
    """The mixed trial move (a learned diagonal-mixture independence draw with probability 0.3, else
    the Gaussian single-coordinate move) on a bimodal isotropic mixture, 96 walkers with automatic
    calibration and decorrelation: identical decisions, calibrated steps and acceptance with the
    oracle, positions and estimates within 1e-12 relative (errors 1e-10)."""
    from mcitest import OBS_INDICATOR, PDF_GMM_ISO
    print("learned mixed move", rng, flush=True)
    d, W, nmc = 4, 96, 16 * 64
    mu = np.array([[-8.0, 0, 0, 0], [8.0, 0, 0, 0]])
    learned = (np.array([0.75, 0.25]), mu + 0.3, np.array([[1.2, 1.0, 0.9, 1.1], [1.5, 1.4, 1.6, 1.2]]), 0.3)
    cfg = dict(ndim=d, seed=404, ndecorr=100, move=(1, 1, 1, None), learned=learned,
               pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.8, 1.3, *mu.ravel()])],
               obs=[(OBS_INDICATOR, [0], 16, 1, 0, ESTIM_UNCORR), (OBS_X2_SUM, [], 16, 1, 0, ESTIM_UNCORR)])
    res = {}
    for prefix in ("orc", "mci"):
        c = chain_setup(prefix, cfg, nwalkers=W)
        c.setRng(rng)
        x = np.zeros((W, d))
        x[:, 0] = 8.0
        c.setX(x)
        a, e = c.integrate(nmc)
        res[prefix] = (a, e, *c.walkerResults(), c.getX(), c.getWalkerSteps(), c.walkerAcceptance())
    o, g = res["orc"], res["mci"]
    assert np.array_equal(o[6][0], g[6][0]) and np.array_equal(o[6][1], g[6][1]), "acceptance"
    assert np.array_equal(o[5], g[5]), "steps"
    for k in (0, 2, 4):
        check_close(g[k], o[k], 1e-12)
    for k in (1, 3):
        check_close(g[k], o[k], 1e-10)
    # the mode weights from walkers that all started in the heavy mode
    assert abs(g[0][0] - 0.7) < 5 * g[1][0] + 0.02, g[0]


def test_learned_mixed_move_physics_at_size():
    """The mixed move on 65,536 walkers (device only): P(x_0 > 0) = 0.7 and E|x|^2 = 64 + 4 (0.3 0.64
    + 0.7 1.69) within 3 standard errors of the walker means, from walkers all in the light mode."""
    from mcitest import OBS_INDICATOR, PDF_GMM_ISO
    d, W = 4, 65536
    mu = np.array([[-8.0, 0, 0, 0], [8.0, 0, 0, 0]])
    learned = (np.array([0.5, 0.5]), mu, np.ones((2, d)), 0.2)
    cfg = dict(ndim=d, seed=405, ndecorr=200, move=(0, 1, 1, None), learned=learned,
               pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.8, 1.3, *mu.ravel()])],
               obs=[(OBS_INDICATOR, [0], 64, 1, 0, ESTIM_UNCORR), (OBS_X2_SUM, [], 64, 1, 0, ESTIM_UNCORR)])
    c = chain_setup("mci", cfg, nwalkers=W)
    c.setRng(RNG_PHILOX)
    x = np.zeros((W, d))
    x[:, 0] = -8.0
    c.setX(x)
    c.integrate(64 * 64)
    wa, _ = c.walkerResults()
    exact = np.array([0.7, 64 + 4 * (0.3 * 0.64 + 0.7 * 1.69)])
    z = (wa.mean(axis=0) - exact) / (wa.std(axis=0, ddof=1) / np.sqrt(W))
    print("z", z, flush=True)
    assert np.all(np.abs(z) < 3), z


@pytest.mark.parametrize("d,blocksize,coupling,W", [(16, 20, COUPLING_INDEPENDENT, 256), (16, 20, COUPLING_SHARED, 256),
                                                     (4, 0, COUPLING_INDEPENDENT, 200), (8, 1, COUPLING_INDEPENDENT, 130)])
def test_reference_streams_on_the_specialised_kernel(d, blocksize, coupling, W):
    """The reference streams (one mt19937_64 per walker) on the specialised kernel, against the
    generic kernel on the same streams: with a power-of-two coordinate count a step takes three
    outputs, and the block sums are the reference's per-step sums, so everything is bit-identical
    -- positions, step sizes, counters, every block mean or trace value and the estimates -- with
    Block, Simple (blocksize 0) and Full (blocksize 1) accumulators, X2, X2_SUM and EXP_X2, both
    couplings and a ragged last wave.  (The golden reference chains of this family, e.g. cfg2_walker
    and the eight walkers of walkers.json, run through this kernel too.)"""
    from mcitest import OBS_EXP_X2
    print("reference streams, specialised vs generic", d, blocksize, coupling, W, flush=True)
    sig = np.random.default_rng(d).uniform(0.5, 2.0, d)
    est = ESTIM_MJ if blocksize > 1 else (ESTIM_NOOP if blocksize == 0 else ESTIM_MJ)
    cfg = dict(ndim=d, seed=271 + d, ndecorr=300, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, 0.5 / sig ** 2)],
               obs=[(OBS_X2, [], blocksize, 1, 1 if blocksize > 0 else 0, est), (OBS_X2_SUM, [2.0], blocksize, 1, 1 if blocksize > 0 else 0, est),
                    (OBS_EXP_X2, [float(d)], blocksize, 1, 1 if blocksize > 0 else 0, est)])
    nmc = 1024 if blocksize == 1 else 64 * 20  # the MJBlocker on Full traces takes 2^k values
    out = {}
    for name in ("fast", "generic"):
        c = chain_setup("mci", cfg, nwalkers=W)
        c.setRng(RNG_MT)
        c.setCoupling(coupling)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a, e = c.integrate(nmc)
        a2, e2 = c.integrate(nmc // 2, False, False)
        data = [c.observableData(i, n) for i, n in enumerate((d, 1, 1))]
        out[name] = (a, e, a2, e2, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance(), *data)
    f, g = out["fast"], out["generic"]
    for k in range(len(f)):
        assert np.array_equal(f[k], g[k]), k


@pytest.mark.parametrize("rng", [RNG_PHILOX, RNG_MT])
@pytest.mark.parametrize("d,blocksize,coupling,W,iso", [(16, 20, COUPLING_INDEPENDENT, 256, False), (16, 20, COUPLING_SHARED, 200, False),
                                                         (5, 0, COUPLING_INDEPENDENT, 130, False), (7, 1, COUPLING_INDEPENDENT, 100, False),
                                                         (3, 1, COUPLING_SHARED, 70, False), (3, 20, COUPLING_INDEPENDENT, 90, True),
                                                         (20, 20, COUPLING_INDEPENDENT, 100, False), (23, 1, COUPLING_SHARED, 70, False)])
def test_all_coordinate_moves_on_the_specialised_kernel(rng, d, blocksize, coupling, W, iso):
    """The reference's default move (every coordinate at once, one step size) on the specialised
    all-coordinate kernels (walker in registers up to d = 16, in LDS beyond), against the generic
    kernel on the same streams: positions, step sizes, counters, every block mean or trace value and
    the estimates bit-identical, with Block, Simple (blocksize 0) and Full (blocksize 1)
    accumulators, X2, X2_SUM and EXP_X2, both couplings, both stream modes, ragged last waves, the
    diagonal Gaussian and exp(-|x|^2); the native streams also against the oracle's walkers.  (The
    golden reference chains cfg1, ut2_sequence, ut3_sequence, cfg2_allmove_walker, allmove_full and
    allmove_simple run through these kernels as well.)"""
    from mcitest import PDF_GAUSS_ISO
    print("all-coordinate moves, specialised vs generic", rng, d, blocksize, coupling, W, iso, flush=True)
    sig = np.random.default_rng(d).uniform(0.5, 2.0, d)
    est = ESTIM_NOOP if blocksize == 0 else ESTIM_MJ
    eq = 1 if blocksize > 0 else 0
    pdf = (PDF_GAUSS_ISO, []) if iso else (PDF_GAUSS_DIAG, 0.5 / sig ** 2)
    cfg = dict(ndim=d, seed=913 + d, ndecorr=300, move=(0, 0, 1, None), pdfs=[pdf],
               obs=[(OBS_X2, [], blocksize, 1, eq, est), (OBS_X2_SUM, [2.0], blocksize, 1, eq, est),
                    (OBS_EXP_X2, [float(d)], blocksize, 1, eq, est)])
    nmc = 1024 if blocksize == 1 else 64 * 20
    out = {}
    for name in ("fast", "generic", "orc"):
        if name == "orc" and rng != RNG_PHILOX:
            continue
        c = chain_setup("orc" if name == "orc" else "mci", cfg, nwalkers=W)
        c.setRng(rng)
        c.setCoupling(coupling)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a, e = c.integrate(nmc)
        a2, e2 = c.integrate(nmc // 2, False, False)
        data = [] if name == "orc" else [c.observableData(i, n) for i, n in enumerate((d, 1, 1))]
        out[name] = (c.getX(), c.getWalkerSteps(), *c.walkerAcceptance(), a, e, a2, e2, *c.walkerResults(), *data)
    f, g = out["fast"], out["generic"]
    for k in range(len(f)):
        assert np.array_equal(f[k], g[k]), k
    if "orc" in out:
        o = out["orc"]
        for k in range(4):
            assert np.array_equal(f[k], o[k]), k


@pytest.mark.parametrize("veclen,d", [(0, 6), (1, 8)])
def test_specialised_moves_after_generator_positions_diverge(veclen, d):
    """Reference streams whose walkers sit at different generator positions (Gaussian moves take a
    varying number of outputs per step), then uniform all-coordinate (veclen 0) or single-coordinate
    (veclen 1) moves on the specialised kernels: their per-lane-position paths, bit-identical to the
    generic kernel on the same sequence."""
    from mcitest import SRRD_GAUSSIAN
    W = 150
    print("specialised moves after diverged generator positions", veclen, d, W, flush=True)
    sig = np.random.default_rng(d).uniform(0.5, 2.0, d)
    cfg = dict(ndim=d, seed=4242, ndecorr=100, move=(SRRD_GAUSSIAN, 0, 1, None), pdfs=[(PDF_GAUSS_DIAG, 0.5 / sig ** 2)],
               obs=[(OBS_X2, [], 20, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 20, 1, 1, ESTIM_MJ)])
    out = {}
    for name in ("fast", "generic"):
        c = chain_setup("mci", cfg, nwalkers=W)
        c.setRng(RNG_MT)
        if name == "generic":
            c.setKernelPath(KERNELS_GENERIC)
        a0, e0 = c.integrate(64 * 20)
        c.setTrialMove(0, veclen, 1, None)
        a, e = c.integrate(64 * 20)
        a2, e2 = c.integrate(32 * 20, False, False)
        out[name] = (a0, e0, a, e, a2, e2, *c.walkerResults(), c.getWalkerSteps(), c.getX(), *c.walkerAcceptance(),
                     *[c.observableData(i, n) for i, n in enumerate((d, 1))])
    f, g = out["fast"], out["generic"]
    for k in range(len(f)):
        assert np.array_equal(f[k], g[k]), k
