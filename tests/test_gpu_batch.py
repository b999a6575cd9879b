"""Batched mixture integrals on the GPU: bit-exact against the oracle on small batches, graph
replays equal to eager runs, and the full 4,096-integral set against the analytic values."""
import numpy as np
import pytest

from batchtest import orc_batch, small_mixtures

pytestmark = pytest.mark.gpu


def _gpu(ndim, nmc, w, s, mu, seed, **kw):
    from mcintegratorplusplus_amd import BatchIntegrals
    b = BatchIntegrals(ndim, nmc, w, s, mu)
    b.setSeed(seed)
    for k, v in kw.items():
        getattr(b, k)(*v) if isinstance(v, tuple) else getattr(b, k)(v)
    return b


@pytest.mark.parametrize("mode,ndecorr,nblocks", [(0, 0, 16), (1, 37, 8)])
def test_batch_bit_exact_with_oracle(mode, ndecorr, nblocks):
    ndim, nmc, w, s, mu, exact = small_mixtures(n=40, seed=13, dmax=70, nmc=(1000, 20000))
    print("batch parity: 40 integrals, mode", mode, "ndecorr", ndecorr, "blocks", nblocks, flush=True)
    b = _gpu(ndim, nmc, w, s, mu, 99, setErrorMode=mode, setNdecorrelationSteps=ndecorr, setBlocks=nblocks,
             setGraph=False)
    avg, err = b.run()
    R, T, off = b.plan()
    o = orc_batch(ndim, nmc, w, s, mu, T, seed=99, nblocks=nblocks, ndecorr=ndecorr, mode=mode)
    assert np.array_equal(R, o["R"]) and np.array_equal(off, o["off"])
    steps, crate = b.steps()
    assert np.array_equal(steps, o["steps"]) and np.array_equal(crate, o["calib_rate"])
    assert np.array_equal(b.blockMeans(nblocks), o["bmeans"])
    wa, we = b.walkerResults()
    assert np.array_equal(wa, o["wavg"]) and np.array_equal(we, o["werr"])
    assert np.array_equal(b.acceptance(), o["accrate"])
    assert np.array_equal(avg, o["avg"]) and np.array_equal(err, o["err"])
    # physics with the standard error of the walker means (the per-walker MJBlocker errors of the
    # MPIMCI combine cannot see a walker's mode: walkers start in a mode drawn from the weights
    # and rarely leave it, so that combine understates the error of a multimodal integral)
    se = np.array([wa[off[i]:off[i] + R[i]].std(ddof=1) / np.sqrt(R[i]) for i in range(len(R))])
    z = (avg - exact) / se
    assert np.all(np.abs(z) < 5), z


@pytest.mark.parametrize("dmax", [40, 200])
def test_batch_graph_replays_equal_eager(dmax):
    """Graph replays equal the eager run bit for bit.  With dmax 200 the integrals fall in all
    three dimension classes, whose launches the graph runs as parallel branches."""
    ndim, nmc, w, s, mu, _ = small_mixtures(n=32, seed=21, dmax=dmax, nmc=(1000, 8000))
    print("graph replay: 32 integrals, dmax", dmax, flush=True)
    eager = _gpu(ndim, nmc, w, s, mu, 5, setGraph=False)
    a0, e0 = eager.run()
    st0 = eager.steps()
    bm0 = eager.blockMeans(16)
    g = _gpu(ndim, nmc, w, s, mu, 5, setGraph=True)
    other = _gpu(ndim[:8], nmc[:8], w[:8], s[:8], mu[:8], 6, setGraph=False)
    for _ in range(4):
        a, e = g.run()
        assert np.array_equal(a, a0) and np.array_equal(e, e0)
        st = g.steps()
        assert np.array_equal(st[0], st0[0]) and np.array_equal(st[1], st0[1])
        assert np.array_equal(g.blockMeans(16), bm0)
        other.run()  # unrelated eager work between replays
    assert g.timings()[3] > 0


def test_batch_full_size_against_analytic():
    from mcintegratorplusplus_amd import mixture_workload
    ndim, nmc, w, s, mu, exact = mixture_workload(n=4096, seed=4096)
    print("full batch: 4096 integrals", flush=True)
    b = _gpu(ndim, nmc, w, s, mu, 1337)
    b.autotune(0)
    avg, err = b.run()
    R, T, _ = b.plan()
    assert np.all(R * T >= nmc)
    z = (avg - exact) / err
    print("walkers", b.nwalkers(), "mean z^2", np.mean(z * z), "max |z|", np.max(np.abs(z)), "ms", b.timings(), flush=True)
    assert np.mean(np.abs(z) < 3) > 0.99
    assert 0.8 < np.mean(z * z) < 1.25
    acc = b.acceptance()
    assert np.all((acc > 0.25) & (acc < 0.75))


def test_batch_autotune_measured():
    from mcintegratorplusplus_amd import mixture_workload
    ndim, nmc, w, s, mu, exact = mixture_workload(n=512, seed=77)
    print("autotune: 512 integrals", flush=True)
    b = _gpu(ndim, nmc, w, s, mu, 3)
    b.autotune(3)
    R, T, _ = b.plan()
    assert np.all(R * T >= nmc)
    avg, err = b.run()
    z = (avg - exact) / err
    assert np.mean(np.abs(z) < 3) > 0.97
