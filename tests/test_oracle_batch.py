"""The oracle of the batched mixture integrals, pinned by what the mathematics fixes: its FP32
exp/log against libm in double, the analytic E|x|^2 of each mixture, the FP64 sums of the same
samples (Kahan accuracy), the textbook standard error of independent walker means, and the
planner's invariants (host logic of the product library, no GPU needed)."""
import numpy as np
import pytest

from batchtest import _orc, orc_batch, orc_walker_order, small_mixtures


def _ulps(got, ref):
    ref32 = ref.astype(np.float32)
    spacing = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / spacing


def test_fp32_exp_within_one_ulp():
    lib = _orc()
    x = np.concatenate([np.linspace(-87.0, 88.0, 200001), np.linspace(-1e-3, 1e-3, 2001), [0.0]]).astype(np.float32)
    got = np.array([lib.orc_bexp(float(v)) for v in x], dtype=np.float32)
    assert np.max(_ulps(got, np.exp(x.astype(np.float64)))) <= 1.0
    assert lib.orc_bexp(0.0) == 1.0 and lib.orc_bexp(-100.0) == 0.0 and np.isinf(lib.orc_bexp(89.0))


def test_fp32_log_within_two_ulp():
    lib = _orc()
    bits = np.linspace(0x00800000, 0x7f000000, 200001).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.abs(x - 1.0) > 1e-3]  # relative error of log is unbounded only at x = 1 exactly
    got = np.array([lib.orc_blog(float(v)) for v in x], dtype=np.float32)
    assert np.max(_ulps(got, np.log(x.astype(np.float64)))) <= 2.0
    assert lib.orc_blog(1.0) == 0.0


def _plan(nmc, nblocks=16, rmin=16):
    # the planner's rule for one cap: T = blocks x 2^k, the largest with T <= nmc / rmin
    T = []
    for n in nmc:
        t = nblocks
        while t * 2 <= max(n // rmin, nblocks):
            t *= 2
        T.append(t)
    return np.array(T, np.int64)


def test_batch_oracle_physics_and_kahan():
    ndim, nmc, w, s, mu, exact = small_mixtures(n=16, seed=3, dmax=24, nmc=(2000, 8000))
    T = _plan(nmc)
    o = orc_batch(ndim, nmc, w, s, mu, T, seed=11, fp64_shadow=True)
    z = (o["avg"] - exact) / o["err"]
    assert np.all(np.abs(z) < 3), z
    assert 0.3 < np.mean(z * z) < 2.5
    # Kahan-compensated FP32 block sums against FP64 sums of the same FP32 samples
    rel = np.abs(o["bmeans"] - o["bmeans_fp64"]) / np.abs(o["bmeans_fp64"])
    assert np.max(rel) < 1e-6, np.max(rel)
    # calibrated steps and rates
    assert np.all(o["steps"] > 0) and np.all(np.abs(o["calib_rate"] - 0.5) < 0.05)
    assert np.all((o["accrate"] > 0.3) & (o["accrate"] < 0.7))


def test_batch_oracle_error_is_standard_error_of_walker_means():
    ndim, nmc, w, s, mu, _ = small_mixtures(n=8, seed=5, dmax=8, nmc=(1500, 5000))
    T = _plan(nmc)
    o = orc_batch(ndim, nmc, w, s, mu, T, seed=2)
    for i in range(len(ndim)):
        a = o["wavg"][o["off"][i]:o["off"][i] + o["R"][i]]
        assert np.isclose(o["avg"][i], a.mean(), rtol=1e-13)
        assert np.isclose(o["err"][i], a.std(ddof=1) / np.sqrt(len(a)), rtol=1e-10)
    # the MPIMCI combine of per-walker MJBlocker errors (src/MPIMCI.cpp:85-92)
    o2 = orc_batch(ndim, nmc, w, s, mu, T, seed=2, mode=1)
    assert np.array_equal(o2["avg"], o["avg"])
    for i in range(len(ndim)):
        e = o["werr"][o["off"][i]:o["off"][i] + o["R"][i]]
        assert np.isclose(o2["err"][i], np.sqrt(np.sum(e * e)) / len(e), rtol=1e-12)


def test_batch_oracle_block_means_average_to_walker_mean():
    ndim, nmc, w, s, mu, _ = small_mixtures(n=6, seed=9, dmax=6, nmc=(1000, 3000))
    T = _plan(nmc)
    o = orc_batch(ndim, nmc, w, s, mu, T, seed=4)
    assert np.allclose(o["bmeans"].mean(axis=0), o["wavg"], rtol=1e-12)


def test_planner_host_logic():
    # the product's planner (no device needed for the cost model)
    from mcintegratorplusplus_amd import BatchIntegrals, mixture_workload
    ndim, nmc, w, s, mu, _ = mixture_workload(n=512, seed=1)
    b = BatchIntegrals(ndim, nmc, w, s, mu)
    b.autotune(0)
    R, T, off = b.plan()
    assert np.all(R * T >= nmc)
    assert np.all(T % 16 == 0) and np.all(((T // 16) & (T // 16 - 1)) == 0)
    assert np.all((R >= 64) | (T == 16))
    # walker order: integrals sorted by (T, d, K) descending, as the oracle orders them
    R2, off2, W = orc_walker_order(ndim, np.array([len(x) for x in w]), nmc, T)
    assert np.array_equal(R, R2) and np.array_equal(off, off2) and W == b.nwalkers()
    # an explicit plan is validated
    with pytest.raises(Exception):
        b.setPlan(np.full(len(ndim), 24))
    b.setPlan(np.full(len(ndim), 64))
    R3, T3, _ = b.plan()
    assert np.all(T3 == 64) and np.all(R3 == (nmc + 63) // 64)


def test_batch_create_validates_inputs():
    from mcintegratorplusplus_amd import BatchIntegrals, MciError
    with pytest.raises(MciError):
        BatchIntegrals([3], [1000], [[1.0]], [[-1.0]], [[[0, 0, 0]]])
    with pytest.raises(MciError):
        BatchIntegrals([300], [1000], [[1.0]], [[1.0]], [np.zeros((1, 300))])
    with pytest.raises(MciError):
        BatchIntegrals([2], [1000], [[1.0] * 5], [[1.0] * 5], [np.zeros((5, 2))])
