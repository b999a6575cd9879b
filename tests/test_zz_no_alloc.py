"""mci_estimate_device, the allocation-free entry point, against mci_estimate on the same data: equal
averages and errors bit for bit, on two calls in a row (the second with a smaller input, so that
the shared scratch array still holds the first call's values), for the estimator that uses the
scratch array (the general MJBlocker, more than 64 values per component) and for the others.
Device buffers come from the HIP runtime the library itself links (not torch's own runtime)."""
import ctypes as C

import numpy as np
import pytest

from mcitest import ESTIM_CORR, ESTIM_FC, ESTIM_MJ, ESTIM_UNCORR, MciError, estimate

pytestmark = pytest.mark.gpu

_DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def hip():
    from mcintegratorplusplus_amd import _lib
    lib = C.CDLL(_lib.library_path())
    rt = C.CDLL("libamdhip64.so.7")  # the runtime the library is linked against (already loaded with it)
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    lib.mci_estimate_device.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mci_estimate_device.restype = C.c_int
    lib.mci_last_error.restype = C.c_char_p
    return lib, rt


def device_estimate(hip, estim, x):
    lib, rt = hip
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, nd = x.shape
    bufs = [C.c_void_p() for _ in range(3)]
    sizes = [x.nbytes, 8 * nd, 8 * nd]
    for b, sz in zip(bufs, sizes):
        assert rt.hipMalloc(C.byref(b), sz) == 0
    try:
        assert rt.hipMemcpy(bufs[0], x.ctypes.data, x.nbytes, 1) == 0  # hipMemcpyHostToDevice
        st = lib.mci_estimate_device(estim, n, nd, bufs[0], bufs[1], bufs[2], None)
        if st != 0:
            raise MciError(st, lib.mci_last_error().decode())
        assert rt.hipDeviceSynchronize() == 0
        a, e = np.zeros(nd), np.zeros(nd)
        assert rt.hipMemcpy(a.ctypes.data, bufs[1], 8 * nd, 2) == 0  # hipMemcpyDeviceToHost
        assert rt.hipMemcpy(e.ctypes.data, bufs[2], 8 * nd, 2) == 0
        return a, e
    finally:
        for b in bufs:
            rt.hipFree(b)


def ar1(n, nd, seed):
    rng = np.random.default_rng(seed)
    eps = rng.normal(size=(n, nd))
    x = np.zeros((n, nd))
    for i in range(1, n):
        x[i] = 0.85 * x[i - 1] + eps[i]
    return x + np.arange(nd)


@pytest.mark.parametrize("estim", [ESTIM_MJ, ESTIM_CORR, ESTIM_FC, ESTIM_UNCORR])
def test_no_alloc_estimate_matches_mci_estimate_twice(hip, estim):
    print("no-alloc estimator", estim, flush=True)
    sizes = [(4096, 3), (1024, 2)] if estim in (ESTIM_MJ, ESTIM_CORR) else [(3000, 3), (777, 2)]
    for k, (n, nd) in enumerate(sizes):
        x = ar1(n, nd, 17 + k)
        a_dev, e_dev = device_estimate(hip, estim, x)
        a_ref, e_ref = estimate("mci", estim, x)
        assert np.array_equal(a_dev, a_ref) and np.array_equal(e_dev, e_ref), (estim, n, nd)
        assert np.all(e_dev > 0)


def test_no_alloc_estimate_rejects_beyond_its_array(hip):
    lib, _ = hip
    # checked before any device work: no buffer is touched
    st = lib.mci_estimate_device(ESTIM_MJ, 1 << 20, 4, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None)
    assert st == 1 and b"at most" in lib.mci_last_error()
    st = lib.mci_estimate_device(ESTIM_MJ, 1000, 1, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None)
    assert st == 1 and b"power of two" in lib.mci_last_error()
