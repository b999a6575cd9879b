"""The oracle of the learned independence move (oracle/mci_lp_oracle.c) pinned by what the
mathematics fixes: its sine/cosine against numpy, its Box-Muller normals distributed as N(0, 1),
and the move itself leaving the target invariant (walkers that start from exact draws of the
target stay distributed as the target after many moves of a poor proposal; a wrong sign in the
Hastings factor or a proposal density that does not match the draw fails this)."""
import ctypes as C

import numpy as np
import pytest
from scipy import stats

from mcitest import load

_FP, _DP = C.POINTER(C.c_float), C.POINTER(C.c_double)


def lib():
    L = load("orc")
    L.orc_bsincos2pi.argtypes = [C.c_float, _FP, _FP]
    L.orc_normal4.argtypes = [C.c_uint32] * 6 + [_FP]
    L.orc_lp_indep_move.argtypes = [C.c_int, C.c_int, _FP, C.c_int, _DP, _DP, _DP, C.c_uint64, C.c_int64, _FP, C.c_int,
                                    C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_uint8), _DP]
    L.orc_lp_indep_move.restype = C.c_int
    return L


def test_sincos_2pi_u():
    L = lib()
    u = np.concatenate([np.arange(0, 1 << 24, 4099), [0, 1 << 22, 1 << 23, 3 << 22, (1 << 24) - 1]]).astype(np.float64) * 2.0 ** -24
    s, c = C.c_float(), C.c_float()
    err = 0.0
    for v in u:
        L.orc_bsincos2pi(float(v), C.byref(s), C.byref(c))
        err = max(err, abs(s.value - np.sin(2 * np.pi * v)), abs(c.value - np.cos(2 * np.pi * v)))
    assert err < 3e-7, err


def test_box_muller_normals():
    L = lib()
    z = np.zeros(4, np.float32)
    out = []
    for i in range(25000):
        L.orc_normal4(0x10000001 + (i % 7), i, 3, 0, 1234, 0, z.ctypes.data_as(_FP))
        out.append(z.copy())
    x = np.concatenate(out).astype(np.float64)
    assert abs(x.mean()) < 4 / np.sqrt(len(x))
    assert abs(x.var() - 1) < 4 * np.sqrt(2 / len(x))
    assert stats.kstest(x, "norm").pvalue > 1e-3


def mixture(d, KT, seed):
    rng = np.random.default_rng(seed)
    w = rng.dirichlet(np.ones(KT))
    mu = rng.normal(0, 2.0, (KT, d))
    cov = np.zeros((KT, d, d))
    for k in range(KT):
        a = rng.normal(size=(d, d)) * 0.3
        cov[k] = a @ a.T + np.diag(rng.uniform(0.3, 1.0, d))
    prec = np.linalg.inv(cov)
    c = np.log(w) - 0.5 * np.array([np.linalg.slogdet(2 * np.pi * cv)[1] for cv in cov])
    return w, mu, cov, prec, c


def proposal(d, K, target, seed, scale):
    """a diagonal mixture near the target, in the layout of mci_lp_indep_move: mu, 1/sigma, sigma
    [8][d], lc [8], cw [8]; components k cover target component k mod KT, shifted and widened"""
    w, tmu, cov, _, _ = target
    KT = len(w)
    rng = np.random.default_rng(seed)
    mu = np.zeros((8, d), np.float32)
    sg = np.ones((8, d), np.float32)
    for k in range(K):
        mu[k] = tmu[k % KT] + rng.normal(0, 0.3, d)
        sg[k] = np.sqrt(np.diag(cov[k % KT])) * scale * rng.uniform(0.8, 1.2, d)
    pi = rng.dirichlet(4 * np.ones(K))
    isg = (1 / sg).astype(np.float32)
    lc = np.full(8, -np.inf, np.float32)
    lc[:K] = np.log(pi) + np.sum(np.log(isg[:K].astype(np.float64)), axis=1) - 0.5 * d * np.log(2 * np.pi)
    cw = np.ones(8, np.float32)
    cw[:K] = np.cumsum(pi)
    cw[K - 1:] = 1.0
    return np.concatenate([mu.ravel(), isg.ravel(), sg.ravel(), lc, cw]).astype(np.float32)


def run(L, d, K, mix, target, x, nl, seed=77, t0=5):
    w, mu, cov, prec, c = target
    n = x.shape[1]
    xx = np.ascontiguousarray(x, np.float32)
    acc = np.zeros(n, np.int32)
    dec = np.zeros(n * nl, np.uint8)
    mar = np.zeros(n * nl)
    st = L.orc_lp_indep_move(d, K, mix.ctypes.data_as(_FP), len(w), np.ascontiguousarray(mu).ctypes.data_as(_DP),
                             np.ascontiguousarray(prec).ctypes.data_as(_DP), np.ascontiguousarray(c).ctypes.data_as(_DP),
                             seed, n, xx.ctypes.data_as(_FP), nl, t0, acc.ctypes.data_as(C.POINTER(C.c_int32)),
                             dec.ctypes.data_as(C.POINTER(C.c_uint8)), mar.ctypes.data_as(_DP))
    assert st == 0
    return xx, acc, dec.reshape(n, nl), mar.reshape(n, nl)


def exact_draws(target, n, seed):
    w, mu, cov, _, _ = target
    rng = np.random.default_rng(seed)
    k = rng.choice(len(w), n, p=w)
    L = np.linalg.cholesky(cov)
    z = rng.normal(size=(n, mu.shape[1]))
    return (mu[k] + np.einsum("nij,nj->ni", L[k], z)).T.copy()


@pytest.mark.parametrize("d,KT,K", [(8, 3, 4), (5, 2, 2)])
def test_independence_move_leaves_the_target_invariant(d, KT, K):
    L = lib()
    target = mixture(d, KT, seed=d)
    mix = proposal(d, K, target, seed=d + 1, scale=1.3)
    n, nl = 4000, 40
    x0 = exact_draws(target, n, seed=3)
    x1, acc, dec, mar = run(L, d, K, mix, target, x0, nl)
    rate = acc.sum() / (n * nl)
    assert 0.01 < rate < 0.9, rate
    assert np.array_equal(acc, dec.sum(axis=1)) and np.array_equal(dec == 1, mar > 0)
    # the target's moments: E|x|^2 and E x_0 under the mixture; the walkers are independent
    w, mu, cov, _, _ = target
    ex2 = np.sum(w * (np.sum(mu * mu, axis=1) + np.trace(cov, axis1=1, axis2=2)))
    ex0 = np.sum(w * mu[:, 0])
    r2 = np.sum(x1.astype(np.float64) ** 2, axis=0)
    z2 = (r2.mean() - ex2) / (r2.std() / np.sqrt(n))
    z0 = (x1[0].mean() - ex0) / (x1[0].std() / np.sqrt(n))
    assert abs(z2) < 4 and abs(z0) < 4, (z2, z0)
