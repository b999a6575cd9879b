"""The network oracle (oracle/mlp_oracle.py) pinned by what the mathematics fixes: backprop in
exact float64 against central finite differences, the mixture head against scipy's normal
log-density, and the bfloat16 rounding against hand-made values."""
import os
import sys

import numpy as np
import pytest
from scipy.special import logsumexp
from scipy.stats import norm

from mcitest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402


def _params(d, H, K, seed):
    rng = np.random.default_rng(seed)
    n = sum(int(np.prod(s)) for _, s in M.sizes(d, H, K))
    return rng.normal(0.0, 0.4, n)


@pytest.mark.parametrize("scale", [1.0, 4.0])
def test_backprop_matches_finite_differences(scale):
    d, H, K, n = 3, 8, 2, 6
    rng = np.random.default_rng(1)
    p = _params(d, H, K, 2)
    p[-K * (1 + 2 * d):] *= scale  # scale 4: output biases that push log-scales beyond the clamp
    x, y = rng.normal(size=(n, d)) * 2, rng.normal(size=(n, d)) * 2
    ol = M._split(M.forward(p, x, d, H, K, False)[0]["OUT"], d, K)[2]
    assert np.min(np.abs(np.abs(ol) - M.LS_MAX)) > 1e-3  # no output at the clamp's kink
    if scale > 1.0:
        assert np.any(np.abs(ol) > M.LS_MAX) and np.any(np.abs(ol) < M.LS_MAX)
    loss, g = M.loss_grad(p, x, y, d, H, K, emulate=False)
    num = np.zeros_like(p)
    h = 1e-6
    for i in range(p.size):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num[i] = (M.loss_grad(pp, x, y, d, H, K, emulate=False)[0] - M.loss_grad(pm, x, y, d, H, K, emulate=False)[0]) / (2 * h)
    assert np.max(np.abs(g - num)) < 1e-6 * max(1.0, np.max(np.abs(num)))
    assert np.isclose(loss, -np.mean(M.mixture_logq(M.forward(p, x, d, H, K, False)[0]["OUT"], y, d, K)), rtol=1e-14)


def test_mixture_density_matches_scipy():
    d, K, n = 4, 3, 5
    rng = np.random.default_rng(3)
    out = rng.normal(size=(n, K * (1 + 2 * d))) * 2.5  # some log-scales beyond the clamp
    y = rng.normal(size=(n, d)) * 3
    got = M.mixture_logq(out, y, d, K)
    a = out[:, 2 * d * K:]
    pi = np.exp(a - a.max(1, keepdims=True))
    pi /= pi.sum(1, keepdims=True)
    comp = out[:, :2 * d * K].reshape(n, K, 2 * d)
    mu = 4.0 * comp[:, :, :d]
    assert np.any(np.abs(comp[:, :, d:]) > 3.0)
    sig = np.exp(np.clip(comp[:, :, d:], -3.0, 3.0))
    want = logsumexp(np.log(pi) + np.sum(norm.logpdf(y[:, None, :], mu, sig), axis=2), axis=1)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2.0 ** -7       # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2.0 ** -8 * 3.0e-3  # half an ulp of 8 significant bits
