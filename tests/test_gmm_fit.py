"""The EM fit the learned proposal's network starts from (host code in the library, no device):
pinned by what the mathematics fixes — one component is the sample mean and variance, the fit is
a fixed point of textbook EM, the returned log-likelihood is the mixture density's, and on eight
clusters about 24 apart every cluster gets its own component (greedy k-means++ seeding)."""
import numpy as np
import pytest
from scipy.special import logsumexp


def _draw(n, seed):
    from mcintegratorplusplus_amd.learned import mixture_target
    w, mu, cov, _ = mixture_target(d=32, ncomp=8, seed=5)
    rng = np.random.default_rng(seed)
    k = rng.choice(len(w), size=n, p=w)
    L = np.linalg.cholesky(cov)
    return w, mu, cov, mu[k] + np.einsum("nij,nj->ni", L[k], rng.normal(size=(n, mu.shape[1])))


def _loglik(x, mu, var, pi):
    lp = np.log(pi)[None] - 0.5 * np.sum(np.log(2 * np.pi * var)[None] + (x[:, None, :] - mu[None]) ** 2 / var[None], axis=2)
    return lp, logsumexp(lp, axis=1)


def test_one_component_is_the_sample_moments():
    from mcintegratorplusplus_amd.learned import gmm_fit_diag
    x = np.random.default_rng(3).normal(1.5, 2.0, (999, 5)) * np.arange(1, 6)
    mu, var, pi, ll = gmm_fit_diag(x, 1, 1, 7)
    # (the component's count carries a 1e-9 guard against empty components)
    assert np.allclose(mu[0], x.mean(axis=0), rtol=1e-10, atol=0)
    assert np.allclose(var[0], x.var(axis=0) + 1e-6, rtol=1e-8)
    assert pi[0] == pytest.approx(1.0)
    assert ll == pytest.approx(float(np.mean(_loglik(x, mu, var, pi)[1])), rel=1e-12)


def test_separated_clusters_each_get_a_component_and_the_fit_is_an_em_fixed_point():
    from mcintegratorplusplus_amd.learned import gmm_fit_diag
    w, mu_t, cov, x = _draw(16384, 1)
    for seed in (1, 2, 3):
        mu, var, pi, ll = gmm_fit_diag(x, 8, 60, seed)
        order = [int(np.argmin(np.sum((mu - m) ** 2, axis=1))) for m in mu_t]
        assert sorted(order) == list(range(8)), (seed, order)
        for k, j in enumerate(order):
            # 2,048 samples per cluster: standard errors about 0.02 (means), 3 % (variances), 0.003 (weights)
            assert np.max(np.abs(mu[j] - mu_t[k])) < 0.15
            assert np.max(np.abs(var[j] / np.diag(cov[k]) - 1)) < 0.2
            assert abs(pi[j] - w[k]) < 0.015
        # one more EM iteration, written from the textbook update, leaves the fit where it is
        lp, lse = _loglik(x, mu, var, pi)
        r = np.exp(lp - lse[:, None])
        nk = r.sum(axis=0)
        m1 = (r.T @ x) / nk[:, None]
        v1 = (r.T @ (x * x)) / nk[:, None] - m1 * m1 + 1e-6
        assert np.max(np.abs(m1 - mu)) < 1e-6 and np.max(np.abs(v1 / var - 1)) < 1e-6
        assert np.max(np.abs(nk / len(x) - pi)) < 1e-8
        assert ll == pytest.approx(float(np.mean(lse)), rel=1e-10)


def test_fit_is_deterministic_and_validates_its_arguments():
    from mcintegratorplusplus_amd.learned import gmm_fit_diag
    from mcintegratorplusplus_amd.mci import MciError
    x = np.random.default_rng(4).normal(size=(300, 3))
    a = gmm_fit_diag(x, 3, 20, 11)
    b = gmm_fit_diag(x, 3, 20, 11)
    for u, v in zip(a[:3], b[:3]):
        assert np.array_equal(u, v)
    with pytest.raises(MciError):
        gmm_fit_diag(x[:2], 3, 20, 11)
    with pytest.raises(MciError):
        gmm_fit_diag(x, 3, 1001, 11)
