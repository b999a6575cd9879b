"""The learned proposal on the GPU: the BF16 MFMA network forward and gradient against the CPU
oracle (oracle/mlp_oracle.py, same bfloat16 rounding points), and the sampler: plain chains and
the mixed chain (untrained and trained network) unbiased against the analytic values, with the
trained network's gain in error bar at equal steps."""
import os
import sys

import numpy as np
import pytest

from mcitest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("d,H,K,n", [(32, 128, 8, 512), (8, 64, 3, 300)])
def test_network_forward_and_gradient_match_oracle(d, H, K, n):
    from mcintegratorplusplus_amd.learned import Network
    print("network", d, H, K, n, flush=True)
    net = Network(d, H, K, seed=7)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(11)
    x = rng.normal(0.0, 3.0, (n, d)).astype(np.float32)
    y = (x + rng.normal(0.0, 1.0, (n, d))).astype(np.float32)
    out = net.forward(x)
    f, _ = M.forward(p, x, d, H, K, emulate=True)
    assert _rel(out, f["OUT"]) < 2e-3
    loss, g = net.lossGrad(x, y)
    lo, go = M.loss_grad(p, x, y, d, H, K, emulate=True)
    assert abs(loss - lo) < 1e-3 * abs(lo)
    off = 0
    for name, shp in M.sizes(d, H, K):
        m = int(np.prod(shp))
        assert _rel(g[off:off + m], go[off:off + m]) < 2e-2, name
        off += m
    # parameters round-trip
    q = p.astype(np.float32) * 0.5
    net.setParams(q)
    assert np.array_equal(net.params(), q)


@pytest.mark.parametrize("d,H,K,n", [(32, 128, 8, 1000), (8, 64, 3, 200), (16, 128, 5, 130)])
def test_fused_move_network_matches_oracle(d, H, K, n):
    """the network as the learned move's fused kernel evaluates it (last layer in registers, rows
    reordered as (mean, log-scale) pairs, logits from a separate block) against the oracle"""
    from mcintegratorplusplus_amd.learned import LearnedProposalSampler, mixture_target
    print("fused network", d, H, K, n, flush=True)
    w, mu, cov, _ = mixture_target(d=d, ncomp=3, seed=2)
    s = LearnedProposalSampler(w, mu, cov, 256)
    s.setNetwork(H, K)
    p = s.network().params().astype(np.float64)
    rng = np.random.default_rng(d + K)
    x = rng.normal(0.0, 3.0, (n, d)).astype(np.float32)
    got = s.evalNetwork(x)
    f, _ = M.forward(p, x, d, H, K, emulate=True)
    out = f["OUT"]
    assert got.shape == out.shape
    assert np.all(np.isfinite(got))
    # the move keeps (mean, 1 / sigma) as halves rounded toward zero (relative error < 2^-10), so
    # the means agree to that and the log-scales to about 2^-10 absolute; the logits are FP32
    for k in range(K):
        b = 2 * d * k
        mu_ref = M.MU_SCALE * out[:, b:b + d]
        ls_ref = np.clip(out[:, b + d:b + 2 * d], -M.LS_MAX, M.LS_MAX)
        mu, ls = got[:, b:b + d], got[:, b + d:b + 2 * d]
        assert np.all(np.abs(mu - mu_ref) <= 2.0 ** -9 * np.abs(mu_ref) + 2e-3 * np.abs(mu_ref).max()), k
        assert np.max(np.abs(ls - ls_ref)) < 3e-3, k
    assert _rel(got[:, 2 * d * K:], out[:, 2 * d * K:]) < 2e-3


def _small_target():
    from mcintegratorplusplus_amd.learned import mixture_target
    return mixture_target(d=8, ncomp=3, seed=3, mean_sd=2.0, eig=(0.2, 1.0))


def _z(avg, err, exact):
    return (np.asarray(avg) - np.asarray(exact)) / np.asarray(err)


def test_sampler_plain_and_untrained_mixed_are_unbiased():
    from mcintegratorplusplus_amd.learned import MIXED_FP32, PLAIN_FP32, PLAIN_FP64, LearnedProposalSampler
    w, mu, cov, exact = _small_target()
    s = LearnedProposalSampler(w, mu, cov, 8192)
    s.setSeed(21)
    s.setNetwork(64, 3)
    for mode in (PLAIN_FP64, PLAIN_FP32, MIXED_FP32):
        print("sampler mode", mode, flush=True)
        avg, err = s.run(mode, 2048, 32, 0.5)
        z = _z(avg, err, exact)
        r = s.results()
        print(mode, avg, err, z, r["acc_rw"], r["acc_learned"], s.timings(), flush=True)
        assert np.all(np.abs(z) < 3), (mode, z)
        assert 0.3 < r["acc_rw"] < 0.7


def test_ragged_walker_count_and_small_dimension():
    """1,000 walkers (the last workgroup of the learned moves holds 40 live walkers of 64) and
    d = 16 with 5 network components: every chain stays unbiased, the independence chain (the
    network's mixture at the snapshot mean) included."""
    from mcintegratorplusplus_amd.learned import INDEP_FP32, MIXED_FP32, PLAIN_FP32, mixture_target, LearnedProposalSampler
    w, mu, cov, exact = mixture_target(d=16, ncomp=4, seed=9, mean_sd=2.0, eig=(0.2, 1.0))
    s = LearnedProposalSampler(w, mu, cov, 1000)
    s.setSeed(3)
    s.setNetwork(128, 5)
    s.collect(4, 128)
    s.train(100, 1024, 2e-3)
    for mode, pm in ((PLAIN_FP32, 0.5), (MIXED_FP32, 0.5), (MIXED_FP32, 1.0), (INDEP_FP32, 0.5), (INDEP_FP32, 1.0)):
        avg, err = s.run(mode, 4096, 32, pm)
        r = s.results()
        print("ragged", mode, pm, avg, err, r["acc_learned"], "z", _z(avg, err, exact), flush=True)
        assert np.all(np.isfinite(avg)) and np.all(err > 0)
        assert np.all(np.abs(_z(avg, err, exact)) < 3), (mode, pm)
    m = s.walkerMeans()
    assert m.shape == (2, 1000) and np.all(np.isfinite(m))


def test_em_start_fits_the_target_mixture():
    """The EM start of the network's last layer on samples of a 3-component target whose
    components lie about 8 apart: before any training, the network's mixture at any input has,
    up to the order of its components, the target's weights (1/3), means and per-coordinate
    log-scales (0.5 log of the covariance diagonal), to within the sampling error of 2,048
    walkers and the 0.02-scaled input dependence of the last layer."""
    from mcintegratorplusplus_amd.learned import LearnedProposalSampler
    w, mu, cov, exact = _small_target()
    s = LearnedProposalSampler(w, mu, cov, 2048)
    s.setSeed(7)
    s.setNetwork(64, 3)
    s.collect(2, 256)
    s.initFromSamples(60)
    d, K = 8, 3
    x = np.random.default_rng(0).normal(0.0, 2.0, (16, d))
    out = s.evalNetwork(x)
    for r in range(x.shape[0]):
        o = out[r]
        means = np.array([o[2 * d * k: 2 * d * k + d] for k in range(K)])
        ls = np.array([o[2 * d * k + d: 2 * d * k + 2 * d] for k in range(K)])
        lg = o[2 * d * K: 2 * d * K + K]
        pi = np.exp(lg - lg.max())
        pi /= pi.sum()
        order = [int(np.argmin(np.sum((means - m) ** 2, axis=1))) for m in mu]
        assert sorted(order) == list(range(K)), order
        for k, j in enumerate(order):
            assert np.max(np.abs(means[j] - mu[k])) < 0.3, (r, k, means[j], mu[k])
            assert np.max(np.abs(ls[j] - 0.5 * np.log(np.diag(cov[k])))) < 0.25, (r, k)
            assert abs(pi[j] - w[k]) < 0.06, (r, k, pi)


def test_trained_network_reduces_the_error_bar():
    from mcintegratorplusplus_amd.learned import INDEP_FP32, MIXED_FP32, PLAIN_FP32, LearnedProposalSampler
    w, mu, cov, exact = _small_target()
    s = LearnedProposalSampler(w, mu, cov, 8192)
    s.setSeed(5)
    s.setNetwork(64, 3)
    print("collect + train", flush=True)
    s.collect(8, 256)
    l0, l1 = s.train(400, 2048, 2e-3)
    print("loss", l0, l1, s.timings(), flush=True)
    assert l1 < l0
    a_p, e_p = s.run(PLAIN_FP32, 2048, 32)
    a_m, e_m = s.run(MIXED_FP32, 2048, 32, 0.5)
    r = s.results()
    print("plain", a_p, e_p, "mixed", a_m, e_m, "learned acceptance", r["acc_learned"], "z", _z(a_m, e_m, exact), flush=True)
    assert np.all(np.abs(_z(a_m, e_m, exact)) < 3)
    assert r["acc_learned"] > 0.05
    assert e_m[0] < 0.7 * e_p[0]
    # independence proposal: the network trained with one fixed input
    s.setTrainingPairs(0.0)
    s.setFixedInput(True)
    s.collect(8, 256)
    s.train(400, 2048, 2e-3)
    a_i, e_i = s.run(INDEP_FP32, 2048, 32, 0.5)
    r = s.results()
    print("independence", a_i, e_i, "learned acceptance", r["acc_learned"], s.timings(), "z", _z(a_i, e_i, exact), flush=True)
    assert np.all(np.abs(_z(a_i, e_i, exact)) < 3)
    assert r["acc_learned"] > 0.005
    assert e_i[0] < 0.7 * e_p[0]


@pytest.mark.parametrize("d,n", [(32, 300), (16, 200)])
def test_independence_move_matches_oracle(d, n):
    """The independence learned move (k_lp_indep) against its CPU restatement (oracle/mci_lp_oracle.c),
    fed the same positions, proposal and Philox words: the same component, normals (the library's
    own log, square root and sine/cosine), proposals and proposal densities bit for bit, hence the
    same decisions wherever the MH margin exceeds the FP32 error of the device's target density
    (2e-3); on every walker without such a near tie the positions and accepted counts are identical."""
    import ctypes as C
    from mcintegratorplusplus_amd.learned import LearnedProposalSampler, mixture_target
    from mcitest import load
    print("indep move vs oracle", d, n, flush=True)
    w, mu, cov, _ = mixture_target(d=d, ncomp=8 if d == 32 else 4, seed=5)
    s = LearnedProposalSampler(w, mu, cov, 4096)
    s.setSeed(11)
    s.setNetwork(128, 8)
    s.setFixedInput(True)
    s.setTrainingPairs(0.0)
    s.collect(4, 256)
    s.initFromSamples(30)
    rng = np.random.default_rng(d)
    k = rng.choice(len(w), n, p=w)
    x0 = (mu[k] + np.einsum("nij,nj->ni", np.linalg.cholesky(cov)[k], rng.normal(size=(n, d)))).astype(np.float32)
    nl, t0 = 32, 1000
    xd, accd, mix = s.indepMove(x0, nl, t0)
    L = load("orc")
    _FP, _DP = C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.orc_lp_indep_move.argtypes = [C.c_int, C.c_int, _FP, C.c_int, _DP, _DP, _DP, C.c_uint64, C.c_int64, _FP, C.c_int,
                                    C.c_uint64, C.POINTER(C.c_int32), C.POINTER(C.c_uint8), _DP]
    prec = np.ascontiguousarray(np.linalg.inv(cov))
    c = np.log(w) - 0.5 * np.array([np.linalg.slogdet(2 * np.pi * cv)[1] for cv in cov])
    xo = np.ascontiguousarray(x0.T)
    acco = np.zeros(n, np.int32)
    dec = np.zeros(n * nl, np.uint8)
    mar = np.zeros(n * nl)
    assert L.orc_lp_indep_move(d, 8, mix.ctypes.data_as(_FP), len(w), np.ascontiguousarray(mu).ctypes.data_as(_DP), prec.ctypes.data_as(_DP),
                               np.ascontiguousarray(c).ctypes.data_as(_DP), 11, n, xo.ctypes.data_as(_FP), nl, t0,
                               acco.ctypes.data_as(C.POINTER(C.c_int32)), dec.ctypes.data_as(C.POINTER(C.c_uint8)),
                               mar.ctypes.data_as(_DP)) == 0
    near = np.any(np.abs(mar.reshape(n, nl)) < 2e-3, axis=1)
    ok = ~near
    print(f"accepted {acco.sum()} of {n * nl} moves; walkers with a near tie: {near.sum()}", flush=True)
    assert acco.sum() > 0.02 * n * nl
    assert near.mean() < 0.05
    assert np.array_equal(accd[ok], acco[ok])
    assert np.array_equal(xd.T[:, ok], xo[:, ok])


def test_trained_mixture_as_the_trial_move_of_a_context():
    """The learned proposal trained on an isotropic three-mode target in d = 8, its mixture handed to
    an MCI context over the same target (the catalogue's isotropic mixture density) as the learned
    half of a mixed trial move: from walkers that all start in one mode, P(x_0 > 0) and E|x|^2 land
    within 3 standard errors of the walker means; without the learned half the walkers stay in that
    mode and miss P(x_0 > 0)."""
    from scipy.stats import norm

    from mcintegratorplusplus_amd import constants as K
    from mcintegratorplusplus_amd.learned import LearnedProposalSampler
    from mcintegratorplusplus_amd.mci import MCI
    d, W = 8, 16384
    w = np.array([0.2, 0.3, 0.5])
    s = np.array([0.6, 0.9, 1.2])
    mu = np.random.default_rng(3).normal(0.0, 4.0, (3, d))
    mu[:, 0] = [-6.0, 5.0, 7.0]
    exact = np.array([np.sum(w * norm.cdf(mu[:, 0] / s)), np.sum(w * (np.sum(mu * mu, axis=1) + d * s * s))])
    lp = LearnedProposalSampler(w, mu, np.array([sk * sk * np.eye(d) for sk in s]), W)
    lp.setSeed(5)
    lp.setNetwork(64, 4)
    lp.setFixedInput(True)
    lp.collect(8, 256)
    lp.initFromSamples(60)
    lp.train(400, 4096, 1e-3)
    lw, lm, ls = lp.mixture()
    print("learned mixture weights", np.round(lw, 3), flush=True)
    assert lm.shape == (4, d) and ls.shape == (4, d) and abs(lw.sum() - 1) < 1e-6

    def run(learned):
        c = MCI(d, W)
        c.setSeed(8)
        c.setRandomMode(K.RNG_PHILOX)
        c.addSamplingFunction(K.PDF_GMM_ISO, [3, *w, *s, *mu.ravel()])
        c.setTrialMove(K.SRRD_UNIFORM, 1)
        if learned:
            c.setLearnedMove(lw, lm, ls, 0.25)
        c.addObservable(K.OBS_INDICATOR, [0], 64, 1, False, estimator=K.ESTIM_UNCORRELATED)
        c.addObservable(K.OBS_X2_SUM, [], 64, 1, False, estimator=K.ESTIM_UNCORRELATED)
        c.setX(np.tile(mu[0], (W, 1)))
        c.integrate(64 * 64)
        wa, _ = c.walkerResults()
        # (walkers that never leave their mode have a spread of 0: a tiny floor keeps z finite)
        return (wa.mean(axis=0) - exact) / np.maximum(wa.std(axis=0, ddof=1) / np.sqrt(W), 1e-300)

    z = run(True)
    z0 = run(False)
    print("z with the learned half", z, "without", z0, flush=True)
    assert np.all(np.abs(z) < 3), z
    assert abs(z0[0]) > 20, z0


@pytest.mark.parametrize("n,d,K,iters", [(16384, 32, 8, 60), (5000, 8, 3, 40), (999, 5, 1, 1), (300, 3, 3, 0)])
def test_device_em_fit_matches_the_host_fit(n, d, K, iters):
    """The mixture fit the network starts from, on the device (seeding distances and potentials, EM
    in FP64 kernels): the same seeds as the host fit (the squared distances are bit-identical, so
    the candidates are the same draws) and, after the EM iterations, the same mixture to rounding
    (sums over samples in another order); deterministic from run to run."""
    from mcintegratorplusplus_amd.learned import gmm_fit_diag, mixture_target
    print("device EM", n, d, K, iters, flush=True)
    if d == 32:
        w, mu, cov, _ = mixture_target(d=32, ncomp=8, seed=5)
        rng = np.random.default_rng(1)
        k = rng.choice(len(w), n, p=w)
        L = np.linalg.cholesky(cov)
        x = mu[k] + np.einsum("nij,nj->ni", L[k], rng.normal(size=(n, d)))
    else:
        x = np.random.default_rng(d).normal(size=(n, d)) * np.arange(1, d + 1) + np.repeat(np.arange(3), -(-n // 3))[:n, None] * 6.0
    host = gmm_fit_diag(x, K, iters, seed=3)
    dev = gmm_fit_diag(x, K, iters, seed=3, device=True)
    dev2 = gmm_fit_diag(x, K, iters, seed=3, device=True)
    for a, b in zip(dev[:3], dev2[:3]):
        assert np.array_equal(a, b)
    if iters == 0:  # the seeds themselves: samples, equal bit for bit
        assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
    for a, b, tol in zip(host[:3], dev[:3], (1e-9, 1e-8, 1e-9)):
        assert np.allclose(b, a, rtol=tol, atol=tol * np.max(np.abs(a))), np.max(np.abs(b - a))
    assert dev[3] == pytest.approx(host[3], rel=1e-11)
