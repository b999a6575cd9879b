"""The learned-proposal configuration at full size on the GPU: d = 32 target, a mixture of 8 rotated
anisotropic Gaussians (means N(0, 3^2 I), covariances R diag(U(0.1, 1)) R^T), observables |x|^2 and
1[x_0 > 0]; 65,536 walkers x 10,240 steps; the network trained on 2^20 samples of the plain chain
(16 snapshots of every walker, 512 steps apart).

- every chain (plain FP64, conditional learned proposal, independence learned proposal) lands within
  3 error bars of the analytic values (error bar: the standard error of the 65,536 walker means);
- the independence chain reaches at least 10x the ESS/s of the plain FP64 chain with every one-off
  cost counted (collection of the training samples, the EM start, the GPU training);
- the trained network gains over the same network untrained (the method without learning)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS, W = 10240, 65536


def target():
    from mcintegratorplusplus_amd.learned import mixture_target
    w, mu, cov, (ex2, pind) = mixture_target(d=32, ncomp=8, seed=5)
    ex4 = float(np.sum(w * [(np.trace(c) + m @ m) ** 2 + 2 * np.trace(c @ c) + 4 * m @ c @ m for m, c in zip(mu, cov)]))
    var = np.array([ex4 - ex2 ** 2, pind * (1 - pind)])
    return w, mu, cov, np.array([ex2, pind]), var


def seconds(s):
    t = s.timings()
    return (t["sampling"] + t["calibration"]) / 1000


def train(s, fixed):
    """collect, EM start, 2 x 1,000 Adam steps; host clock around the synchronous calls"""
    s.setNetwork(128, 8)
    s.setFixedInput(fixed)
    s.setTrainingPairs(0.0)
    t0 = time.perf_counter()
    s.collect(16, 512)
    s.initFromSamples(60)
    for lr in (1e-3, 5e-4):
        s.train(1000, 4096, lr)
    return time.perf_counter() - t0


def test_learned_proposal_configuration_at_full_size():
    from mcintegratorplusplus_amd.learned import INDEP_FP32, MIXED_FP32, PLAIN_FP64, LearnedProposalSampler
    w, mu, cov, exact, var = target()
    s = LearnedProposalSampler(w, mu, cov, W)
    s.setSeed(1)
    print("plain FP64", flush=True)
    a_p, e_p = s.run(PLAIN_FP64, NSTEPS, 64)
    ess_p = var / e_p ** 2 / seconds(s)
    # the method without learning: the independence move with the network's mixture untrained (at
    # the fixed input, the mean of a snapshot of the plain chain)
    s.setNetwork(128, 8)
    s.setFixedInput(True)
    s.collect(2, 512)
    a_u, e_u = s.run(INDEP_FP32, NSTEPS, 64, 0.5)
    acc_u = s.results()["acc_learned"]
    print("conditional", flush=True)
    t_cond = train(s, fixed=False)
    a_c, e_c = s.run(MIXED_FP32, NSTEPS, 64, 0.25)
    ess_c = var / e_c ** 2 / seconds(s)
    print("independence", flush=True)
    t_ind = train(s, fixed=True)
    a_i, e_i = s.run(INDEP_FP32, NSTEPS, 64, 0.5)
    sec_i = seconds(s)
    acc_i = s.results()["acc_learned"]
    ess_i = var / e_i ** 2 / sec_i
    ess_i_all = var / e_i ** 2 / (sec_i + t_ind)
    for name, a, e in (("plain", a_p, e_p), ("untrained", a_u, e_u), ("conditional", a_c, e_c), ("independence", a_i, e_i)):
        z = (a - exact) / e
        print(f"{name}: z {np.round(z, 2)}, err {e}", flush=True)
        assert np.all(np.abs(z) < 3), (name, z)
    print(f"ESS/s gain vs plain FP64: conditional {ess_c / ess_p}, independence {ess_i / ess_p}, "
          f"independence with the one-off {t_ind:.3f} s counted {ess_i_all / ess_p}; "
          f"learned acceptance untrained {acc_u:.4f}, trained {acc_i:.4f}", flush=True)
    assert np.all(ess_i_all / ess_p >= 10)
    assert np.all(ess_c / ess_p >= 10)
    # what training gains: far higher acceptance of learned moves and far smaller error bars at equal steps
    assert acc_i > 20 * max(acc_u, 1e-4)
    assert np.all(e_i < 0.3 * e_u)
