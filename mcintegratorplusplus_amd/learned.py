"""The learned proposal (mci_mlp_* and mci_lp_* of include/mci_amd.h) from Python.

Network: an MLP x -> K-component diagonal Gaussian mixture q(. | x) with BF16 MFMA forward and
backward on the GPU.  LearnedProposalSampler: a mixture-of-Gaussians target sampled by the plain
random-walk chain (FP64 or FP32) or by the chain that mixes random-walk moves with MH-corrected
moves drawn from the network, which it trains on samples of the plain chain.
"""
import ctypes as C

import numpy as np
from scipy.stats import norm

from .mci import _check, _lib as _mci_lib

_P, _I, _I64, _U64, _D = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_DP, _FP, _IP = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)

PLAIN_FP64, PLAIN_FP32, MIXED_FP32, INDEP_FP32 = 0, 1, 2, 3

_SIGS = {
    "mci_mlp_create": [_I, _I, _I, _U64, C.POINTER(_P)], "mci_mlp_destroy": [_P],
    "mci_mlp_nparams": [_P, C.POINTER(C.c_int64)], "mci_mlp_get_params": [_P, _FP, _I64],
    "mci_mlp_set_params": [_P, _FP, _I64], "mci_mlp_forward": [_P, _I64, _FP, _FP],
    "mci_mlp_loss_grad": [_P, _I64, _FP, _FP, _DP, _FP],
    "mci_lp_create": [_I, _I, _DP, _DP, _DP, _I64, C.POINTER(_P)], "mci_lp_destroy": [_P],
    "mci_lp_set_seed": [_P, _U64], "mci_lp_set_network": [_P, _I, _I],
    "mci_lp_run": [_P, _I, _I64, _I, _D, _DP, _DP], "mci_lp_collect": [_P, _I, _I64],
    "mci_lp_train": [_P, _I, _I, _D, _DP, _DP], "mci_lp_get_results": [_P, _DP, _I],
    "mci_lp_get_timings": [_P, _DP, _I], "mci_lp_get_walker_means": [_P, _DP, _I64],
    "mci_lp_get_network": [_P, C.POINTER(_P)], "mci_lp_set_training_pairs": [_P, _D],
    "mci_lp_eval_network": [_P, _I64, _FP, _FP], "mci_lp_set_fixed_input": [_P, _I],
    "mci_lp_init_network_from_samples": [_P, _I],
    "mci_lp_indep_move": [_P, _I64, _FP, _I, _U64, C.POINTER(C.c_int32), _FP],
    "mci_lp_get_mixture": [_P, _I, C.POINTER(C.c_int), _DP, _DP, _DP],
    "mci_gmm_fit_diag": [_DP, _I64, _I, _I, _I, C.c_uint64, _DP, _DP, _DP, _DP],
    "mci_gmm_fit_diag_device": [_DP, _I64, _I, _I, _I, C.c_uint64, _DP, _DP, _DP, _DP],
}
_bound = None


def _lib():
    global _bound
    if _bound is None:
        lib = _mci_lib()
        for name, args in _SIGS.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _I
        _bound = lib
    return _bound


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Network:
    """MLP d -> H -> H -> K (1 + 2 d) (SiLU), BF16 weights and activations, FP32 accumulation."""

    def __init__(self, d=32, hidden=128, ncomp=8, seed=1, _handle=None):
        self.d, self.hidden, self.ncomp = d, hidden, ncomp
        self.nout = ncomp * (1 + 2 * d)
        if _handle is not None:
            self._h, self._own = _handle, False
        else:
            h = _P()
            _check(_lib().mci_mlp_create(d, hidden, ncomp, seed, C.byref(h)))
            self._h, self._own = h, True

    def __del__(self):
        h = getattr(self, "_h", None)
        if getattr(self, "_own", False) and h is not None and h.value:
            _lib().mci_mlp_destroy(h)
            self._h = None

    def nparams(self):
        n = C.c_int64()
        _check(_lib().mci_mlp_nparams(self._h, C.byref(n)))
        return n.value

    def params(self):
        p = np.zeros(self.nparams(), np.float32)
        _check(_lib().mci_mlp_get_params(self._h, p.ctypes.data_as(_FP), p.size))
        return p

    def setParams(self, p):
        p = _f(p)
        _check(_lib().mci_mlp_set_params(self._h, p.ctypes.data_as(_FP), p.size))

    def forward(self, x):
        x = _f(x)
        out = np.zeros((x.shape[0], self.nout), np.float32)
        _check(_lib().mci_mlp_forward(self._h, x.shape[0], x.ctypes.data_as(_FP), out.ctypes.data_as(_FP)))
        return out

    def lossGrad(self, x, y):
        x, y = _f(x), _f(y)
        loss = C.c_double()
        g = np.zeros(self.nparams(), np.float32)
        _check(_lib().mci_mlp_loss_grad(self._h, x.shape[0], x.ctypes.data_as(_FP), y.ctypes.data_as(_FP),
                                        C.byref(loss), g.ctypes.data_as(_FP)))
        return loss.value, g


class LearnedProposalSampler:
    def __init__(self, weights, means, covs, nwalkers):
        self.weights, self.means, self.covs = _d(weights), _d(means), _d(covs)
        self.ncomp, self.d = self.means.shape
        self.nwalkers = nwalkers
        self.hidden, self.net_ncomp = 128, 8
        h = _P()
        _check(_lib().mci_lp_create(self.d, self.ncomp, self.weights.ctypes.data_as(_DP), self.means.ctypes.data_as(_DP),
                                    self.covs.ctypes.data_as(_DP), nwalkers, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib().mci_lp_destroy(h)
            self._h = None

    def setSeed(self, seed):
        _check(_lib().mci_lp_set_seed(self._h, seed))

    def setNetwork(self, hidden=128, ncomp=8):
        _check(_lib().mci_lp_set_network(self._h, hidden, ncomp))
        self.hidden, self.net_ncomp = hidden, ncomp

    def run(self, mode, nsteps, nblocks=64, pmix=0.5):
        a, e = np.zeros(2), np.zeros(2)
        _check(_lib().mci_lp_run(self._h, mode, nsteps, nblocks, pmix, a.ctypes.data_as(_DP), e.ctypes.data_as(_DP)))
        return a, e

    def collect(self, nsnap, lag):
        _check(_lib().mci_lp_collect(self._h, nsnap, lag))

    def setTrainingPairs(self, same_fraction):
        _check(_lib().mci_lp_set_training_pairs(self._h, same_fraction))

    def initFromSamples(self, em_iters=40):
        """start the network's last layer at an EM fit of its mixture to the last snapshot"""
        _check(_lib().mci_lp_init_network_from_samples(self._h, em_iters))

    def setFixedInput(self, on=True):
        """train with one fixed input: the network's mixture there is an independence proposal"""
        _check(_lib().mci_lp_set_fixed_input(self._h, 1 if on else 0))

    def train(self, iters, batch=4096, lr=1e-3):
        l0, l1 = C.c_double(), C.c_double()
        _check(_lib().mci_lp_train(self._h, iters, batch, lr, C.byref(l0), C.byref(l1)))
        return l0.value, l1.value

    def results(self):
        r = np.zeros(8)
        _check(_lib().mci_lp_get_results(self._h, r.ctypes.data_as(_DP), 8))
        return dict(avg=r[0:2], err=r[2:4], err_mpimci=r[4:6], acc_rw=r[6], acc_learned=r[7])

    def timings(self):
        ms = np.zeros(4)
        _check(_lib().mci_lp_get_timings(self._h, ms.ctypes.data_as(_DP), 4))
        return dict(calibration=ms[0], sampling=ms[1], estimation=ms[2], training=ms[3])

    def walkerMeans(self):
        m = np.zeros(2 * self.nwalkers)
        _check(_lib().mci_lp_get_walker_means(self._h, m.ctypes.data_as(_DP), m.size))
        return m.reshape(2, self.nwalkers)

    def evalNetwork(self, x):
        """the mixture as the learned move computes it at positions x [n][d]: [n][K (1 + 2 d)]
        (per component d means, d log-scales; then K logits)"""
        x = _f(x)
        n = x.shape[0]
        net = self.network()
        out = np.zeros((net.nout, n), np.float32)
        xt = np.ascontiguousarray(x.T)
        _check(_lib().mci_lp_eval_network(self._h, n, xt.ctypes.data_as(_FP), out.ctypes.data_as(_FP)))
        return out.T.copy()

    def indepMove(self, x, nl, t0=0):
        """nl independence moves of walkers at positions x [n][d] (FP32), steps t0 .. t0 + nl - 1 of
        walker streams 0 .. n - 1: (new positions [n][d], accepted moves [n], the proposal as the move
        uses it: means, 1/sigma, sigma [8][d], log normalisers [8], cumulative weights [8])"""
        x = _f(x)
        n, d = x.shape
        xt = np.ascontiguousarray(x.T)
        acc = np.zeros(n, np.int32)
        mix = np.zeros(3 * 8 * d + 16, np.float32)
        _check(_lib().mci_lp_indep_move(self._h, n, xt.ctypes.data_as(_FP), nl, t0, acc.ctypes.data_as(C.POINTER(C.c_int32)),
                                        mix.ctypes.data_as(_FP)))
        return xt.T.copy(), acc, mix

    def mixture(self):
        """the trained independence proposal (weights [K], means [K][d], sigmas [K][d]), the form
        MCI.setLearnedMove takes"""
        kc = self.net_ncomp
        w, m, s = np.zeros(kc), np.zeros((kc, self.d)), np.zeros((kc, self.d))
        k = C.c_int()
        _check(_lib().mci_lp_get_mixture(self._h, kc, C.byref(k), w.ctypes.data_as(_DP), m.ctypes.data_as(_DP),
                                         s.ctypes.data_as(_DP)))
        return w[:k.value].copy(), m[:k.value].copy(), s[:k.value].copy()

    def network(self):
        h = _P()
        _check(_lib().mci_lp_get_network(self._h, C.byref(h)))
        return Network(self.d, self.hidden, self.net_ncomp, _handle=h)


def gmm_fit_diag(x, ncomp, em_iters=60, seed=1, device=False):
    """diagonal Gaussian mixture fitted to x [n][d] by EM from greedy k-means++ seeds: (means [K][d],
    variances [K][d], weights [K], mean log-likelihood).  device=False: host code; device=True: the
    same fit with its sample sums on the GPU (the one the learned proposal's network starts from)"""
    x = _d(x)
    n, d = x.shape
    mu, var, pi, ll = np.zeros((ncomp, d)), np.zeros((ncomp, d)), np.zeros(ncomp), C.c_double()
    fn = _lib().mci_gmm_fit_diag_device if device else _lib().mci_gmm_fit_diag
    _check(fn(x.ctypes.data_as(_DP), n, d, ncomp, em_iters, seed, mu.ctypes.data_as(_DP), var.ctypes.data_as(_DP),
              pi.ctypes.data_as(_DP), C.byref(ll)))
    return mu, var, pi, ll.value


def mixture_target(d=32, ncomp=8, seed=5, mean_sd=3.0, eig=(0.1, 1.0)):
    """Mixture of ncomp equally weighted Gaussians with means N(0, mean_sd^2 I) and covariances
    R diag(U(eig)) R^T, R Haar-random orthogonal (QR of a Gaussian matrix, signs fixed).
    Returns (weights, means, covs, exact) with exact = (E|x|^2, P(x_0 > 0))."""
    rng = np.random.default_rng(seed)
    w = np.full(ncomp, 1.0 / ncomp)
    mu = rng.normal(0.0, mean_sd, (ncomp, d))
    covs = np.zeros((ncomp, d, d))
    for k in range(ncomp):
        q, r = np.linalg.qr(rng.normal(size=(d, d)))
        q = q * np.sign(np.diag(r))
        lam = rng.uniform(eig[0], eig[1], d)
        covs[k] = (q * lam) @ q.T
        covs[k] = 0.5 * (covs[k] + covs[k].T)
    ex2 = float(np.sum(w * (np.sum(mu * mu, axis=1) + np.trace(covs, axis1=1, axis2=2))))
    pind = float(np.sum(w * norm.cdf(mu[:, 0] / np.sqrt(covs[:, 0, 0]))))
    return w, mu, covs, (ex2, pind)
