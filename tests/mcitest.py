"""Test helpers: one ctypes wrapper that drives either the HIP library (prefix
``mci``) or the CPU oracle (prefix ``orc``) through the identical C-ABI, and a
runner for the reference harness binary (oracle/_ref/mci_ref).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use the
oracle side of this module.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "liborc.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "mci_ref")
REF_BIN_FAST = os.path.join(ROOT, "oracle", "_ref", "mci_ref_fast")
GOLDEN = os.path.join(ROOT, "tests", "golden")

# constants of include/mci_amd.h
OK, E_INVALID, E_DOMAIN, E_RUNTIME, E_DEVICE = 0, 1, 2, 3, 4
ESTIM_NOOP, ESTIM_UNCORR, ESTIM_CORR, ESTIM_FC, ESTIM_MJ = 0, 1, 2, 3, 4
MOVE_ALL, MOVE_VEC = 0, 1
SRRD_UNIFORM = 0
SRRD_GAUSSIAN = 1
PDF_GAUSS_ISO, PDF_GAUSS_DIAG, PDF_EXP_DIAG, PDF_GMM_ISO = 1, 2, 3, 4
OBS_X2, OBS_X2_SUM, OBS_X, OBS_X_SUM, OBS_CONST, OBS_EXP_X2, OBS_INDICATOR, OBS_GAUSS_X2 = 1, 2, 3, 4, 5, 6, 7, 8
OBS_PDF_LOG, OBS_PDF_PROTO, OBS_DEP_SQUARE = 9, 10, 11
RNG_MT, RNG_PHILOX = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
COUPLING_INDEPENDENT, COUPLING_SHARED = 0, 1
KERNELS_AUTO, KERNELS_GENERIC = 0, 1

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_U64 = C.c_uint64
_D = C.c_double
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_I64P = C.POINTER(C.c_int64)
_U64P = C.POINTER(C.c_uint64)
_U8P = C.POINTER(C.c_uint8)

# name -> argtypes after the context pointer (None = no context argument)
SIGS = {
    "set_seed": [_U64],
    "set_walker_seeds": [_U64P, _I64],
    "set_rng": [_I],
    "set_precision": [_I],
    "set_coupling": [_I],
    "set_target_acceptance_rate": [_D],
    "set_nfind_mrt2_iterations": [_I],
    "set_ndecorrelation_steps": [_I64],
    "set_x": [_DP, _I64],
    "get_x": [_DP, _I64],
    "center_x": [],
    "move_x": [],
    "new_random_x": [],
    "set_mrt2step": [_I, _D],
    "set_mrt2steps": [_DP, _I],
    "get_mrt2step": [_I, _DP],
    "get_walker_mrt2steps": [_DP, _I64],
    "get_n_step_sizes": [_IP],
    "reset_domain": [],
    "set_irange": [_DP, _DP],
    "domain_is_finite": [_IP],
    "set_trial_move": [_I, _I, _I, _IP],
    "set_trial_move_type": [_I],
    "set_trial_move_multistep": [_I, _I, _I, _I, _IP],
    "add_multistep_sampling_function": [_I, _I, _DP, _I],
    "clear_multistep_sampling_functions": [],
    "set_learned_move": [_I, _DP, _DP, _DP, _D],
    "clear_learned_move": [],
    "get_change_rate": [_DP],
    "add_sampling_function": [_I, _I, _DP, _I],
    "pop_sampling_function": [],
    "clear_sampling_functions": [],
    "get_npdf": [_IP],
    "add_observable": [_I, _I, _DP, _I, _I, _I, _I, _I],
    "pop_observable": [],
    "clear_observables": [],
    "get_nobs": [_IP],
    "get_nobsdim": [_IP],
    "integrate": [_I64, _DP, _DP, _I, _I],
    "get_walker_results": [_DP, _DP, _I64],
    "get_acceptance": [_I64P, _I64P, _DP],
    "get_walker_acceptance": [_I64P, _I64P, _I64],
}
ORACLE_ONLY = {
    "get_warnings": [_I64P],
    "trace_enable": [_DP, _U8P, _I64],
    "get_accu_data": [_I, _I64, _DP, _I64, _I64P],
}
MCI_ONLY = {
    "get_state_size": [_I64P],
    "get_state": [C.c_void_p, _I64],
    "set_state": [C.c_void_p, _I64],
    "set_kernel_path": [_I],
    "get_timings": [_DP, _I],
    "get_observable_data": [_I, _DP, _I64, _I64P],
    "set_callback": [C.c_void_p, C.c_void_p, _I64],
    "clear_callback": [],
    "store_observables_on_file": [C.c_char_p, _I],
    "clear_observable_file": [],
    "store_walker_positions_on_file": [C.c_char_p, _I],
    "clear_walker_file": [],
    "add_sampling_function_device": [C.c_char_p, C.c_char_p, _I, _I, _DP, _I],
    "add_observable_device": [C.c_char_p, C.c_char_p, _I, _I, _DP, _I, _I, _I, _I, _I],
    "build_device_functors": [],
    "set_trial_move_device": [C.c_char_p, C.c_char_p, _I, _I, _IP, _D, _DP, _DP, _I],
}


class MciError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[status {code}] {msg}")
        self.code = code


_LIBS = {}


def load(prefix):
    """Load the C-ABI library for 'mci' (HIP product) or 'orc' (oracle)."""
    if prefix in _LIBS:
        return _LIBS[prefix]
    if prefix == "orc":
        path = ORACLE_LIB
    else:
        from mcintegratorplusplus_amd import _lib
        path = _lib.library_path()
    lib = C.CDLL(path)
    extra = ORACLE_ONLY if prefix == "orc" else MCI_ONLY
    for name, args in list(SIGS.items()) + list(extra.items()):
        f = getattr(lib, f"{prefix}_{name}")
        f.argtypes = [_P] + args
        f.restype = _I
    getattr(lib, f"{prefix}_create").argtypes = [_I, _I64, C.POINTER(_P)]
    getattr(lib, f"{prefix}_create").restype = _I
    getattr(lib, f"{prefix}_destroy").argtypes = [_P]
    getattr(lib, f"{prefix}_last_error").restype = C.c_char_p
    getattr(lib, f"{prefix}_estimate").argtypes = [_I, _I64, _I, _DP, _DP, _DP]
    getattr(lib, f"{prefix}_estimate").restype = _I
    if prefix == "orc":
        lib.orc_mt64_nth.argtypes = [_U64, _I64]
        lib.orc_mt64_nth.restype = _U64
        lib.orc_mt64_words.argtypes = [_U64, _I64, _U64P]
        lib.orc_canonical53.argtypes = [_U64]
        lib.orc_canonical53.restype = _D
        lib.orc_lemire64.argtypes = [_U64P, _I64, _U64, _I64P]
        lib.orc_lemire64.restype = _U64
        lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.orc_det_sum.argtypes = [_DP, _I64]
        lib.orc_det_sum.restype = _D
        lib.orc_block_estimate.argtypes = [_I64, _I, _DP, _I64, _DP, _DP]
        lib.orc_block_estimate.restype = _I
    _LIBS[prefix] = lib
    return lib


def dptr(a):
    return a.ctypes.data_as(_DP)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Chain:
    """A context of either backend, with the reference's method vocabulary."""

    def __init__(self, prefix, ndim, nwalkers=1):
        self.prefix = prefix
        self.lib = load(prefix)
        self.ndim = ndim
        self.W = nwalkers
        h = _P()
        self._check(getattr(self.lib, f"{prefix}_create")(ndim, nwalkers, C.byref(h)))
        self.ctx = h

    def __del__(self):
        try:
            getattr(self.lib, f"{self.prefix}_destroy")(self.ctx)
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            msg = getattr(self.lib, f"{self.prefix}_last_error")().decode()
            raise MciError(st, msg)

    def call(self, name, *args):
        self._check(getattr(self.lib, f"{self.prefix}_{name}")(self.ctx, *args))

    # settings
    def setSeed(self, s): self.call("set_seed", s)
    def setWalkerSeeds(self, seeds):
        a = np.ascontiguousarray(seeds, dtype=np.uint64)
        self.call("set_walker_seeds", a.ctypes.data_as(_U64P), len(a))
    def setRng(self, k): self.call("set_rng", k)
    def setPrecision(self, p): self.call("set_precision", p)
    def setCoupling(self, m): self.call("set_coupling", m)
    def setKernelPath(self, p): self.call("set_kernel_path", p)
    def setTargetAcceptanceRate(self, r): self.call("set_target_acceptance_rate", r)
    def setNfindMRT2Iterations(self, n): self.call("set_nfind_mrt2_iterations", n)
    def setNdecorrelationSteps(self, n): self.call("set_ndecorrelation_steps", n)

    def setX(self, x):
        a = f64(x).ravel()
        self.call("set_x", dptr(a), a.size)

    def getX(self):
        a = np.zeros(self.W * self.ndim)
        self.call("get_x", dptr(a), a.size)
        return a.reshape(self.W, self.ndim)

    def centerX(self): self.call("center_x")
    def moveX(self): self.call("move_x")
    def newRandomX(self): self.call("new_random_x")
    def setMRT2Step(self, s, idx=-1): self.call("set_mrt2step", idx, s)

    def getMRT2Step(self, i=0):
        v = _D()
        self.call("get_mrt2step", i, C.byref(v))
        return v.value

    def getWalkerSteps(self):
        n = _I()
        self.call("get_n_step_sizes", C.byref(n))
        a = np.zeros(self.W * n.value)
        self.call("get_walker_mrt2steps", dptr(a), a.size)
        return a.reshape(self.W, n.value)

    def setIRange(self, lb, ub):
        lo = f64(np.broadcast_to(lb, (self.ndim,)))
        hi = f64(np.broadcast_to(ub, (self.ndim,)))
        self.call("set_irange", dptr(lo), dptr(hi))

    def resetDomain(self): self.call("reset_domain")

    def setTrialMove(self, srrd=SRRD_UNIFORM, veclen=0, ntypes=1, typeEnds=None):
        te = None
        if typeEnds is not None:
            arr = (C.c_int * len(typeEnds))(*typeEnds)
            te = C.cast(arr, _IP)
        self.call("set_trial_move", srrd, veclen, ntypes, te)

    def setTrialMoveType(self, t): self.call("set_trial_move_type", t)

    def setTrialMoveMultistep(self, nsteps, srrd=SRRD_UNIFORM, veclen=1, ntypes=1, typeEnds=None):
        te = None
        if typeEnds is not None:
            arr = (C.c_int * len(typeEnds))(*typeEnds)
            te = C.cast(arr, _IP)
        self.call("set_trial_move_multistep", nsteps, srrd, veclen, ntypes, te)

    def setLearnedMove(self, weights, means, sigmas, pmix):
        w, m, s = f64(weights), f64(means), f64(sigmas)
        self.call("set_learned_move", len(w), dptr(w), dptr(m), dptr(s), pmix)

    def clearLearnedMove(self): self.call("clear_learned_move")

    def setCallback(self, fn, capacity=1 << 20):
        """fn(x [ndim] array, acc, rej, rate) for every callback call of an integrate"""
        class View(C.Structure):
            _fields_ = [("call", C.c_int64), ("ndim", C.c_int), ("x", C.POINTER(C.c_double)),
                        ("acc", C.c_int64), ("rej", C.c_int64), ("rate", C.c_double)]
        proto = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(View))

        def tramp(user, v):
            v = v.contents
            fn(np.ctypeslib.as_array(v.x, (v.ndim,)).copy(), v.acc, v.rej, v.rate)
        self._cb = proto(tramp)  # kept alive with the chain
        self.call("set_callback", C.cast(self._cb, C.c_void_p), None, capacity)

    def storeObservablesOnFile(self, path, freq): self.call("store_observables_on_file", str(path).encode(), freq)
    def storeWalkerPositionsOnFile(self, path, freq): self.call("store_walker_positions_on_file", str(path).encode(), freq)

    def addMultistepSamplingFunction(self, kind, params=()):
        p = f64(params)
        self.call("add_multistep_sampling_function", kind, self.ndim, dptr(p) if p.size else None, p.size)

    def getChangeRate(self):
        v = _D()
        self.call("get_change_rate", C.byref(v))
        return v.value

    def addSamplingFunction(self, kind, params=()):
        p = f64(params)
        self.call("add_sampling_function", kind, self.ndim, dptr(p) if p.size else None, p.size)

    def addSamplingFunctionDevice(self, name, source, nproto, params=()):
        p = f64(params)
        self.call("add_sampling_function_device", name.encode(), source.encode(), self.ndim, nproto, dptr(p) if p.size else None, p.size)

    def addObservableDevice(self, name, source, nobs, params=(), blocksize=1, nskip=1, flag_equil=None, estim=None):
        if flag_equil is None:
            flag_equil = blocksize > 0
        if estim is None:
            estim = ESTIM_CORR if blocksize == 1 else (ESTIM_UNCORR if blocksize > 0 else ESTIM_NOOP)
        p = f64(params)
        self.call("add_observable_device", name.encode(), source.encode(), self.ndim, nobs, dptr(p) if p.size else None, p.size,
                  blocksize, nskip, int(bool(flag_equil)), estim)

    def setTrialMoveDevice(self, name, source, nproto, steps, typeEnds=None, change_rate=1.0, params=()):
        st = f64(steps)
        te = None
        if typeEnds is not None:
            arr = (C.c_int * len(typeEnds))(*typeEnds)
            te = C.cast(arr, _IP)
        p = f64(params)
        self.call("set_trial_move_device", name.encode(), source.encode(), nproto, st.size, te, change_rate, dptr(st),
                  dptr(p) if p.size else None, p.size)

    def popSamplingFunction(self): self.call("pop_sampling_function")
    def clearSamplingFunctions(self): self.call("clear_sampling_functions")

    def addObservable(self, kind, params=(), blocksize=1, nskip=1, flag_equil=None, estim=None):
        if flag_equil is None:
            flag_equil = blocksize > 0
        if estim is None:  # the reference's default overload (MCIntegrator.hpp:161-164)
            estim = ESTIM_CORR if blocksize == 1 else (ESTIM_UNCORR if blocksize > 0 else ESTIM_NOOP)
        p = f64(params)
        self.call("add_observable", kind, self.ndim, dptr(p) if p.size else None, p.size,
                  blocksize, nskip, int(bool(flag_equil)), estim)

    def popObservable(self): self.call("pop_observable")
    def clearObservables(self): self.call("clear_observables")

    def getNObsDim(self):
        v = _I()
        self.call("get_nobsdim", C.byref(v))
        return v.value

    def integrate(self, nmc, do_find=True, do_decorr=True):
        nd = max(self.getNObsDim(), 1)
        a, e = np.zeros(nd), np.zeros(nd)
        self.call("integrate", nmc, dptr(a), dptr(e), int(do_find), int(do_decorr))
        return a, e

    def walkerResults(self):
        nd = self.getNObsDim()
        a, e = np.zeros(self.W * nd), np.zeros(self.W * nd)
        self.call("get_walker_results", dptr(a), dptr(e), a.size)
        return a.reshape(self.W, nd), e.reshape(self.W, nd)

    def acceptance(self):
        acc, rej, rate = _I64(), _I64(), _D()
        self.call("get_acceptance", C.byref(acc), C.byref(rej), C.byref(rate))
        return acc.value, rej.value, rate.value

    def observableData(self, i, nobs):
        """finalised accumulator data of observable i (nobs components), shape [W, nstore, nobs]
        (HIP library only)"""
        ns = _I64()
        self.call("get_observable_data", i, None, 0, C.byref(ns))
        buf = np.zeros(self.W * ns.value * nobs)
        self.call("get_observable_data", i, dptr(buf), buf.size, C.byref(ns))
        return buf.reshape(self.W, ns.value, nobs)

    def getState(self):
        n = _I64()
        self.call("get_state_size", C.byref(n))
        buf = (C.c_char * n.value)()
        self.call("get_state", buf, n.value)
        return bytes(buf)

    def setState(self, state):
        buf = C.create_string_buffer(bytes(state), len(state))
        self.call("set_state", buf, len(state))

    def walkerAcceptance(self):
        a = np.zeros(self.W, dtype=np.int64)
        r = np.zeros(self.W, dtype=np.int64)
        self.call("get_walker_acceptance", a.ctypes.data_as(_I64P), r.ctypes.data_as(_I64P), self.W)
        return a, r


def estimate(prefix, estim, x):
    """Run a standalone estimator on x of shape [n, ndim] (or [n])."""
    lib = load(prefix)
    x = f64(x)
    if x.ndim == 1:
        x = x[:, None]
    n, nd = x.shape
    a, e = np.zeros(nd), np.zeros(nd)
    st = getattr(lib, f"{prefix}_estimate")(estim, n, nd, dptr(x), dptr(a), dptr(e))
    if st != OK:
        raise MciError(st, getattr(lib, f"{prefix}_last_error")().decode())
    return a, e


# ---- reference harness ----

def hexf(s):
    return float.fromhex(s)


def hexv(lst):
    return np.array([float.fromhex(s) for s in lst])


def ref_run(lines, binary=REF_BIN, timeout=600):
    """Run the reference harness on a script; returns the parsed JSON records."""
    script = "\n".join(lines) + "\n"
    out = subprocess.run([binary], input=script, capture_output=True, text=True, timeout=timeout, check=True)
    return [json.loads(l) for l in out.stdout.splitlines() if l.startswith("{")]


def fmt(v):
    return float(v).hex()


def script_setup(cfg):
    """Reference-harness script lines for a configuration dict (see tests/golden/make_golden.py)."""
    L = [f"create {cfg['ndim']}"]
    if "seed" in cfg:
        L.append(f"seed {cfg['seed']}")
    if "target" in cfg:
        L.append(f"target {fmt(cfg['target'])}")
    if "nfind" in cfg:
        L.append(f"nfind {cfg['nfind']}")
    if "ndecorr" in cfg:
        L.append(f"ndecorr {cfg['ndecorr']}")
    if "move" in cfg:
        srrd, veclen, ntypes, te = cfg["move"]
        L.append(f"move {srrd} {veclen} {ntypes} " + " ".join(str(t) for t in (te or [])))
    if "umove" in cfg:
        kind, params = cfg["umove"]
        L.append(f"umove {kind} {len(params)} " + " ".join(fmt(p) for p in params))
    if "multistep" in cfg:
        nsteps, srrd, veclen, ntypes, te, mpdfs = cfg["multistep"]
        L.append(f"msmove {nsteps} {srrd} {veclen} {ntypes} " + " ".join(str(t) for t in (te or [])))
        for kind, params in mpdfs:
            L.append(f"mspdf {kind} {len(params)} " + " ".join(fmt(p) for p in params))
        L.append("msapply")
    if "irange" in cfg:
        lb, ub = cfg["irange"]
        L.append("irange " + " ".join(fmt(v) for v in lb) + " " + " ".join(fmt(v) for v in ub))
    for kind, params in cfg.get("pdfs", []):
        L.append(f"pdf {kind} {len(params)} " + " ".join(fmt(p) for p in params))
    for o in cfg.get("obs", []):
        kind, params, bs, nskip, eq, est = o
        L.append(f"obs {kind} {len(params)} " + " ".join(fmt(p) for p in params) + f" {bs} {nskip} {eq} {est}")
    if "x0" in cfg:
        L.append("setx " + " ".join(fmt(v) for v in cfg["x0"]))
    if "step" in cfg:
        L.append(f"step {fmt(cfg['step'])}")
    return L


def chain_setup(prefix, cfg, nwalkers=1):
    """The same configuration on a C-ABI backend."""
    c = Chain(prefix, cfg["ndim"], nwalkers)
    if "seed" in cfg:
        c.setSeed(cfg["seed"])
    if "target" in cfg:
        c.setTargetAcceptanceRate(cfg["target"])
    if "nfind" in cfg:
        c.setNfindMRT2Iterations(cfg["nfind"])
    if "ndecorr" in cfg:
        c.setNdecorrelationSteps(cfg["ndecorr"])
    if "move" in cfg:
        srrd, veclen, ntypes, te = cfg["move"]
        c.setTrialMove(srrd, veclen, ntypes, te)
    if "umove" in cfg:  # a trial move outside the SRRD family: its device functor source
        import device_functors as DF
        kind, params = cfg["umove"]
        name, src, nproto, te, rate, steps = DF.MOVES[kind]
        d = cfg["ndim"]
        c.setTrialMoveDevice(name, src, nproto, steps(params), te(d), rate(d))
    if "multistep" in cfg:
        nsteps, srrd, veclen, ntypes, te, mpdfs = cfg["multistep"]
        c.setTrialMoveMultistep(nsteps, srrd, veclen, ntypes, te)
        for kind, params in mpdfs:
            c.addMultistepSamplingFunction(kind, params)
    if "learned" in cfg:
        c.setLearnedMove(*cfg["learned"])
    if "irange" in cfg:
        c.setIRange(*cfg["irange"])
    for kind, params in cfg.get("pdfs", []):
        if kind >= 100:  # a plugin outside the catalogue: its device functor source
            import device_functors as DF
            name, src, nproto = DF.PDFS[kind]
            c.addSamplingFunctionDevice(name, src, nproto(cfg["ndim"]), params)
        else:
            c.addSamplingFunction(kind, params)
    for o in cfg.get("obs", []):
        kind, params, bs, nskip, eq, est = o
        if kind >= 100:
            import device_functors as DF
            name, src, nobs = DF.OBS[kind]
            c.addObservableDevice(name, src, nobs(cfg["ndim"]), params, bs, nskip, eq, est)
        else:
            c.addObservable(kind, params, bs, nskip, eq, est)
    if "x0" in cfg:
        c.setX(cfg["x0"])
    if "step" in cfg:
        c.setMRT2Step(cfg["step"])
    return c
