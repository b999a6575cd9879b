"""Benchmark of the Metropolis path on one MI355X.

Workload (the GPU configuration of the benchmark set, FP64): 65,536 walkers x 10,240 steps,
d = 16 anisotropic Gaussian p(x) ~ exp(-sum x_i^2 / (2 sigma_i^2)) with sigma_i ~ U(0.5, 2) from
seed 1337 (tests/golden/sigmas.json), single-coordinate uniform moves with the step calibrated
to 50 % acceptance (walkers coupled like MPI ranks: shared step, MIN_STAT 100), 1,000
decorrelation steps, observables x_i^2 (16) and |x|^2 in 64 blocks of 160 steps, MJBlocker
errors per walker, combined over walkers as MPIMCI does.

One "step" of this benchmark = one complete MCI::integrate of that workload (calibration +
decorrelation + sampling + estimation + combine), inputs resident on the GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `value` is walker-steps/s of the production chain (65,536 x 10,240 per
integrate) over the whole integrate; ESS/s, the roofline of the sampling kernel and the
reference CPU path timed on this host's cores are reported beside it.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W, NMC, D, BLOCK, NDECORR = 65536, 10240, 16, 160, 1000
# algorithmic work per walker-step of the sampling kernel (FP64): 4 (move) + 3 (delta log p)
# + 2 (|x|^2 update) + 17 (block accumulation of 17 observable components) flops
FLOPS_PER_STEP = 26
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3  # TFLOP/s, MI355X FP32 vector
SEED = 1337


def cpu_share():
    """CPU processes for the reference CPU path: the cores this process may run on, capped by the
    host's CPU share for one GPU (OMP_NUM_THREADS on the GPU host, 16: worker pools there are
    sized to it whatever the machine shows).  Returns (processes, cores visible)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(visible, share)), visible


def sigmas():
    with open(os.path.join(ROOT, "tests", "golden", "sigmas.json")) as f:
        return np.array([float.fromhex(v) for v in json.load(f)["sigma16"]])


def make_mci(walkers, precision=0):
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    m = MCI(D, walkers)
    m.setSeed(SEED)
    m.setRandomMode(K.RNG_PHILOX)
    m.setPrecision(precision)
    m.setWalkerCoupling(K.COUPLING_SHARED)
    m.setTrialMove(K.SRRD_UNIFORM, 1, 1)
    m.setNdecorrelationSteps(NDECORR)
    m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
    m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
    m.addObservable(K.OBS_X2_SUM, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
    return m


def ess_per_obs(err, walkers_steps):
    """ESS_j = s_j^2 / err_j^2 with the per-step variance s^2 of x_i^2 (2 sigma_i^4) and of |x|^2
    (sum 2 sigma_i^4) under the Gaussian target.  err is the standard error of the mean of the
    independent walkers' means (calibrated: the deviations from the analytic values are N(0, 1) in
    units of it); the MPIMCI combine of the walkers' MJBlocker errors on 64 blocks of 160 steps is
    about 14 % smaller on this workload (its blocking test has little power on 64 blocks), and
    would overstate ESS by about 1.34x."""
    sig = sigmas()
    s2 = np.concatenate([2 * sig ** 4, [np.sum(2 * sig ** 4)]])
    return s2 / err ** 2


def cpu_reference_gauss(dim, sig, obs_script, budget_s):
    """The reference CPU path (oracle/_ref/mci_ref_fast: the reference library built from its
    own sources with -O3 -flto) on a diagonal Gaussian in `dim` dimensions, timed on this host:
    P processes, each integrating its share of independent walkers of the workload (one MCI per
    walker with its own calibration and NDECORR decorrelation steps, no MPI).  Returns the
    walkers, walker-steps, wall seconds and the sum of the walkers' squared errors."""
    binary = os.path.join(ROOT, "oracle", "_ref", "mci_ref_fast")
    if not os.path.exists(binary):
        return None
    a = " ".join(float(v).hex() for v in 0.5 / sig ** 2)
    script = f"create {dim}\nmove 0 1 1\nndecorr {NDECORR}\npdf 2 {dim} {a}\n{obs_script}endbench\n"
    procs, _ = cpu_share()
    # calibrate the per-walker cost with one process, then size the sample to the budget
    t0 = time.time()
    subprocess.run([binary], input=f"bench 4 {NMC} 9001\n" + script, capture_output=True, text=True, check=True)
    per_walker = max((time.time() - t0) / 4, 1e-4)
    per_proc = int(min(100000, max(4, budget_s / (procs * per_walker))))  # ~budget_s of CPU work in total
    t0 = time.time()
    ps = [subprocess.Popen([binary], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
          for _ in range(procs)]
    for i, p in enumerate(ps):
        p.stdin.write(f"bench {per_proc} {NMC} {100000 + i * per_proc}\n" + script)
        p.stdin.close()
    outs = [json.loads([l for l in p.stdout.read().splitlines() if l.startswith("{")][-1]) for p in ps]
    for p in ps:
        p.wait()
    wall = time.time() - t0
    nw = sum(o["walkers"] for o in outs)
    s1 = np.sum([[float.fromhex(v) for v in o["avg_sum"]] for o in outs], axis=0)
    s2 = np.sum([[float.fromhex(v) for v in o["avg2_sum"]] for o in outs], axis=0)
    spread = np.sqrt(np.maximum(s2 / nw - (s1 / nw) ** 2, 0.0) * nw / (nw - 1) / nw)  # standard error of the walker means
    return dict(walkers=nw, steps=sum(o["walker_steps"] for o in outs), wall=wall, procs=procs,
                err2=np.sum([[float.fromhex(v) for v in o["err2_sum"]] for o in outs], axis=0), spread_err=spread)


def cpu_baseline(budget_s=20.0):
    """The reference CPU path on the benchmark workload (d = 16, x_i^2 and |x|^2)."""
    r = cpu_reference_gauss(D, sigmas(), f"obs 1 0 {BLOCK} 1 1 4\nobs 2 0 {BLOCK} 1 1 4\n", budget_s)
    if r is None:
        return None
    nw, steps, wall, procs = r["walkers"], r["steps"], r["wall"], r["procs"]
    ess = ess_per_obs(r["spread_err"], steps)
    return dict(value=steps / wall, unit="walker-steps/s", cores=procs, kind="reference",
                ess_per_s=float(np.min(ess)) / wall,
                sample=f"{nw} walkers x {NMC} steps (+calibration, {NDECORR} decorrelation steps each) in {procs} "
                       f"processes, {wall:.1f} s wall; reference built from its sources with -O3 -flto",
                seconds=wall)


def cpu_single_core_cfg1():
    """The reference's own CPU case, timed on one core of this host (the harness model of
    benchmark/common/MCIBenchmarks.hpp:34-45): d = 3, p ~ exp(-|x|^2), x_i^2 in 16 blocks of 62,500
    with MJBlocker, default all-coordinate uniform move with automatic calibration and automatic
    decorrelation, 1 walker, 10^6 steps, seed 1337.  steps/s over the whole integrate and over the
    sampling alone (a second integrate without calibration and decorrelation); ESS = Var(x_i^2) /
    err^2 with the exact per-step variance 1/2."""
    binary = os.path.join(ROOT, "oracle", "_ref", "mci_ref_fast")
    if not os.path.exists(binary):
        return None
    nmc = 1000000
    script = (f"create 3\nseed 1337\npdf 1 0\nobs 1 0 62500 1 1 4\nintegrate {nmc} 1 1\nintegrate {nmc} 0 0\n")
    out = subprocess.run([binary], input=script, capture_output=True, text=True, check=True).stdout
    recs = [json.loads(l) for l in out.splitlines() if l.startswith("{")]
    full, samp = recs[0], recs[1]
    err = np.array([float.fromhex(v) for v in samp["err"]])
    ess = 0.5 / err ** 2
    return {"workload": "d=3 exp(-|x|^2), x_i^2 in 16 blocks of 62500, MJBlocker, all-move, 1 walker x 1e6 steps, seed 1337",
            "cores": 1, "integrate_s": full["seconds"], "steps_per_s_integrate": nmc / full["seconds"],
            "sampling_s": samp["seconds"], "steps_per_s_sampling": nmc / samp["seconds"],
            "ess_per_s_sampling": float(np.min(ess)) / samp["seconds"],
            "avg": [float.fromhex(v) for v in samp["avg"]], "err": list(err)}


def run_full_traces(nmc=8192):
    """The benchmark workload with Full accumulators (blocksize 1, the reference's default
    addObservable(obs)): 65,536 walkers x 8,192 steps, every step's x_i^2 (16) and |x|^2 stored,
    Correlated estimator (MJBlocker on 8,192 values per walker and component): 73 GB of traces in
    HBM.  Reports the trace kernel's HBM write rate (algorithmic bytes / kernel time) beside the
    8 TB/s peak, and the whole integrate."""
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    m = MCI(D, W)
    m.setSeed(SEED)
    m.setRandomMode(K.RNG_PHILOX)
    m.setWalkerCoupling(K.COUPLING_SHARED)
    m.setTrialMove(K.SRRD_UNIFORM, 1, 1)
    m.setNdecorrelationSteps(NDECORR)
    m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
    m.addObservable(K.OBS_X2, (), 1, 1, True, estimator=K.ESTIM_CORRELATED)
    m.addObservable(K.OBS_X2_SUM, (), 1, 1, True, estimator=K.ESTIM_CORRELATED)
    m.integrate(nmc)
    t0 = time.perf_counter()
    avg, err = m.integrate(nmc)
    ms = 1000 * (time.perf_counter() - t0)
    kt = m.getTimings()
    trace_bytes = W * nmc * (D + 1) * 8
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    z = (avg - exact) / err
    return {"workload": f"{W} walkers x {nmc} steps, d={D}, x_i^2 and |x|^2 as Full traces (blocksize 1), Correlated estimator",
            "trace_bytes": trace_bytes, "trace_kernel_ms": kt[2], "hbm_gbps": trace_bytes / (kt[2] / 1000) / 1e9,
            "hbm_frac_of_8tbs": trace_bytes / (kt[2] / 1000) / 8e12, "integrate_ms": ms,
            "phase_ms": {"calibration": kt[0], "decorrelation": kt[1], "sampling_kernel": kt[2], "estimation": kt[3]},
            "walker_steps_per_s": W * nmc / (ms / 1000), "max_abs_z": float(np.max(np.abs(z))), "mean_z2": float(np.mean(z * z))}


def run_reference_streams():
    """The benchmark workload on the reference's random streams (one mt19937_64 per walker, the
    library's default, which reproduces the reference's chains bit for bit), on the specialised
    kernel and, for comparison, on the generic one: whole integrate and sampling kernel.  The
    reference-stream kernel's traffic is the walkers' generator states: three outputs per step,
    each reading two state words and writing one (24 B)."""
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    out = {"workload": f"{W} walkers x {NMC} steps, d={D}, as the benchmark, mt19937_64 streams (MCI_RNG_MT19937_64)"}
    for name, path in (("specialised", K.KERNELS_AUTO), ("generic", K.KERNELS_GENERIC)):
        m = MCI(D, W)
        m.setSeed(SEED)
        m.setRandomMode(K.RNG_MT19937_64)
        m.setKernelPath(path)
        m.setWalkerCoupling(K.COUPLING_SHARED)
        m.setTrialMove(K.SRRD_UNIFORM, 1, 1)
        m.setNdecorrelationSteps(NDECORR)
        m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
        m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
        m.addObservable(K.OBS_X2_SUM, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
        if name == "specialised":
            m.integrate(NMC)
        t0 = time.perf_counter()
        avg, err = m.integrate(NMC)
        ms = 1000 * (time.perf_counter() - t0)
        kt = m.getTimings()
        wa, _ = m.walkerResults()
        z = (avg - exact) / (wa.std(axis=0, ddof=1) / np.sqrt(W))
        out[name] = {"integrate_ms": ms, "walker_steps_per_s": W * NMC / (ms / 1000), "sampling_kernel_ms": kt[2],
                     "max_abs_z": float(np.max(np.abs(z)))}
        if name == "specialised":
            out[name]["generator_state_gbps"] = 3 * 24 * W * NMC / (kt[2] / 1000) / 1e9
            out[name]["avg"] = avg
    out["speedup_specialised_vs_generic"] = out["generic"]["integrate_ms"] / out["specialised"]["integrate_ms"]
    out["specialised"].pop("avg")
    return out


def run_all_coordinate_moves():
    """The benchmark's target, observables and walkers with the reference's default move (every
    coordinate at once, one step size, SRRDAllMove) instead of the single-coordinate move, on both
    random-stream modes, on the specialised all-coordinate kernel and, for comparison, on the
    generic one: whole integrate and sampling kernel.  A step reads and writes the walker's 16
    coordinates in LDS; the native streams' Philox blocks (nine per step at d = 16) bound it."""
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    out = {"workload": f"{W} walkers x {NMC} steps, d={D}, as the benchmark with all-coordinate moves (veclen 0)"}
    for rname, rng in (("native", K.RNG_PHILOX), ("reference_streams", K.RNG_MT19937_64)):
        r = {}
        for name, path in (("specialised", K.KERNELS_AUTO), ("generic", K.KERNELS_GENERIC)):
            m = MCI(D, W)
            m.setSeed(SEED)
            m.setRandomMode(rng)
            m.setKernelPath(path)
            m.setWalkerCoupling(K.COUPLING_SHARED)
            m.setTrialMove(K.SRRD_UNIFORM, 0, 1)
            m.setNdecorrelationSteps(NDECORR)
            m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
            m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
            m.addObservable(K.OBS_X2_SUM, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
            if name == "specialised":
                m.integrate(NMC)
            t0 = time.perf_counter()
            avg, err = m.integrate(NMC)
            ms = 1000 * (time.perf_counter() - t0)
            kt = m.getTimings()
            wa, _ = m.walkerResults()
            z = (avg - exact) / (wa.std(axis=0, ddof=1) / np.sqrt(W))
            r[name] = {"integrate_ms": ms, "walker_steps_per_s": W * NMC / (ms / 1000), "sampling_kernel_ms": kt[2],
                       "max_abs_z": float(np.max(np.abs(z))), "acceptance": float(m.getAcceptanceRate())}
        r["speedup_specialised_vs_generic"] = r["generic"]["integrate_ms"] / r["specialised"]["integrate_ms"]
        out[rname] = r
    # beyond the register width: d = 64 (the walker in LDS), native streams, the specialised kernel
    d64 = 64
    s64 = np.linspace(0.5, 2.0, d64)
    m = MCI(d64, W)
    m.setSeed(SEED)
    m.setRandomMode(K.RNG_PHILOX)
    m.setWalkerCoupling(K.COUPLING_SHARED)
    m.setTrialMove(K.SRRD_UNIFORM, 0, 1)
    m.setNdecorrelationSteps(NDECORR)
    m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / s64 ** 2)
    m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
    m.integrate(NMC)
    t0 = time.perf_counter()
    avg, err = m.integrate(NMC)
    ms = 1000 * (time.perf_counter() - t0)
    wa, _ = m.walkerResults()
    z = (avg - s64 ** 2) / (wa.std(axis=0, ddof=1) / np.sqrt(W))
    out["native_d64"] = {"workload": f"{W} walkers x {NMC} steps, d=64, sigma_i = linspace(0.5, 2, 64), x_i^2 in 64 blocks of 160",
                         "integrate_ms": ms, "sampling_kernel_ms": m.getTimings()[2], "walker_steps_per_s": W * NMC / (ms / 1000),
                         "max_abs_z": float(np.max(np.abs(z)))}
    return out


DIAG_SRC = r"""
__device__ void diag_proto(int ndim, MciVec x, const double *par, MciVec proto)
{
    for (int i = 0; i < ndim; ++i) { const double v = x[i]; proto[i] = (v * v) * par[i]; }
}
__device__ double diag_acceptance(int nproto, MciVec po, MciVec pn, const double *par)
{
    double e = 0.0;
    for (int i = 0; i < nproto; ++i) { e = e + po[i]; }
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pn[i]; }
    return exp(e - s);
}
__device__ double diag_sampling(int nproto, MciVec pv, const double *par)
{
    double s = 0.0;
    for (int i = 0; i < nproto; ++i) { s = s + pv[i]; }
    return exp(-s);
}
"""
XSQ_SRC = r"""
__device__ void xsq_observable(int ndim, MciVec x, const double *par, MciVec out)
{
    for (int i = 0; i < ndim; ++i) { const double v = x[i]; out[i] = v * v; }
}
"""


def run_device_functors():
    """The benchmark's target and x_i^2 written by the caller as device functors (compiled at run
    time), against the same catalogue entries on the generic kernel and on the specialised
    all-coordinate kernel: 65,536 walkers x 2,560 steps, d = 16, all-coordinate moves, 16 blocks of
    160, shared calibration, 1,000 decorrelation steps."""
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    a = 0.5 / sig ** 2
    nmc = 2560
    out = {"workload": f"{W} walkers x {nmc} steps, d={D}, all-coordinate moves, x_i^2 in 16 blocks of 160"}
    res = {}
    for name in ("device_functors", "catalogue_generic", "catalogue_specialised"):
        m = MCI(D, W)
        m.setSeed(SEED)
        m.setRandomMode(K.RNG_PHILOX)
        m.setWalkerCoupling(K.COUPLING_SHARED)
        m.setTrialMove(K.SRRD_UNIFORM, 0, 1)
        m.setNdecorrelationSteps(NDECORR)
        if name == "catalogue_generic":
            m.setKernelPath(K.KERNELS_GENERIC)
        if name == "device_functors":
            m.addSamplingFunctionDevice("diag", DIAG_SRC, D, a)
            m.addObservableDevice("xsq", XSQ_SRC, D, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
            t0 = time.perf_counter()
            m.buildDeviceFunctors()
            out["build_s"] = time.perf_counter() - t0
        else:
            m.addSamplingFunction(K.PDF_GAUSS_DIAG, a)
            m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
        m.integrate(nmc)
        t0 = time.perf_counter()
        avg, err = m.integrate(nmc)
        ms = 1000 * (time.perf_counter() - t0)
        res[name] = avg
        out[name] = {"integrate_ms": ms, "sampling_kernel_ms": m.getTimings()[2], "walker_steps_per_s": W * nmc / (ms / 1000)}
    out["device_equals_catalogue_generic"] = bool(np.array_equal(res["device_functors"], res["catalogue_generic"]))
    out["device_equals_catalogue_specialised"] = bool(np.array_equal(res["device_functors"], res["catalogue_specialised"]))
    return out


def run_reference_defaults(shared=True):
    """A drop-in user's integrate with the reference's defaults: MCI(16) with 65,536 walkers, the
    default move (all coordinates, one step size), the default random streams (mt19937_64 per
    walker), automatic step calibration and automatic decorrelation, and addObservable(obs) with
    its defaults (a Full accumulator, the Correlated estimator) for x_i^2, on the benchmark's target;
    8,192 steps (MJBlocker on the traces).  Walkers as ranks sharing calibration and decorrelation
    (shared), or as independent MCI objects (the default coupling: each walker calibrates and
    decorrelates by itself, on 800-step rounds).  Phases by HIP events."""
    from mcintegratorplusplus_amd import MCI, constants as K
    sig = sigmas()
    nmc = 8192
    out = {"workload": f"{W} walkers x {nmc} steps, d={D}, reference defaults (all-coordinate move, mt19937_64, "
                       "automatic calibration and decorrelation, Full accumulator + Correlated estimator), "
                       + ("shared coupling" if shared else "independent walkers")}
    m = MCI(D, W)
    m.setSeed(SEED)
    m.setWalkerCoupling(K.COUPLING_SHARED if shared else K.COUPLING_INDEPENDENT)
    m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
    m.addObservable(K.OBS_X2)
    m.integrate(nmc)
    t0 = time.perf_counter()
    avg, err = m.integrate(nmc)
    ms = 1000 * (time.perf_counter() - t0)
    kt = m.getTimings()
    z = (avg - sig ** 2) / err
    out.update({"integrate_ms": ms, "walker_steps_per_s": W * nmc / (ms / 1000),
                "phase_ms": {"calibration": kt[0], "decorrelation": kt[1], "sampling": kt[2], "estimation": kt[3]},
                "max_abs_z": float(np.max(np.abs(z)))})
    return out


def profile_summary():
    """HBM bytes per launch of the sampling kernel (FETCH_SIZE + WRITE_SIZE, each scaled by a
    1 GiB copy probe) and its VALU-busy fraction, from the committed rocprofv3 summary."""
    p = os.path.join(ROOT, "profiles", "traffic_cfg2_sample.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


def run_precision_form(with_cpu=True):
    """d = 64, 262,144 walkers x 10,240 steps (x_i^2 and exp(-|x|^2/64) in 64 blocks of 160,
    MJBlocker, shared calibration, 1,000 decorrelation steps): the FP32 chain (Kahan block sums)
    against the FP64 chain on the same Philox words and against the analytic values."""
    from mcintegratorplusplus_amd import MCI, constants as K
    with open(os.path.join(ROOT, "tests", "golden", "sigmas.json")) as f:
        sig = np.array([float.fromhex(v) for v in json.load(f)["sigma64"]])
    Wp, D64 = 262144, 64
    exact = np.concatenate([sig ** 2, [np.prod((1 + 2 * sig ** 2 / D64) ** -0.5)]])
    out = {}
    for prec in (0, 1):
        m = MCI(D64, Wp)
        m.setSeed(4242)
        m.setRandomMode(K.RNG_PHILOX)
        m.setPrecision(prec)
        m.setWalkerCoupling(K.COUPLING_SHARED)
        m.setTrialMove(K.SRRD_UNIFORM, 1, 1)
        m.setNdecorrelationSteps(NDECORR)
        m.addSamplingFunction(K.PDF_GAUSS_DIAG, 0.5 / sig ** 2)
        m.addObservable(K.OBS_X2, (), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
        m.addObservable(K.OBS_EXP_X2, (float(D64),), BLOCK, 1, True, estimator=K.ESTIM_MJBLOCKER)
        m.integrate(NMC)
        t0 = time.perf_counter()
        avg, err = m.integrate(NMC)
        ms = 1000 * (time.perf_counter() - t0)
        wa, _ = m.walkerResults()
        out[prec] = (avg, err, ms, m.getTimings(), wa.std(axis=0, ddof=1) / np.sqrt(Wp))
    (a64, e64, ms64, kt64, s64), (a32, e32, ms32, kt32, s32) = out[0], out[1]
    # k_fast32 against the FP32 VALU peak, counted as 270 flops per walker-step: proposal 4,
    # Delta log p 3, |x|^2 update 2, exp argument 1, and 4 (a Kahan add) for each of the 65
    # observable components every step.  The kernel sums by dwell time (a component's term enters
    # its Kahan sum when the value changes), so it executes far fewer of those 260.
    f32 = 270 * Wp * NMC / (kt32[2] / 1000) / 1e12
    return {"workload": "d=64 anisotropic Gaussian, 262144 walkers x 10240 steps, x_i^2 + exp(-|x|^2/64)",
            "fp64": {"ms": ms64, "walker_steps_per_s": Wp * NMC / (ms64 / 1000), "sampling_kernel_ms": kt64[2],
                     "max_abs_z": float(np.max(np.abs(a64 - exact) / s64)),
                     "max_abs_z_mjblocker_errors": float(np.max(np.abs(a64 - exact) / e64))},
            "fp32": {"ms": ms32, "walker_steps_per_s": Wp * NMC / (ms32 / 1000), "sampling_kernel_ms": kt32[2],
                     "max_abs_z": float(np.max(np.abs(a32 - exact) / s32)),
                     "max_abs_z_mjblocker_errors": float(np.max(np.abs(a32 - exact) / e32)),
                     "roofline": {"bound": "valu", "achieved": f32, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                  "frac": f32 / PEAK_FP32_TFLOPS, "kernel": "k_fast32 sampling kernel (10240 steps x 262144 walkers)",
                                  "kernel_ms": kt32[2], "flops_per_walker_step": 270,
                                  "hbm_gbps": 65 * 8 * Wp * (NMC // BLOCK) / (kt32[2] / 1000) / 1e9}},
            "error_bar": "max_abs_z: standard error of the walker means; the MJBlocker combine runs low at d = 64",
            "fp32_vs_fp64_max_rel_diff": float(np.max(np.abs(a32 - a64) / np.abs(a64))),
            "fp32_vs_fp64_max_diff_over_err": float(np.max(np.abs(a32 - a64) / s64)),
            "reference_cpu": cpu_precision_form(sig) if with_cpu else None}


def cpu_precision_form(sig, budget_s=8.0):
    """The reference CPU path on the precision-form workload (d = 64, x_i^2 and exp(-|x|^2/64)):
    walker-steps/s of a bounded sample of independent walkers."""
    r = cpu_reference_gauss(64, sig, f"obs 1 0 {BLOCK} 1 1 4\nobs 6 1 0x1p+6 {BLOCK} 1 1 4\n", budget_s)
    if r is None:
        return None
    return {"walkers": r["walkers"], "steps_per_walker": NMC, "processes": r["procs"], "wall_s": r["wall"],
            "walker_steps_per_s": r["steps"] / r["wall"]}


def cpu_batched_sample(ndim, nmc, w, s, mu, budget_s=8.0):
    """The reference CPU path on the batched workload: each integral as the reference computes one
    (one MCI, the isotropic-mixture plugin of oracle/ref_harness.cpp, nmc single-coordinate steps
    with its own findMRT2Step and automatic decorrelation, |x|^2 in a simple accumulator), on a
    subsample of every k-th integral sized to about budget_s of CPU work, P processes."""
    binary = os.path.join(ROOT, "oracle", "_ref", "mci_ref_fast")
    if not os.path.exists(binary):
        return None
    n = len(ndim)

    def block(i):
        K = len(w[i])
        p = np.concatenate([[K], w[i], s[i], np.asarray(mu[i]).ravel()])
        a = " ".join(float(v).hex() for v in p)
        return (f"bench 1 {int(nmc[i])} {9000 + i}\ncreate {int(ndim[i])}\nmove 0 1 1\npdf 4 {len(p)} {a}\n"
                f"obs 2 0 0 1 0 0\nendbench\n")

    # per-step cost from a small probe, then a subsample of every k-th integral
    probe = list(range(0, n, max(1, n // 16)))
    t0 = time.time()
    subprocess.run([binary], input="".join(block(i) for i in probe), capture_output=True, text=True, check=True)
    rate1 = sum(int(nmc[i]) for i in probe) / max(time.time() - t0, 1e-3)
    procs, _ = cpu_share()
    k = max(1, int(np.ceil(float(np.sum(nmc)) / (rate1 * procs * budget_s))))
    chosen = list(range(0, n, k))
    shares = [chosen[r::procs] for r in range(procs)]
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        ins = []
        for r, sh in enumerate(shares):
            fn = os.path.join(tmp, f"in{r}.txt")
            with open(fn, "w") as f:
                f.write("".join(block(i) for i in sh))
            ins.append(fn)
        t0 = time.time()
        ps = []
        for r, fn in enumerate(ins):  # inputs and outputs through files: no pipe can fill up
            ps.append(subprocess.Popen([binary], stdin=open(fn), stdout=open(os.path.join(tmp, f"out{r}.txt"), "w")))
        for pr in ps:
            pr.wait()
        wall = time.time() - t0
    steps = int(sum(int(nmc[i]) for i in chosen))
    rate = steps / wall
    return {"integrals": len(chosen), "of": n, "walker_steps": steps, "processes": procs, "wall_s": wall,
            "walker_steps_per_s": rate, "seconds_for_all": float(np.sum(nmc)) / rate}


def run_batched(reps=5, with_cpu=True):
    """4,096 independent Gaussian-mixture integrals (d log-uniform in [1, 128], nmc log-uniform in
    [1e3, 1e5], 1-4 isotropic components), |x|^2, FP32 walkers, one hipGraph replay per batch."""
    from mcintegratorplusplus_amd import BatchIntegrals, mixture_workload
    ndim, nmc, w, s, mu, exact = mixture_workload(n=4096, seed=4096)
    b = BatchIntegrals(ndim, nmc, w, s, mu)
    b.setSeed(1337)
    b.autotune(3)
    R, T, _ = b.plan()
    b.setGraph(False)
    b.run()
    phases = b.timings()
    b.setGraph(True)
    b.run()
    t0 = time.perf_counter()
    for _ in range(reps):
        avg, err = b.run()
    ms = 1000 * (time.perf_counter() - t0) / reps
    steps = int(np.sum(R * T))
    z = (avg - exact) / err
    out = {"workload": "4096 mixture integrals, d in [1,128], nmc in [1e3,1e5], |x|^2",
            "ms_per_batch": ms, "walker_steps": steps, "walker_steps_per_s": steps / (ms / 1000),
            "requested_steps": int(np.sum(nmc)), "walkers": int(np.sum(R)),
            "phase_ms_eager": {"calibration": phases[0], "sampling": phases[1], "estimation": phases[2]},
            "frac_abs_z_below_3": float(np.mean(np.abs(z) < 3)), "mean_z2": float(np.mean(z * z))}
    cpu = cpu_batched_sample(ndim, nmc, w, s, mu) if with_cpu else None
    if cpu is not None:
        out["reference_cpu"] = cpu
        out["speedup_vs_reference_cpu"] = cpu["seconds_for_all"] / (ms / 1000)
    return out


def cpu_learned_target(w, mu, cov, var, nsteps, budget_s=10.0):
    """The reference CPU path (oracle/_ref/mci_ref_fast) on the learned proposal's target: the
    full-covariance mixture as a reference sampling-function plugin with incremental
    single-coordinate updates, one MCI per walker (exact start, own step calibration), nsteps
    steps; P processes.  ESS from the spread of the walker means, as on the GPU."""
    binary = os.path.join(ROOT, "oracle", "_ref", "mci_ref_fast")
    if not os.path.exists(binary):
        return None
    K, d = mu.shape
    P = np.linalg.inv(cov)
    c = np.log(w) - 0.5 * np.array([np.linalg.slogdet(2 * np.pi * cv)[1] for cv in cov])
    L = np.linalg.cholesky(cov)
    p = np.concatenate([[K], w, mu.ravel(), P.ravel(), c, L.ravel()])
    a = " ".join(float(v).hex() for v in p)
    bs = max(1, nsteps // 64)
    script = f"create {d}\nmove 0 1 1\nndecorr 0\npdf 5 {len(p)} {a}\nobs 2 0 {bs} 1 1 4\nobs 7 1 0 {bs} 1 1 4\nendbench\n"
    procs, _ = cpu_share()
    t0 = time.time()
    subprocess.run([binary], input=f"bench 4 {nsteps} 5001\n" + script, capture_output=True, text=True, check=True)
    per_walker = max((time.time() - t0) / 4, 1e-4)
    per_proc = int(min(20000, max(8, budget_s / per_walker)))
    t0 = time.time()
    ps = [subprocess.Popen([binary], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True) for _ in range(procs)]
    for i, pr in enumerate(ps):
        pr.stdin.write(f"bench {per_proc} {nsteps} {700000 + i * per_proc}\n" + script)
        pr.stdin.close()
    outs = [json.loads([l for l in pr.stdout.read().splitlines() if l.startswith("{")][-1]) for pr in ps]
    for pr in ps:
        pr.wait()
    wall = time.time() - t0
    nw = sum(o["walkers"] for o in outs)
    s1 = np.sum([[float.fromhex(v) for v in o["avg_sum"]] for o in outs], axis=0)
    s2 = np.sum([[float.fromhex(v) for v in o["avg2_sum"]] for o in outs], axis=0)
    mean = s1 / nw
    se = np.sqrt(np.maximum(s2 / nw - mean * mean, 0.0) / (nw - 1))
    ess_s = var / se ** 2 / wall
    return {"walkers": nw, "steps": nsteps, "processes": procs, "wall_s": wall, "err": list(se),
            "walker_steps_per_s": nw * nsteps / wall, "ess_per_s": list(ess_s)}


def run_learned(with_cpu=True):
    """d = 32 mixture of 8 rotated anisotropic Gaussians, 65,536 walkers x 10,240 steps: the plain
    FP64 random-walk chain against the FP32 chains that mix random-walk moves with moves drawn
    from the trained network (MH-corrected; cycles of 32 steps, a fraction pmix of them learned):
    the conditional proposal q(x' | x) and the independence proposal q(x') (a second network
    trained with one fixed input), and the reference CPU path on the same target (its ESS/s set
    against the best learned configuration).  ESS = Var_p(f) / err^2 with the exact
    variance and err the standard error of the walker means; time = calibration + sampling.  The
    network starts from a 60-iteration EM fit of its 8-component diagonal mixture to the last
    snapshot and is trained on pairs of independent snapshots of the plain FP32 chain (16
    snapshots, 512 steps apart; 2 x 1,000 Adam steps of 4,096 pairs, learning rate 1e-3, 5e-4)."""
    from mcintegratorplusplus_amd.learned import INDEP_FP32, MIXED_FP32, PLAIN_FP64, LearnedProposalSampler, mixture_target
    w, mu, cov, (ex2, pind) = mixture_target(d=32, ncomp=8, seed=5)
    ex4 = float(np.sum(w * [(np.trace(c) + m @ m) ** 2 + 2 * np.trace(c @ c) + 4 * m @ c @ m for m, c in zip(mu, cov)]))
    var = np.array([ex4 - ex2 ** 2, pind * (1 - pind)])
    exact = np.array([ex2, pind])
    Wl = 65536
    s = LearnedProposalSampler(w, mu, cov, Wl)
    s.setSeed(1)
    a_p, e_p = s.run(PLAIN_FP64, NMC, 64)
    t_p = s.timings()
    sec_p = (t_p["sampling"] + t_p["calibration"]) / 1000
    ess_p = var / e_p ** 2 / sec_p
    s.setTrainingPairs(0.0)
    t_c0 = time.perf_counter()
    s.collect(16, 512)
    t_col = time.perf_counter() - t_c0
    t0 = time.perf_counter()
    s.initFromSamples(60)
    t_em = 1000 * (time.perf_counter() - t0)
    t_train, losses = 0.0, []
    for lr in (1e-3, 5e-4):
        l0, l1 = s.train(1000, 4096, lr)
        t_train += s.timings()["training"]
        losses.append(l1)
    out = {"workload": "d=32, 8 rotated anisotropic Gaussians, 65536 walkers x 10240 steps, |x|^2 and 1[x_0>0]",
           "plain_fp64": {"seconds": sec_p, "err": list(e_p), "ess_per_s": list(ess_p), "z": list((a_p - exact) / e_p)},
           "em_start_ms": t_em, "training_ms": t_train, "training_loss_after_each_stage": losses}
    for pm in (0.25, 0.5):
        a_m, e_m = s.run(MIXED_FP32, NMC, 64, pm)
        t_m = s.timings()
        sec_m = (t_m["sampling"] + t_m["calibration"]) / 1000
        ess_m = var / e_m ** 2 / sec_m
        oneoff = t_col + t_em / 1000 + t_train / 1000
        out[f"learned_fp32_pmix{pm}"] = {"seconds": sec_m, "err": list(e_m), "ess_per_s": list(ess_m),
                                         "z": list((a_m - exact) / e_m), "learned_acceptance": s.results()["acc_learned"],
                                         "ess_per_s_gain_vs_plain_fp64": list(ess_m / ess_p),
                                         "ess_per_step_gain": list((e_p / e_m) ** 2),
                                         "oneoff_s": oneoff,
                                         "ess_per_s_gain_vs_plain_fp64_with_oneoff": list(var / e_m ** 2 / (sec_m + oneoff) / ess_p)}
    # the independence form: a second network trained with one fixed input, its mixture there the
    # proposal q(x') of every learned move (a fraction pmix of each cycle of 32 steps).  Three ways to
    # set the network: the EM fit written into its last layer with no training ("em_only"), Adam from
    # the network's random start with no EM fit ("adam_only", 4 x 1,000 steps), and both ("em_adam",
    # the default); the one-off cost of each (collection, EM start, training: host clock around the
    # synchronous calls) is reported and counted in the *_with_oneoff gains.
    def indep_row(key, use_em, stages, pms):
        s.setNetwork(128, 8)
        s.setFixedInput(True)
        t0 = time.perf_counter()
        s.collect(16, 512)
        t_col = time.perf_counter() - t0
        t1 = time.perf_counter()
        if use_em:
            s.initFromSamples(60)
        t_em = time.perf_counter() - t1
        t2 = time.perf_counter()
        ls = [s.train(1000, 4096, lr)[1] for lr in stages]
        t_tr = time.perf_counter() - t2
        oneoff = t_col + t_em + t_tr
        for pm in pms:
            a_i, e_i = s.run(INDEP_FP32, NMC, 64, pm)
            t_i = s.timings()
            sec_i = (t_i["sampling"] + t_i["calibration"]) / 1000
            ess_i = var / e_i ** 2 / sec_i
            ess_all = var / e_i ** 2 / (sec_i + oneoff)
            out[f"independence_{key}_fp32_pmix{pm}"] = {
                "seconds": sec_i, "err": list(e_i), "ess_per_s": list(ess_i), "z": list((a_i - exact) / e_i),
                "learned_acceptance": s.results()["acc_learned"], "ess_per_s_gain_vs_plain_fp64": list(ess_i / ess_p),
                "ess_per_step_gain": list((e_p / e_i) ** 2),
                "oneoff_s": {"collection": t_col, "em_start": t_em, "training": t_tr},
                "ess_per_s_gain_vs_plain_fp64_with_oneoff": list(ess_all / ess_p), "training_loss_after_each_stage": ls}
    indep_row("em_only", True, (), (0.5,))
    indep_row("adam_only", False, (1e-3, 1e-3, 5e-4, 5e-4), (0.5,))
    indep_row("em_adam", True, (1e-3, 5e-4), (0.5, 0.75))
    # the best learned configuration by the smaller of its two ESS/s gains
    runs = [k for k in out if k.startswith(("learned_fp32", "independence_"))]
    best = max(runs, key=lambda k: min(out[k]["ess_per_s_gain_vs_plain_fp64"]))
    out["best"] = best
    cpu = cpu_learned_target(w, mu, cov, var, NMC) if with_cpu else None
    if cpu is not None:
        out["reference_cpu"] = cpu
        out["ess_per_s_gain_vs_reference_cpu"] = list(np.array(out[best]["ess_per_s"]) / np.array(cpu["ess_per_s"]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", type=int, default=0, help="0 = FP64 (the benchmark), 1 = FP32")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--generic", action="store_true", help="force the generic kernels (comparison)")
    ap.add_argument("--workloads", default="all", choices=["all", "none", "precision", "batched", "learned", "full", "mt", "allmove", "functors", "defaults"],
                    help="also measure the precision form, the batched integrals and the learned proposal (all), "
                         "none of them, or one of them")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain device copy of 1 GiB (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 1 << 27  # 2^27 doubles = 1 GiB
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; the walkers of one GPU are the unit of "
                                   "work, and sharding walkers across GPUs is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)

    m = make_mci(W, args.precision)
    if args.generic:
        m.setKernelPath(1)
    for _ in range(args.warmup):
        m.integrate(NMC)
    times, ktimes, results = [], [], None
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        avg, err = m.integrate(NMC)
        times.append(time.perf_counter() - t0)
        ktimes.append(m.getTimings())
        results = (avg, err)
    total = time.perf_counter() - t_all
    wavg, _ = m.walkerResults()  # after the timed region: the walker means of the last integrate
    err_spread = wavg.std(axis=0, ddof=1) / np.sqrt(W)
    ms_per_step = 1000 * total / args.steps
    kt = np.mean(ktimes, axis=0)  # calibration, decorrelation, sampling kernel, estimation, total (ms)
    value = W * NMC / (ms_per_step / 1000)
    avg, err = results
    ess = ess_per_obs(err_spread, W * NMC)
    ess_mj = ess_per_obs(err, W * NMC)
    sig = sigmas()
    exact = np.concatenate([sig ** 2, [np.sum(sig ** 2)]])
    achieved = FLOPS_PER_STEP * W * NMC / (kt[2] / 1000) / 1e12
    prof = profile_summary() if args.precision == 0 else {}
    line = {
        "metric": "MC steps/s and ESS/s per GPU at matched error bar vs plain FP64 Metropolis",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32",
        "data": "synthetic",
        "config": {"workload": f"{W} walkers x {NMC} steps, d={D} anisotropic Gaussian (sigma_i~U(0.5,2), seed 1337), "
                               f"single-coordinate uniform moves, shared step calibration, {NDECORR} decorrelation "
                               f"steps, x_i^2 and |x|^2 in 64 blocks of {BLOCK}, MJBlocker, MPIMCI combine",
                   "walkers": W, "nmc": NMC, "ndim": D, "rng": "philox4x32-10",
                   "parallelism": "one walker per lane, 1 GPU"},
        "ess_per_s": float(np.min(ess)) / (ms_per_step / 1000),
        "ess_error_bar": "standard error of the 65536 walker means (the MJBlocker combine is 14 % smaller here)",
        "ess_per_s_with_mjblocker_errors": float(np.min(ess_mj)) / (ms_per_step / 1000),
        "phase_ms": {"calibration": kt[0], "decorrelation": kt[1], "sampling_kernel": kt[2], "estimation": kt[3]},
        "max_abs_z_vs_analytic": float(np.max(np.abs(avg - exact) / err_spread)),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64, "traffic": prof.get("bytes_per_launch"),
                     "kernel": "k_fast64 sampling kernel (10240 steps x 65536 walkers)",
                     "kernel_ms": kt[2], "flops_per_walker_step": FLOPS_PER_STEP,
                     "valu_busy_rocprof": prof.get("valu_busy"),
                     "instructions_per_wave_step_rocprof": (prof.get("instructions_per_walker_step") or {}).get("SQ_INSTS"),
                     # one wave per SIMD: each instruction of the wave takes an issue slot of 4 cycles;
                     # the share of the step's cycles those slots fill (rocprof counts)
                     "issue_slots_frac": (4 * prof["instructions_per_walker_step"]["SQ_INSTS"] / prof["cycles_per_wave_step"])
                     if "cycles_per_wave_step" in prof else None,
                     "hbm_gbps": (prof["bytes_per_launch"] / (kt[2] / 1000) / 1e9) if "bytes_per_launch" in prof else None},
    }
    if args.workloads != "none" and args.precision == 0 and not args.generic:
        runs = {"precision_fp32": ("precision", lambda: run_precision_form(not args.no_cpu)), "batched": ("batched", lambda: run_batched(with_cpu=not args.no_cpu)),
                "learned_proposal": ("learned", lambda: run_learned(not args.no_cpu)), "full_traces": ("full", run_full_traces),
                "reference_streams": ("mt", run_reference_streams), "all_coordinate_moves": ("allmove", run_all_coordinate_moves),
                "device_functors": ("functors", run_device_functors), "reference_defaults": ("defaults", run_reference_defaults),
                "reference_defaults_independent": ("defaults", lambda: run_reference_defaults(False))}
        line["other_workloads"] = {}
        for k, (name, f) in runs.items():
            if args.workloads in ("all", name):
                print(f"bench: {k}", file=sys.stderr, flush=True)
                line["other_workloads"][k] = f()
    if not args.no_cpu:
        cb = cpu_baseline()
        if cb is not None:
            line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
            line["cpu_baseline"]["ess_per_s"] = cb["ess_per_s"]
            line["cpu_baseline"]["host_cpus_visible"] = cpu_share()[1]
            line["cpu_baseline"]["cores_note"] = ("processes = the host's CPU share for one GPU (OMP_NUM_THREADS), "
                                                  "at most the CPUs this process may use")
            c1 = cpu_single_core_cfg1()
            if c1 is not None:
                line["cpu_baseline"]["cfg1_single_core"] = c1
            line["speedup_vs_cpu_steps"] = value / cb["value"]
            line["speedup_vs_cpu_ess"] = line["ess_per_s"] / cb["ess_per_s"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
