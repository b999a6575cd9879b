"""Writes the golden vectors under tests/golden/ by running the REFERENCE itself.

The reference is built from its own sources by oracle/Makefile into
oracle/_ref/mci_ref (-O2 -ffp-contract=off).  This script feeds it scenario
scripts (oracle/ref_harness.cpp) and stores inputs and outputs as JSON with
hex-float doubles.  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py

Nothing here is imported by the tests; they read the JSON files only.
"""
import base64
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from mcitest import (ESTIM_CORR, ESTIM_FC, ESTIM_MJ, ESTIM_NOOP, ESTIM_UNCORR, OBS_CONST, OBS_DEP_SQUARE, OBS_EXP_X2,  # noqa: E402
                     OBS_GAUSS_X2, OBS_INDICATOR, OBS_PDF_LOG, OBS_PDF_PROTO, OBS_X, OBS_X2, OBS_X2_SUM, OBS_X_SUM, PDF_EXP_DIAG, PDF_GAUSS_DIAG,
                     PDF_GAUSS_ISO, PDF_GMM_ISO, REF_BIN, ROOT, fmt, ref_run, script_setup)

FLAGS = "g++ 11.4 -std=c++14 -O2 -ffp-contract=off -include stdexcept (oracle/Makefile target _ref/mci_ref)"


def b64(a):
    """little-endian float64 bytes, base64 (bit-exact and compact)"""
    return base64.b64encode(np.ascontiguousarray(a, dtype="<f8").tobytes()).decode()


def probe(cmd):
    out = subprocess.run([os.path.join(ROOT, "oracle", "build", "stdrng_probe")], input=cmd + "\n",
                         capture_output=True, text=True, check=True).stdout.split()
    return out[:-1]


def sigmas(n):
    """sigma_i ~ U(0.5, 2): std::uniform_real_distribution<double>(0.5, 2) on std::mt19937_64(1337)."""
    return [float.fromhex(v) for v in probe(f"real 1337 {n} 0.5 2.0")]


def cfg2_a(n):
    return [0.5 / (s * s) for s in sigmas(n)]


def scenarios():
    S = []
    # configuration 1 of the benchmark set: d=3 Gaussian, XYZSquared blocks of 62500, MJBlocker, seed 1337
    S.append(dict(name="cfg1", cfg=dict(ndim=3, seed=1337, pdfs=[(PDF_GAUSS_ISO, [])],
                                        obs=[(OBS_X2, [], 62500, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=1000000, find=1, decorr=1)]))
    # the sequence of test/ut2/main.cpp: bad start without/with calibration, continuation, PBC range
    S.append(dict(name="ut2_sequence", cfg=dict(ndim=3, seed=5649871, pdfs=[(PDF_GAUSS_ISO, [])],
                                                obs=[(OBS_X2_SUM, [3.0], 1, 1, 1, ESTIM_CORR)]),
                  calls=[dict(nmc=10000, find=0, decorr=0, setx=[5., -5., 10.]),
                         dict(nmc=10000, find=1, decorr=1, setx=[5., -5., 10.]),
                         dict(nmc=10000, find=0, decorr=0),
                         dict(nmc=10000, find=0, decorr=0, irange=([-5.] * 3, [5.] * 3))]))
    # test/ut3/main.cpp: a 1-d and a 3-d observable together
    S.append(dict(name="ut3_sequence", cfg=dict(ndim=3, seed=5649871, pdfs=[(PDF_GAUSS_ISO, [])],
                                                obs=[(OBS_X2_SUM, [3.0], 1, 1, 1, ESTIM_CORR), (OBS_X2, [], 1, 1, 1, ESTIM_CORR)]),
                  calls=[dict(nmc=10000, find=0, decorr=0, setx=[5., -5., 10.]),
                         dict(nmc=10000, find=1, decorr=1, setx=[5., -5., 10.]),
                         dict(nmc=10000, find=0, decorr=0)]))
    # test/ut4/main.cpp: fixed calibration/decorrelation, power-of-two NMC, blocks and nskip
    S.append(dict(name="ut4_fixed", cfg=dict(ndim=3, seed=1337, nfind=20, ndecorr=2000, pdfs=[(PDF_GAUSS_ISO, [])],
                                             obs=[(OBS_X2_SUM, [3.0], 1, 1, 1, ESTIM_CORR),
                                                  (OBS_X2_SUM, [3.0], 16, 1, 1, ESTIM_UNCORR),
                                                  (OBS_X2_SUM, [3.0], 8, 2, 0, ESTIM_CORR)], x0=[5., -5., 10.]),
                  calls=[dict(nmc=16384, find=1, decorr=1), dict(nmc=16384, find=0, decorr=0)]))
    # test/ut5/main.cpp style: vector moves with types and nskip
    for veclen, ntypes, te in [(1, 1, None), (1, 2, [1, 3]), (3, 1, None), (0, 2, [1, 3])]:
        nskip = 3 if veclen == 1 else 1
        S.append(dict(name=f"ut5_vec{veclen}_types{ntypes}",
                      cfg=dict(ndim=3, seed=1337, pdfs=[(PDF_GAUSS_DIAG, [])], move=(0, veclen, ntypes, te),
                               obs=[(OBS_X2_SUM, [3.0], 1, nskip, 1, ESTIM_CORR), (OBS_X2, [], 1, nskip, 1, ESTIM_CORR)]),
                      calls=[dict(nmc=8192 * nskip, find=1, decorr=0)]))
    # configuration 2 shape, one walker: d=16 anisotropic Gaussian, single-coordinate moves,
    # 1000 decorrelation steps, 64 blocks of 160, MJBlocker
    S.append(dict(name="cfg2_walker", cfg=dict(ndim=16, seed=1337, ndecorr=1000, move=(0, 1, 1, None),
                                               pdfs=[(PDF_GAUSS_DIAG, cfg2_a(16))],
                                               obs=[(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 160, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=10240, find=1, decorr=1)]))
    # configuration 3 shape, one walker (FP64): d=64, x_i^2 and exp(-|x|^2/64)
    S.append(dict(name="cfg3_walker", cfg=dict(ndim=64, seed=4242, ndecorr=1000, move=(0, 1, 1, None),
                                               pdfs=[(PDF_GAUSS_DIAG, cfg2_a(64))],
                                               obs=[(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_EXP_X2, [64.0], 160, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=10240, find=1, decorr=1)]))
    # the reference's default move (all coordinates, one step size) on configuration 2's target:
    # Block accumulators, then a continuation; d = 5 with Full traces; d = 7 with a Simple accumulator
    S.append(dict(name="cfg2_allmove_walker", cfg=dict(ndim=16, seed=1337, ndecorr=1000, pdfs=[(PDF_GAUSS_DIAG, cfg2_a(16))],
                                                       obs=[(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 160, 1, 1, ESTIM_MJ),
                                                            (OBS_EXP_X2, [16.0], 160, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=10240, find=1, decorr=1), dict(nmc=5120, find=0, decorr=0)]))
    S.append(dict(name="allmove_full", cfg=dict(ndim=5, seed=77, pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0, 0.25, 0.75])],
                                                obs=[(OBS_X2, [], 1, 1, 1, ESTIM_CORR), (OBS_X2_SUM, [5.0], 1, 1, 1, ESTIM_CORR),
                                                     (OBS_EXP_X2, [5.0], 1, 1, 1, ESTIM_CORR)]),
                  calls=[dict(nmc=4096, find=1, decorr=1)]))
    S.append(dict(name="allmove_simple", cfg=dict(ndim=7, seed=4711, ndecorr=500, pdfs=[(PDF_GAUSS_DIAG, cfg2_a(7))],
                                                  obs=[(OBS_X2, [], 0, 1, 0, ESTIM_NOOP), (OBS_X2_SUM, [7.0], 0, 1, 0, ESTIM_NOOP)]),
                  calls=[dict(nmc=3000, find=1, decorr=1)]))
    # remaining catalogue kinds, periodic domain, mixture target, Simple/Noop, FCBlocker, Uncorrelated
    S.append(dict(name="catalogue_exp", cfg=dict(ndim=4, seed=7, nfind=20, ndecorr=2000, move=(0, 2, 2, [2, 4]),
                                                 pdfs=[(PDF_EXP_DIAG, [1., 2., 0.5, 1.])],
                                                 obs=[(OBS_X, [], 16, 2, 0, ESTIM_UNCORR), (OBS_X_SUM, [], 0, 1, 0, ESTIM_NOOP),
                                                      (OBS_EXP_X2, [], 8, 1, 1, ESTIM_CORR), (OBS_CONST, [1.3], 1, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=4096, find=1, decorr=1)]))
    S.append(dict(name="catalogue_gmm_pbc", cfg=dict(ndim=2, seed=99, irange=([-1., -2.], [1., 3.]), move=(0, 1, 1, None),
                                                     pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.5, 1.2, 0.2, 0.1, -0.5, 1.0])],
                                                     obs=[(OBS_X2, [], 64, 1, 1, ESTIM_MJ), (OBS_INDICATOR, [1], 1, 1, 1, ESTIM_FC)]),
                  calls=[dict(nmc=64 * 128, find=1, decorr=1)]))
    S.append(dict(name="two_pdfs_allmove", cfg=dict(ndim=3, seed=31337, pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0]), (PDF_EXP_DIAG, [])],
                                                    obs=[(OBS_X2, [], 32, 1, 1, ESTIM_CORR)]),
                  calls=[dict(nmc=32 * 256, find=1, decorr=1)]))
    # no sampling function: uniform random sampling in a periodic box with volume scaling
    S.append(dict(name="nopdf_volume", cfg=dict(ndim=3, seed=5649871, irange=([-5.] * 3, [5.] * 3),
                                                obs=[(OBS_X2_SUM, [3.0], 1, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=10000, find=0, decorr=0)]))
    # the last integral of test/ut2/main.cpp: GaussXSquared without sampling function in [-5, 5]^3
    S.append(dict(name="ut2_nopdf_gauss_x2", cfg=dict(ndim=3, seed=5649871, irange=([-5.] * 3, [5.] * 3),
                                                      obs=[(OBS_GAUSS_X2, [], 1, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=10000, find=0, decorr=0)]))
    # every move distribution (SRRDType) with createSymRRD's defaults, as an all-coordinate move
    # and as a single-coordinate move with two step types; the second call continues with the
    # distributions' saved normals
    for srrd, nm in enumerate(SRRD_NAMES):
        if srrd == 0:
            continue
        for veclen, ntypes, te in ((0, 1, None), (1, 2, [1, 3])):
            S.append(dict(name=f"move_{nm}_{'all' if veclen == 0 else 'vec'}",
                          cfg=dict(ndim=3, seed=1000 + srrd, ndecorr=200, move=(srrd, veclen, ntypes, te),
                                   pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0])], obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR)]),
                          calls=[dict(nmc=2048, find=1, decorr=1), dict(nmc=1024, find=0, decorr=0)]))
    # MultiStepMove: the default one (ndim uniform single-coordinate steps, no own sampling function),
    # one with a cheap Gaussian of its own around a Gaussian-mixture target, and one around a
    # Gaussian move of all coordinates with two own sampling functions in a periodic box
    S.append(dict(name="multistep_default", cfg=dict(ndim=3, seed=4711, ndecorr=100, multistep=(3, 0, 1, 1, None, []),
                                                      pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0])], obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=1024, find=1, decorr=1), dict(nmc=512, find=0, decorr=0)]))
    # (a fixed step: the calibration of a multi-step move whose own density rejects every inner
    # step sees acceptance 1 and grows the step without end, in the reference as here)
    S.append(dict(name="multistep_own_pdf", cfg=dict(ndim=2, seed=815, ndecorr=100, step=0.8,
                                                      multistep=(4, 0, 1, 2, [1, 2], [(PDF_GAUSS_DIAG, [0.3, 0.3])]),
                                                      pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.5, 1.2, 0.2, 0.1, -0.5, 1.0])],
                                                      obs=[(OBS_X2, [], 16, 1, 0, ESTIM_UNCORR), (OBS_INDICATOR, [0], 16, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=2048, find=0, decorr=1), dict(nmc=512, find=0, decorr=0)]))
    S.append(dict(name="multistep_gauss_two_pdfs_pbc", cfg=dict(ndim=3, seed=99, ndecorr=100, irange=([-2.] * 3, [2.] * 3),
                                                                 multistep=(2, 1, 0, 1, None, [(PDF_GAUSS_ISO, []), (PDF_EXP_DIAG, [0.5, 0.5, 0.5])]),
                                                                 pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0])], obs=[(OBS_X2_SUM, [3.0], 16, 1, 0, ESTIM_UNCORR)]),
                  calls=[dict(nmc=1024, find=1, decorr=1)]))
    # dependent observables (DependentObservableInterface.hpp): a sampling function's log value and
    # proto values as its observation callback sees them, with different nskip (the callback runs
    # every gcd of them, MCIntegrator.cpp:294-305), and the squares of an earlier observable's last
    # values; automatic decorrelation takes the equilibration observables among them
    S.append(dict(name="depobs_diag_vec", cfg=dict(ndim=4, seed=2718, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0, 0.25])],
                                                   obs=[(OBS_X2, [], 16, 2, 1, ESTIM_MJ), (OBS_PDF_PROTO, [0], 16, 4, 1, ESTIM_MJ),
                                                        (OBS_PDF_LOG, [0], 1, 6, 0, ESTIM_CORR), (OBS_DEP_SQUARE, [0], 8, 4, 1, ESTIM_UNCORR)],
                                                   x0=[2.0, -1.0, 0.5, 3.0]),
                  calls=[dict(nmc=2 ** 15, find=1, decorr=1), dict(nmc=2 ** 14, find=0, decorr=0)]))
    S.append(dict(name="depobs_two_pdfs_all", cfg=dict(ndim=3, seed=31415, pdfs=[(PDF_GAUSS_ISO, []), (PDF_EXP_DIAG, [0.5, 1.0, 1.5])],
                                                       obs=[(OBS_X2_SUM, [], 32, 1, 1, ESTIM_MJ), (OBS_PDF_LOG, [1], 32, 2, 1, ESTIM_UNCORR),
                                                            (OBS_PDF_PROTO, [1], 0, 3, 0, ESTIM_NOOP), (OBS_PDF_LOG, [0], 32, 4, 1, ESTIM_MJ),
                                                            (OBS_DEP_SQUARE, [0], 32, 4, 1, ESTIM_UNCORR)]),
                  calls=[dict(nmc=4096, find=1, decorr=1), dict(nmc=8192, find=0, decorr=0)]))
    S.append(dict(name="depobs_gmm_pbc", cfg=dict(ndim=2, seed=99, irange=([-1., -2.], [1., 3.]), move=(0, 1, 1, None),
                                                  pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.5, 1.2, 0.2, 0.1, -0.5, 1.0])],
                                                  obs=[(OBS_PDF_LOG, [0], 64, 1, 1, ESTIM_MJ), (OBS_DEP_SQUARE, [0], 64, 2, 1, ESTIM_UNCORR)]),
                  calls=[dict(nmc=64 * 128, find=1, decorr=1)]))
    return S


SRRD_NAMES = ["uniform", "gaussian", "student", "cauchy", "exponential", "gamma", "weibull", "lognormal", "chisq", "fisher"]


def run_scenario(s):
    L = script_setup(s["cfg"])
    for call in s["calls"]:
        if "irange" in call:
            lb, ub = call["irange"]
            L.append("irange " + " ".join(fmt(v) for v in lb) + " " + " ".join(fmt(v) for v in ub))
        if "setx" in call:
            L.append("setx " + " ".join(fmt(v) for v in call["setx"]))
        L.append(f"integrate {call['nmc']} {call['find']} {call['decorr']}")
    recs = ref_run(L)
    for r in recs:
        r.pop("seconds", None)
    return recs


def device_functor_scenarios():
    """Chains of plugins outside the library's catalogue, written as a reference user writes them
    (oracle/ref_harness.cpp QuarticPDF: p ~ exp(-sum c_i x_i^4) without an updatedAcceptance of its
    own; MomentsObs: sum |x_i|^3 and max |x_i|, not updateable).  The library runs them as device
    functors compiled from source (tests/device_functors.py)."""
    from device_functors import MOVE_USER_PAIR, MOVE_USER_SCALE, OBS_USER_MOMENTS, PDF_USER_QUARTIC, PDF_USER_QUARTIC_UPD
    S = []
    S.append(dict(name="quartic_allmove", cfg=dict(ndim=4, seed=2024, ndecorr=500, pdfs=[(PDF_USER_QUARTIC, [0.5, 1.0, 2.0, 0.25])],
                                                   obs=[(OBS_USER_MOMENTS, [], 16, 1, 1, ESTIM_MJ), (OBS_X2, [], 16, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=16 * 64, find=1, decorr=1), dict(nmc=16 * 32, find=0, decorr=0)]))
    S.append(dict(name="quartic_vecmove_full", cfg=dict(ndim=3, seed=77, ndecorr=300, move=(0, 1, 1, None),
                                                        pdfs=[(PDF_USER_QUARTIC, [1.0, 1.0, 0.5])],
                                                        obs=[(OBS_USER_MOMENTS, [], 1, 1, 1, ESTIM_CORR)]),
                  calls=[dict(nmc=1024, find=1, decorr=1)]))
    S.append(dict(name="quartic_auto_decorrelation", cfg=dict(ndim=2, seed=5649871, pdfs=[(PDF_USER_QUARTIC, [1.0, 2.0])],
                                                              obs=[(OBS_USER_MOMENTS, [], 8, 1, 1, ESTIM_MJ)], x0=[2.0, -1.5]),
                  calls=[dict(nmc=8 * 128, find=1, decorr=1)]))
    # with updatedAcceptance (QuarticUpdPDF): single-coordinate and two-coordinate moves update the
    # changed proto values only
    for veclen in (1, 2):
        S.append(dict(name=f"quartic_updated_vec{veclen}", cfg=dict(ndim=4, seed=31 + veclen, ndecorr=400, move=(0, veclen, 1, None),
                                                                   pdfs=[(PDF_USER_QUARTIC_UPD, [0.5, 1.0, 2.0, 0.25])],
                                                                   obs=[(OBS_USER_MOMENTS, [], 16, 1, 1, ESTIM_MJ)]),
                      calls=[dict(nmc=16 * 64, find=1, decorr=1)]))
    # trial moves outside the SRRD family (PairMove: two step types, a two-coordinate range, the
    # catalogue's diagonal Gaussian with its single-index update; ScaleMove: the move's own proto
    # value and factor, on the quartic sampling function)
    S.append(dict(name="pairmove_diag", cfg=dict(ndim=4, seed=606, ndecorr=300, umove=(MOVE_USER_PAIR, [0.3, 0.6]),
                                                 pdfs=[(PDF_GAUSS_DIAG, [0.5, 1.0, 2.0, 0.25])],
                                                 obs=[(OBS_X2, [], 16, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 16, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=16 * 64, find=1, decorr=1), dict(nmc=16 * 16, find=0, decorr=0)]))
    S.append(dict(name="scalemove_quartic", cfg=dict(ndim=3, seed=808, ndecorr=200, umove=(MOVE_USER_SCALE, [0.2]),
                                                     pdfs=[(PDF_USER_QUARTIC, [1.0, 0.5, 2.0])], x0=[0.5, -0.4, 0.3],
                                                     obs=[(OBS_USER_MOMENTS, [], 16, 1, 1, ESTIM_MJ)]),
                  calls=[dict(nmc=16 * 64, find=1, decorr=1)]))
    return [dict(name=s["name"], cfg=s["cfg"], calls=s["calls"], results=run_scenario(s)) for s in S]


def file_scenarios():
    """The observable and walker files of the main sampling (MCIntegrator.cpp:495-542) as the
    reference writes them (calibration and decorrelation on: they write nothing)."""
    import tempfile
    F = []
    for name, cfg, nmc, fo, fw in [
        ("files_gauss3", dict(ndim=3, seed=1337, pdfs=[(PDF_GAUSS_ISO, [])],
                              obs=[(OBS_X2, [], 1, 1, 0, ESTIM_UNCORR), (OBS_X2_SUM, [3.0], 4, 2, 0, ESTIM_UNCORR)]), 400, 7, 5),
        ("files_vec_gmm", dict(ndim=2, seed=99, ndecorr=100, move=(0, 1, 1, None),
                               pdfs=[(PDF_GMM_ISO, [2, 0.3, 0.7, 0.5, 1.2, 0.2, 0.1, -0.5, 1.0])],
                               obs=[(OBS_X, [], 1, 3, 0, ESTIM_UNCORR), (OBS_EXP_X2, [], 1, 1, 0, ESTIM_UNCORR)]), 300, 1, 3),
    ]:
        with tempfile.TemporaryDirectory() as d:
            L = script_setup(cfg) + [f"obsfile {d}/obs.txt {fo}", f"wlkfile {d}/wlk.txt {fw}", f"integrate {nmc} 1 1"]
            ref_run(L)
            with open(f"{d}/obs.txt") as fobs, open(f"{d}/wlk.txt") as fwlk:
                F.append(dict(name=name, cfg=cfg, nmc=nmc, freq_obs=fo, freq_wlk=fw, obs=fobs.read(), wlk=fwlk.read()))
    return F


def trace_scenarios():
    """Step-by-step traces of the main sampling (no calibration, no decorrelation)."""
    T = []
    for name, cfg, nmc in [
        ("trace_all3", dict(ndim=3, seed=1337, pdfs=[(PDF_GAUSS_ISO, [])], obs=[(OBS_X2, [], 1, 1, 0, ESTIM_UNCORR)],
                            x0=[0.3, -0.2, 0.1], step=1.1), 256),
        ("trace_vec4", dict(ndim=4, seed=2024, pdfs=[(PDF_GAUSS_DIAG, [0.5, 1., 2., 4.])], move=(0, 1, 1, None),
                            obs=[(OBS_X2, [], 1, 1, 0, ESTIM_UNCORR)], x0=[1., -1., 0.5, 0.], step=0.9), 256),
        ("trace_types5", dict(ndim=6, seed=77, pdfs=[(PDF_EXP_DIAG, [])], move=(0, 2, 3, [2, 4, 6]),
                              obs=[(OBS_X, [], 1, 1, 0, ESTIM_UNCORR)], step=0.7), 256),
    ]:
        L = script_setup(cfg) + ["trace", f"integrate {nmc} 0 0"]
        r = ref_run(L)[-1]
        r.pop("seconds", None)
        T.append(dict(name=name, cfg=cfg, nmc=nmc, result=r))
    return T


def phase_trace_scenarios():
    """What the callback sees through calibration, decorrelation and sampling (MCIntegrator.cpp:265,
    346-347, 373): fixed and automatic calibration and decorrelation, and random sampling without a
    sampling function."""
    T = []
    for name, cfg, nmc, find, decorr in [
        ("phases_fixed", dict(ndim=3, seed=4242, nfind=3, ndecorr=50, pdfs=[(PDF_GAUSS_ISO, [])], obs=[(OBS_X2, [], 1, 1, 0, ESTIM_UNCORR)],
                              x0=[1.0, -0.5, 0.25]), 64, 1, 1),
        ("phases_auto", dict(ndim=2, seed=99, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, [0.5, 2.0])],
                             obs=[(OBS_X2, [], 1, 1, 1, ESTIM_CORR)], x0=[3.0, -2.0]), 64, 1, 1),
        ("phases_nopdf", dict(ndim=2, seed=5, irange=([-1.0, -1.0], [2.0, 1.0]), obs=[(OBS_X2_SUM, [], 1, 1, 0, ESTIM_UNCORR)]), 32, 1, 1),
    ]:
        L = script_setup(cfg) + ["trace", f"integrate {nmc} {find} {decorr}"]
        r = ref_run(L)[-1]
        r.pop("seconds", None)
        T.append(dict(name=name, cfg=cfg, nmc=nmc, find=find, decorr=decorr, result=r))
    return T


def walker_scenarios():
    """Several independent reference MCIs with their own seeds (one per GPU walker)."""
    seeds = [1337, 11, 12345, 987654321, 2 ** 40 + 3, 5649871, 42, 7]
    cfg = dict(ndim=16, ndecorr=1000, move=(0, 1, 1, None), pdfs=[(PDF_GAUSS_DIAG, cfg2_a(16))],
               obs=[(OBS_X2, [], 160, 1, 1, ESTIM_MJ), (OBS_X2_SUM, [], 160, 1, 1, ESTIM_MJ)])
    out = []
    for s in seeds:
        c = dict(cfg, seed=s)
        r = run_scenario(dict(cfg=c, calls=[dict(nmc=10240, find=1, decorr=1)]))[-1]
        out.append(r)
    return dict(cfg=cfg, seeds=seeds, nmc=10240, results=out)


def estimator_sets():
    rng = np.random.default_rng(20240607)
    sets = []
    for n, nd, kind in [(2, 1, "iid"), (4, 2, "iid"), (64, 1, "ar"), (64, 17, "ar"), (128, 3, "ar"), (1000, 2, "ar"),
                        (1024, 1, "ar"), (2048, 3, "ar"), (50, 1, "iid"), (51, 3, "ar"), (16, 3, "iid")]:
        if kind == "iid":
            x = rng.normal(size=(n, nd))
        else:
            x = np.zeros((n, nd))
            e = rng.normal(size=(n, nd))
            for i in range(1, n):
                x[i] = 0.9 * x[i - 1] + e[i]
            x += 2.0
        ests = [ESTIM_UNCORR, ESTIM_NOOP]
        if n >= 50:
            ests.append(ESTIM_FC)
        if n & (n - 1) == 0:
            ests.append(ESTIM_MJ)
        ests.append(ESTIM_CORR if (n >= 50 or n & (n - 1) == 0) else ESTIM_UNCORR)
        res = {}
        for est in ests:
            L = [f"estimate {est} {n} {nd} " + " ".join(fmt(v) for v in x.ravel())]
            r = ref_run(L)[-1]
            res[str(est)] = dict(avg=r["avg"], err=r["err"])
        sets.append(dict(n=n, ndim=nd, x_b64=b64(x.ravel()), results=res))
    return sets


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -C oracle")
    meta = dict(flags=FLAGS, generator="tests/golden/make_golden.py")
    chains = [dict(name=s["name"], cfg=s["cfg"], calls=s["calls"], results=run_scenario(s)) for s in scenarios()]
    with open(os.path.join(HERE, "chains.json"), "w") as f:
        json.dump(dict(meta=meta, scenarios=chains), f, indent=0)
    with open(os.path.join(HERE, "traces.json"), "w") as f:
        json.dump(dict(meta=meta, traces=trace_scenarios(), phase_traces=phase_trace_scenarios()), f, indent=0)
    with open(os.path.join(HERE, "device_functors.json"), "w") as f:
        json.dump(dict(meta=meta, scenarios=device_functor_scenarios()), f, indent=0)
    with open(os.path.join(HERE, "files.json"), "w") as f:
        json.dump(dict(meta=meta, files=file_scenarios()), f, indent=0)
    with open(os.path.join(HERE, "walkers.json"), "w") as f:
        json.dump(dict(meta=meta, **walker_scenarios()), f, indent=0)
    with open(os.path.join(HERE, "estimators.json"), "w") as f:
        json.dump(dict(meta=meta, sets=estimator_sets()), f, indent=0)
    walk = ref_run(["walkdata 1337 4096 2 2.0 0.5"])[-1]
    walk["x_b64"] = b64([float.fromhex(v) for v in walk.pop("x")])
    for k in [k for k in walk if k.startswith("accu_")]:
        walk[k + "_b64"] = b64([float.fromhex(v) for v in walk.pop(k)])
    with open(os.path.join(HERE, "walkdata.json"), "w") as f:
        json.dump(dict(meta=dict(meta, args="seed 1337, nmc 4096, ndim 2, step 2.0, changeProb 0.5"), walk=walk), f)
    with open(os.path.join(HERE, "sigmas.json"), "w") as f:
        json.dump(dict(meta=dict(meta, draw="std::uniform_real_distribution<double>(0.5,2) on std::mt19937_64(1337)"),
                       sigma16=[fmt(v) for v in sigmas(16)], sigma64=[fmt(v) for v in sigmas(64)]), f, indent=0)
    for fn in sorted(os.listdir(HERE)):
        print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    main()
