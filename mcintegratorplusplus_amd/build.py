"""Builds the HIP library in-tree: mcintegratorplusplus_amd/lib/libmci_amd.so (gfx950).

    python -m mcintegratorplusplus_amd.build

Every HIP source is compiled by hipcc for gfx950 only, with FP contraction off (the FP64
chain is bit-exact with the reference; fused operations are written explicitly where
wanted), IEEE correctly rounded FP32 divide and sqrt (the batched FP32 chain is
reproduced bit for bit by the CPU oracle) and packed-FP32 instructions disabled.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
LIB = os.path.join(OUT, "libmci_amd.so")
SOURCES = ["mci_kernels.hip", "mci_fast.hip", "mci_batch.hip", "mci_mlp.hip", "mci_lp.hip", "mci_noalloc.hip", "mci_em.hip", "mci_api.cpp",
           "mci_batch_api.cpp", "mci_lp_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCI_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument"]
# the learned move keeps the last layer's 128 outputs per lane across phases: matrix-core results in
# ordinary VGPRs, so that they are not held twice (accumulation registers and VGPRs)
EXTRA = {"mci_lp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
