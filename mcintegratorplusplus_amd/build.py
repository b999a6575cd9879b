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
           "mci_batch_api.cpp", "mci_lp_api.cpp", "mci_rtc.cpp"]
ROOT = os.path.dirname(HERE)
# headers the run-time build of the step kernels with the caller's device functors needs, under the
# names their #include lines use (mci_rtc.cpp); the last two stand in for system headers that the
# run-time compiler does not carry (its device runtime is built in; FLT_MIN / FLT_MAX from the
# compiler's predefined macros)
RTC_HEADERS = [("mci_step_kernels.hpp", os.path.join(CSRC, "mci_step_kernels.hpp")),
               ("mci_kernels.hpp", os.path.join(CSRC, "mci_kernels.hpp")),
               ("mci_device.hpp", os.path.join(CSRC, "mci_device.hpp")),
               ("mci_f32math.hpp", os.path.join(CSRC, "mci_f32math.hpp")),
               ("../../include/mci_amd.h", os.path.join(ROOT, "include", "mci_amd.h")),
               ("../../include/mci_amd/device_functor.hpp", os.path.join(ROOT, "include", "mci_amd", "device_functor.hpp")),
               ("hip/hip_runtime.h", None),
               ("float.h", "#define FLT_MIN __FLT_MIN__\n#define FLT_MAX __FLT_MAX__\n")]


def _embed():
    """mci_rtc_embed.cpp: the RTC_HEADERS as byte arrays (names and texts for hiprtcCreateProgram)"""
    out = ["// written by build.py: headers of the run-time build of the step kernels", "namespace mcid", "{"]
    for i, (name, src) in enumerate(RTC_HEADERS):
        if src is None:
            text = b"\n"
        elif src.startswith("#define"):
            text = src.encode()
        else:
            with open(src, "rb") as f:
                text = f.read()
        out.append(f"static const char h{i}[] = {{" + ",".join(str(b if b < 128 else b - 256) for b in text) + ",0};")
    out.append(f"extern const int rtc_nheaders = {len(RTC_HEADERS)};")
    out.append("extern const char *const rtc_header_names[] = {" + ",".join(f'"{n}"' for n, _ in RTC_HEADERS) + "};")
    out.append("extern const char *const rtc_header_texts[] = {" + ",".join(f"h{i}" for i in range(len(RTC_HEADERS))) + "};")
    out.append("} // namespace mcid")
    path = os.path.join(OUT, "mci_rtc_embed.cpp")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path
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
    obj = os.path.join(OUT, os.path.splitext(os.path.basename(src))[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + FLAGS + EXTRA.get(src, []) + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    srcs = SOURCES + [_embed()]
    with cf.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
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
