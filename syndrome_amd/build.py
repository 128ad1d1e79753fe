"""Build the HIP library ``libsyndrome_amd.so`` in-tree for gfx950.

Every call recompiles.  Objects are compiled in parallel, then linked with
hipcc.  Flags that matter for exactness: ``-ffp-contract=off`` (no FMA
formation, so each f32 add/sub is the reference's single rounding) and no
denormal flushing.  Packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsyndrome_amd.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["ldpc_fast_f32_bg1.hip", "ldpc_fastx_f32_bg1.hip", "ldpc_fast_i8_bg1.hip", "ldpc_fast_f32_bg2.hip",
           "ldpc_fast_wifi.hip", "ldpc_pack_i8_bg2.hip",
           "ldpc_persist_f32.hip", "ldpc_persist_f32_small.hip", "ldpc_persist_f32_small_bg2.hip", "ldpc_persist_i8.hip", "ldpc_resident.hip", "ldpc_tb.hip", "ldpc_encode.hip", "ldpc_dematch.hip", "ldpc_plain.hip", "ldpc_wide.hip", "ldpc_plan.hip", "ldpc_serve.hip", "ldpc_capi.hip",
           "code_host.cpp"]
ARCH = os.environ.get("SYNDROME_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]
# per-source extras, scheduling for ILP: the pair-packed kernels, whose
# dependent packed instructions need a wait state between them (max-ilp
# interleaves independent work there instead of s_nop), and the BG1 f32
# kernels (BG1 Z=384: 2.94 -> 2.91 ms per launch, no added spills)
EXTRA = {"ldpc_pack_i8_bg2.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"],
         "ldpc_fast_f32_bg1.hip": ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]}
DEVICE = [
    f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    env = os.environ.get("HIPCC")
    if env:
        return env
    return "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else "hipcc"


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    extra = (DEVICE + EXTRA.get(src, [])) if src.endswith(".hip") else []
    cmd = [hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{errs}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
