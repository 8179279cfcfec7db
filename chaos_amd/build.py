"""Build libchaoslab_hip.so (gfx950) in-tree with hipcc.

Each .hip source compiles to an object in parallel, then one shared library is linked next to
this file. Every HIP compile uses -ffp-contract=off (the kernels' sin/cos/log and reducers are
written with explicit fma where fusion is intended, so host and device round identically) and
disables packed FP32 instructions.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchaoslab_hip.so")
SOURCES = ["chain_mixed.hip", "chain_mixed_f32.hip", "chain_n128.hip", "chain_n128f.hip", "chain_n64.hip", "chain_n64f.hip",
           "chain_n32.hip", "chain_n32f.hip", "chain_n16.hip", "chain_n16f.hip", "chain_n8.hip", "chain_n8f.hip", "chain_kernels.hip",
           "propagate.hip", "inst_maps.hip", "inst_flow.hip", "inst_f32.hip", "inst_pipe.hip", "record.hip", "shadow.hip", "mlp.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CL_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-result", "-Wno-unused-value", "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(int(os.environ.get('MAX_JOBS', '8')), len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
