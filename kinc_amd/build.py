"""Build the MI355X Similarity library (kinc_amd/libkinc.so) for gfx950.

The HIP translation units are compiled in parallel with hipcc and linked
into one shared library that exports the C-ABI of include/kinc_sim.h.
Flags that the bit-exact contract depends on:
  -ffp-contract=off                      no FMA contraction (the reference's
                                          x86-64 SSE2 arithmetic has none)
  -fno-gpu-flush-denormals-to-zero       FP32 denormals preserved
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and sqrt
  -target-feature -packed-fp32-ops       scalar FP32 only (no v_pk_*_f32)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libkinc.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("KINC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = [
    "sim_pairs_k5.hip",
    "sim_pairs_k8.hip",
    "sim_util_kernels.hip",
    "kinc_sim.cpp",
    "formats.cpp",
]
CLI_SOURCE = "kinc_cli.cpp"
CLI = os.path.join(HERE, "kinc")


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    cli_obj = _compile(CLI_SOURCE)
    tmp = CLI + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-o", tmp, cli_obj, "-L" + HERE, "-lkinc",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("CLI link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, CLI)
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
