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

# waves per SIMD the pair kernels are register-budgeted for: 4 (128 VGPRs)
# spills (a few reloads per EM iteration) and runs the GMM/Spearman workload
# 1.45x faster than the spill-free 2
PAIR_WAVES = 4

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = [
    "sim_util_kernels.hip",
    "kinc_sim.cpp",
    "formats.cpp",
    "server.cpp",
    "rmt.hip",
]
# pair-kernel instances (KMAX, P): register-slot counts P, and P = 0 for the
# memory-store instance used above the register range (1024 samples for up to
# 5 clusters, 512 for up to 8)
INSTANCES = ([(5, p) for p in (0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16)]
             + [(8, p) for p in (0, 2, 4, 8)] + [(64, 0)])
# the instances of at most 6 slots (N <= 384) are built a second time for 5
# waves per SIMD (96 VGPRs); the library uses those when the per-wave LDS
# leaves room for 20 waves per CU (about 400 samples or fewer): 1.11 M
# against 1.03 M pairs/s at 256 samples, 726k against 696k at 384, while at
# 500 samples, where LDS allows 16, mixing them into the concurrent bins
# cost time
INSTANCES_W5 = [(5, p) for p in (1, 2, 3, 4, 5, 6)] + [(8, p) for p in (2, 4)]
# the instances of 5 slots or more are built a third time for 3 waves per
# SIMD (168 VGPRs); a handle uses those when its per-wave LDS allows at most
# 12 waves per CU anyway (about 640 samples or more with outlier removal or
# Spearman, about 800 without), where the 4-wave builds would spill for no
# extra residency
INSTANCES_WIDE = [(5, p) for p in (5, 6, 7, 8, 10, 12, 14, 16)] + [(8, 4), (8, 8)]
# the phase-census build (per-phase cycle and event counters, profiling only)
# of the instance that carries the 20,000 x 500 GMM workload
CENSUS_INSTANCES = [(5, 7)]
CLI_SOURCE = "kinc_cli.cpp"
CLI = os.path.join(HERE, "kinc")


def _compile(src, defines=(), tag=None):
    out = os.path.join(OBJ, (tag or os.path.splitext(src)[0]) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def _inst_job(k, p, waves, census=False):
    """Compile job of one pair-kernel instance, with its rough relative cost
    (register slots x clusters; the memory-store instance counts as 4 slots)."""
    defines = ["KINC_INST_KMAX=%d" % k, "KINC_INST_P=%d" % p, "KINC_PAIR_WAVES=%d" % waves]
    tag = "pairs_k%d_p%d" % (k, p)
    if census:
        defines.append("KINC_CENSUS=1")
        tag = "pairs_census_k%d_p%d" % (k, p)
    elif waves != PAIR_WAVES:
        tag += "_w%d" % waves
    return (k * (p or 4), ("sim_pairs_inst.hip", tuple(defines), tag))


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    inst = [_inst_job(k, p, PAIR_WAVES) for k, p in INSTANCES]
    inst += [_inst_job(k, p, 5) for k, p in INSTANCES_W5]
    inst += [_inst_job(k, p, 3) for k, p in INSTANCES_WIDE]
    inst += [_inst_job(k, p, PAIR_WAVES, census=True) for k, p in CENSUS_INSTANCES]
    # longest first, so the pool does not end on one large instance
    jobs = [j for _, j in sorted(inst, key=lambda cj: -cj[0])]
    jobs += [(src, ("KINC_PAIR_WAVES=%d" % PAIR_WAVES,), None) for src in SOURCES]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
        "-Wl,-rpath," + os.path.join(ROCM, "lib")]
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
