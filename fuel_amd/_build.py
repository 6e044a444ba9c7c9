"""Build libfuel_amd.so in-tree with hipcc for gfx950.

Every HIP/C++ source under fuel_amd/csrc is compiled to an object with the
flags below and linked into fuel_amd/libfuel_amd.so.  Numerics-relevant flags:
  -ffp-contract=off        no a*b+c contraction: every fused multiply-add in the
                           restated distance is written explicitly (fuel's Rust
                           code never contracts);
  -fhip-fp32-correctly-rounded-divide-sqrt
                           IEEE sqrt/divide, as fuel's f32::sqrt;
  -packed-fp32-ops off     no packed FP32 VALU instructions.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfuel_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("FUEL_AMD_ARCH", "gfx950")

SOURCES = [
    "api.cpp",
    "engine.cpp",
    "engine_mfma.cpp",
    "cluster.cpp",
    "serve.cpp",
    "batch.cpp",
    "priority.cpp",
    "noalloc.hip",
    "resid.hip",
    "scan_exact.hip",
    "scan_mfma.hip",
    "finalize_rerank.hip",
    "prep.hip",
    "range_pairwise.hip",
    "kmeans_pairwise.hip",
    "lof.hip",
    "sort.hip",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags():
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-fno-fast-math",
        "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-Wno-unknown-warning-option",
        f"-I{os.path.join(HERE, '..', 'include')}",
        f"-I{CSRC}",
    ]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD_DIR, src + ".o")
    cmd = [hipcc()] + _flags() + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True, jobs=None):
    """Compile every source (always) and link the shared library."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
