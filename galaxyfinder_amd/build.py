"""Build recipe of the HIP library (gfx950) -- explicit hipcc invocations,
in-tree output ``galaxyfinder_amd/libpgalf_density.so``.

Every source is compiled with
  -ffp-contract=off                      no FMA contraction: the TSC weights,
                                         read-back sums, distances and W4
                                         follow the reference's rounding;
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE float division and sqrt;
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 ops.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpgalf_density.so")
SOURCES = ["scan_sort.hip", "tsc.hip", "smooth.hip", "knn.hip", "peaks.hip", "capi.hip",
           "batch.hip", "catalog.hip", "watershed.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PGALF_ARCH", "gfx950")

CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(os.path.dirname(HERE), "include"),
]


def _compile(src, out_dir, extra):
    obj = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + CFLAGS + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr))
    return obj


def build(verbose=False, extra_flags=(), lib=LIB):
    """Compile every source and link `lib`.  `extra_flags` go to every hipcc
    compile (e.g. -DPGALF_KNN_BUFCAP=32 for the kNN capacities).  Objects go
    to _build/ beside the package library, or to _obj/ beside any other
    `lib` (variant builds stay out of the package)."""
    if os.path.abspath(lib) == LIB:
        out_dir = os.path.join(HERE, "_build")
    else:
        out_dir = os.path.join(os.path.dirname(os.path.abspath(lib)), "_obj")
    os.makedirs(out_dir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, out_dir, extra_flags), SOURCES))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
