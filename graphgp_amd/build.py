"""Build the gfx950 HIP library ``graphgp_amd/libgraphgp_amd.so`` in place.

Every ``csrc/*.hip`` file is compiled with ``hipcc --offload-arch=gfx950`` (in parallel)
and linked into one shared library that exports the C ABI of ``include/graphgp_amd.h``.
Packed FP32 instructions are disabled for all device code; the tree build and the kNN
search are compiled without floating-point contraction so that their comparisons use the
same separately rounded arithmetic as the reference.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgraphgp_amd.so")
ARCH = os.environ.get("GGP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
NO_CONTRACT = {"tree.hip", "knn.hip"}


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return h


def _compile(src: str) -> str:
    name = os.path.basename(src)
    obj = os.path.join(OUT_DIR, name + ".o")
    flags = list(COMMON)
    if name in NO_CONTRACT:
        flags.append("-ffp-contract=off")
    cmd = [_hipcc(), *flags, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, jobs: int | None = None, only=None) -> str:
    """Compile every source (or, with ``only``, just those file names, reusing the other
    objects of an earlier build) and link the library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    todo = srcs if not only else [s for s in srcs if os.path.basename(s) in set(only)]
    jobs = jobs or min(max(1, len(todo)), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OUT_DIR, os.path.basename(s) + ".o") for s in srcs]
    missing = [o for o in objs if not os.path.exists(o)]
    if missing:
        raise RuntimeError(f"objects missing (run a full build): {missing}")
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, only=sys.argv[1:] or None)
    sys.exit(0)
