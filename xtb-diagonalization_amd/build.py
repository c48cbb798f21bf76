"""Build recipe for libxtbeig.so (gfx950): every csrc/*.hip through hipcc, linked into one shared library in-tree;
then libxtbeig_lapack.so, the opt-in library of LAPACK Fortran names forwarding to it (gcc).

Usage: python xtb-diagonalization_amd/build.py [--jobs N] [--verbose]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxtbeig.so")
INTERPOSE_LIB = os.path.join(HERE, "libxtbeig_lapack.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions off: hipcc must not form v_pk_*_f32 from scalar float code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {os.path.basename(src)}:\n{r.stderr}")
    return obj, r.stderr


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)), 8)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    objs = [o for o, _ in results]
    for _, err in results:
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_interpose()
    return LIB


def build_interpose():
    """libxtbeig_lapack.so: the opt-in LAPACK-name forwards (interpose/lapack_names.c), linked against libxtbeig.so
    and finding it beside itself at run time."""
    src = os.path.join(HERE, "interpose", "lapack_names.c")
    tmp = INTERPOSE_LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=c99", "-Wall", "-Werror", "-shared", "-fPIC", src, "-o", tmp,
           "-L" + HERE, "-lxtbeig", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"interposing library failed to build:\n{r.stderr}")
    os.replace(tmp, INTERPOSE_LIB)
    return INTERPOSE_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
