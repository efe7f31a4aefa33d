"""Builds the a5g HIP library for gfx950 in-tree (a5_streamer_amd/liba5g.so),
then the Fortran binding and the 2D streamer driver on it
(a5_streamer_amd/fortran, with flang).

Each source is compiled to an object of its own under
a5_streamer_amd/_build, several at a time, and the objects are linked into
the library.

    python -m a5_streamer_amd.build [--incremental]

--incremental recompiles only the sources older than their object (or than
any header or this file); the default compiles everything.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liba5g.so")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(HERE, "..", "include")

SOURCES = ["a5g_tree.cpp", "a5g_api.hip", "a5g_mg.hip", "a5g_init.hip", "a5g_streamer.hip",
           "a5g_diag.hip", "a5g_adjoint.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # FP64 parity with the reference: no fused multiply-add contraction
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _deps_mtime():
    """Newest modification time of the headers and of this file."""
    paths = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    paths += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    paths.append(os.path.abspath(__file__))
    return max(os.path.getmtime(p) for p in paths)


def _compile(hipcc, src, obj, verbose):
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", INCLUDE, "-c", "-o", obj + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(obj + ".tmp", obj)


def build(verbose=False, incremental=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJ, exist_ok=True)
    deps = _deps_mtime()
    todo, objs = [], []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        stale = (not os.path.exists(obj) or os.path.getmtime(obj) < os.path.getmtime(src)
                 or os.path.getmtime(obj) < deps)
        if stale or not incremental:
            todo.append((src, obj))
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2) - 1), 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_compile, hipcc, s, o, verbose) for s, o in todo]:
            f.result()
    tmp = OUT + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, OUT)
    fdir = os.path.join(HERE, "fortran")
    fc = os.environ.get("FC", "flang")
    # the driver binary links liba5g.so: rebuild it against the new library
    subprocess.check_call(["make", "-s", "-B", "-C", fdir, "FC=" + fc])
    return OUT


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
