"""Builds the a5g HIP library for gfx950 in-tree (a5_streamer_amd/liba5g.so),
then the Fortran binding and the 2D streamer driver on it
(a5_streamer_amd/fortran, with flang).

    python -m a5_streamer_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liba5g.so")

SOURCES = ["a5g_tree.cpp", "a5g_api.hip", "a5g_mg.hip", "a5g_init.hip", "a5g_streamer.hip",
           "a5g_diag.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # FP64 parity with the reference: no fused multiply-add contraction
    "-ffp-contract=off",
    # packed FP32 instructions are kept out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", os.path.join(HERE, "..", "include"),
                                   "-o", tmp] + srcs
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
    build(verbose=True)
    sys.exit(0)
