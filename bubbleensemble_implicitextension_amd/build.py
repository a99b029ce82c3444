"""Builds the gfx950 solver library (libbex.so) in-tree with hipcc.

    python -m bubbleensemble_implicitextension_amd.build [--verbose]

The library is position independent, built for gfx950 only, with -ffp-contract=off so that the
kernel's arithmetic is exactly the order written in the source (the CPU oracle uses the same
order), with packed FP32 instructions disabled, without the atomic optimizer and with loop heads
aligned to 64 bytes.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
REPO = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbex.so")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # The kernels' LDS counters are updated by one lane at a time: wave-level aggregation of
    # those atomics (mbcnt/bcnt/exec juggling around every add) only adds instructions.
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    # Loop heads on 64-byte boundaries: a system's wave re-enters the BiCG loop about 10^6 times
    # per period, and where its head falls in the instruction cache line shows up as ~3 % of
    # that loop's time (measured on the tail system of the N=32 workload).
    "-falign-loops=64",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I" + os.path.join(REPO, "include"),
    "-I" + CSRC,
]


def build_library(verbose=False, extra_flags=(), output=LIB):
    sources = [os.path.join(CSRC, "bex_capi.hip"), os.path.join(CSRC, "bex_mlp.hip")]
    cmd = [HIPCC, *COMMON_FLAGS, *extra_flags, "-shared", "-Wl,--no-undefined", "-o",
           output + ".tmp", *sources]
    if verbose:
        cmd.insert(1, "-Rpass-analysis=kernel-resource-usage")
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stdout.write(res.stdout)
        sys.stderr.write(res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(output + ".tmp", output)
    return output


def build_apps(verbose=False):
    """C++ applications written against include/CoupledSystems_PerBlock_Interface.hpp."""
    app = os.path.join(HERE, "collapse_strength")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(REPO, "include"),
           os.path.join(HERE, "apps", "collapse_strength.cpp"), "-L" + HERE, "-lbex",
           "-Wl,-rpath,$ORIGIN", "-o", app]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stdout.write(res.stdout)
        sys.stderr.write(res.stderr)
    if res.returncode != 0:
        raise RuntimeError("building collapse_strength failed")
    return app


def build_all(verbose=False):
    build_library(verbose=verbose)
    build_apps(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose="--verbose" in sys.argv)
