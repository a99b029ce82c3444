"""Builds the gfx950 solver library (libbex.so) in-tree with hipcc.

    python -m bubbleensemble_implicitextension_amd.build [--verbose]

The library is position independent, built for gfx950 only, with -ffp-contract=off so that the
kernel's arithmetic is exactly the order written in the source (the CPU oracle uses the same
order), and with packed FP32 instructions disabled.
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
