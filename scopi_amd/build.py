"""Builds libscopi_hip.so (gfx950) in-tree with hipcc.

Usage: python -m scopi_amd.build   (or scopi_amd.build.build() from __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libscopi_hip.so")
SOURCES = ["api.hip", "contacts.hip", "apgd.hip", "sort_scan.hip", "superellipsoid.hip"]

# Packed FP32 instructions are disabled for every device pass.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc, *HIPCC_FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libscopi_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_host_tests() -> str:
    """C++ program that drives the slot API mirror (scopi_amd/host) like the reference's tests."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "cpp", "test_slots.cpp")
    out = os.path.join(root, "tests", "cpp", "test_slots")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(HERE, "host"), src, "-L", HERE, "-lscopi_hip",
           "-Wl,-rpath,$ORIGIN/../../scopi_amd", "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_slots")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
