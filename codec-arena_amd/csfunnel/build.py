"""Build libcsfunnel.so (HIP, gfx950) in-tree next to this file.

    python codec-arena_amd/csfunnel/build.py

Strict IEEE f32 on the device and the host: the reference screen is compiled
by Zig in its default strict float mode, which never contracts a*b+c into a
fused multiply-add, so every translation unit is built with -ffp-contract=off.
Packed-FP32 instruction selection is disabled for the device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
REPO = os.path.dirname(PKG_ROOT)
LIB = os.path.join(HERE, "libcsfunnel.so")
CODEC_LIB = os.path.join(HERE, "libcscodec.so")

SOURCES = ["screen_device.hip", "ridge_device.hip", "mlp_device.hip", "rate_probe.hip", "funnel_host.cpp"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
        "-I", os.path.join(REPO, "include"),
        "-o", LIB + ".tmp",
    ] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed (rc={r.returncode})")
    os.replace(LIB + ".tmp", LIB)
    build_codec(verbose)
    return LIB


def build_codec(verbose: bool = False) -> str:
    """The codec label source is plain host C: built with gcc for a baseline
    x86-64-v3 target (it runs on whichever host the package is shipped to)."""
    cmd = [os.environ.get("CC", "gcc"), "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
           "-fPIC", "-shared", "-std=gnu11", "-Wall", "-Wextra", "-I", os.path.join(REPO, "include"),
           "-o", CODEC_LIB + ".tmp", os.path.join(CSRC, "codec_eval.c"), "-lm", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"gcc failed (rc={r.returncode})")
    os.replace(CODEC_LIB + ".tmp", CODEC_LIB)
    return CODEC_LIB


if __name__ == "__main__":
    print(build(verbose=True))
