"""Builds the package's native code in-tree:

* lib/libmercurydpm_dem.so  -- HIP engine for gfx950 (csrc/dem_engine.hip)
* lib/libmercurydpm_host.so -- C++ host mirror of the reference API (host/)

Run ``python -m mercurydpm_fork_amd.build`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: no fused multiply-add, so every operation rounds as in
# the reference's CPU build.  Packed FP32 instructions are disabled for all
# HIP code built here.
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    src = os.path.join(HERE, "csrc", "dem_engine.hip")
    out = os.path.join(LIB, "libmercurydpm_dem.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, src])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_engine()]


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
