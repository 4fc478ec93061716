"""Builds the package's native code in-tree:

* lib/libmercurydpm_dem.so  -- HIP engine for gfx950 (csrc/dem_engine.hip)
* lib/libmercurydpm_host.so -- C++ host API keeping the reference's driver
  interface (host/Mercury3D.h)
* tests/drivers/bin/*        -- drivers written against that API

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
    srcs = [os.path.join(HERE, "csrc", f) for f in ("dem_engine.hip", "closure_mlp.hip", "closure_mfma.hip")]
    out = os.path.join(LIB, "libmercurydpm_dem.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-ffp-contract=off"]
DRIVER_DIR = os.path.join(ROOT, "tests", "drivers")


HOST_SOURCES = [os.path.join(HERE, "host", f) for f in ("Mercury3D.cc", "DPMBaseOutput.cc", "DPMBaseRestart.cc", "Logger.cc",
                                                                "InsertionBoundary.cc", "Chute.cc")]


def build_host():
    """C++ host API (host/*.cc) linked against
    the engine."""
    out = os.path.join(LIB, "libmercurydpm_host.so")
    _run([CXX, *HOST_FLAGS, "-shared", "-o", out, *HOST_SOURCES,
          "-L" + LIB, "-lmercurydpm_dem", "-Wl,-rpath,$ORIGIN"])
    return out


def build_drivers():
    """Driver programs written against the host API (tests/drivers/*.cpp)."""
    outs = []
    bindir = os.path.join(DRIVER_DIR, "bin")
    os.makedirs(bindir, exist_ok=True)
    for f in sorted(os.listdir(DRIVER_DIR)):
        if not f.endswith(".cpp"):
            continue
        out = os.path.join(bindir, f[:-4])
        _run([CXX, *HOST_FLAGS, "-I" + os.path.join(HERE, "host"), "-o", out, os.path.join(DRIVER_DIR, f),
              "-L" + LIB, "-lmercurydpm_host", "-lmercurydpm_dem",
              "-Wl,-rpath,$ORIGIN/../../../mercurydpm_fork_amd/lib"])
        outs.append(out)
    return outs


def build_all():
    return [build_engine(), build_host(), *build_drivers()]


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
