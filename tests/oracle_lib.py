"""Loads the CPU oracle (oracle/liboracle_dem.so) for the tests: it is the
checker the engine is compared against, built from oracle/dem_oracle.c."""
import ctypes
import os
import subprocess

from mercurydpm_fork_amd import dem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle_dem.so")

_lib = None


def load():
    global _lib
    if _lib is None:
        src = os.path.join(ORACLE_DIR, "dem_oracle.c")
        if not os.path.exists(ORACLE_LIB) or os.path.getmtime(ORACLE_LIB) < os.path.getmtime(src):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(ORACLE_LIB)
        dem._declare(lib, "oracle_dem_", False)
        lib.oracle_dem_force_scale.restype = ctypes.c_int32
        lib.oracle_dem_force_scale.argtypes = [ctypes.c_void_p] * 3
        _lib = lib
    return _lib


def engine():
    return dem.Engine(lib=load(), prefix="oracle_dem_")


def force_scale(e):
    """Per-particle sums of |force term| and |torque term| of the last force pass."""
    import numpy as np
    fs = np.zeros(e.n)
    ts = np.zeros(e.n)
    r = e.lib.oracle_dem_force_scale(e.h, fs.ctypes.data_as(ctypes.c_void_p), ts.ctypes.data_as(ctypes.c_void_p))
    assert r == 0
    return fs, ts


ORACLE_OMP_LIB = os.path.join(ORACLE_DIR, "liboracle_dem_omp.so")
_omp = None


def engine_omp(threads):
    """The oracle's OpenMP timing build (force, wall and integration loops on
    `threads` threads; bench.py's multi-core CPU baseline).  Returns the
    engine and the number of threads it runs."""
    global _omp
    if _omp is None:
        src = os.path.join(ORACLE_DIR, "dem_oracle.c")
        if not os.path.exists(ORACLE_OMP_LIB) or os.path.getmtime(ORACLE_OMP_LIB) < os.path.getmtime(src):
            subprocess.run(["make", "-C", ORACLE_DIR, "liboracle_dem_omp.so"], check=True, stdout=subprocess.DEVNULL)
        lib = ctypes.CDLL(ORACLE_OMP_LIB)
        dem._declare(lib, "oracle_dem_", False)
        lib.oracle_dem_set_threads.restype = ctypes.c_int32
        lib.oracle_dem_set_threads.argtypes = [ctypes.c_int32]
        _omp = lib
    got = _omp.oracle_dem_set_threads(int(threads))
    return dem.Engine(lib=_omp, prefix="oracle_dem_"), got
