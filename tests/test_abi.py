"""The C-ABI library loads and exports every function include/*.h declares; argument checks that need no device.  CPU only."""
import ctypes
import os
import re
import subprocess

import pytest

from mercurydpm_fork_amd import dem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mercurydpm_dem.h")
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_functions():
    src = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(dem_\w+)\s*\(", src, re.M)))


def test_header_declares_abi():
    names = declared_functions()
    for n in ("dem_create", "dem_step", "dem_compute_forces", "dem_upload_particles", "dem_download_contacts"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = dem.load_hip_library()
    for n in declared_functions():
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", dem.HIP_LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dem_\w+)", out))
    assert set(declared_functions()) <= exported


def test_abi_version_and_struct_sizes():
    lib = dem.load_hip_library()
    assert lib.dem_abi_version() == 9
    assert ctypes.sizeof(dem.PairParams) == 4 * 4 + 17 * 8
    assert ctypes.sizeof(dem.Stats) == 6 * 8 + 2 * 4 + 6 * 8 + 8 + 3 * 8 + 8
    assert ctypes.sizeof(dem.Wall) == 16 + 10 * 8 + 3 * 4 * 3 * 8
    assert ctypes.sizeof(dem.Periodic) == 8 + 3 * 8


def test_null_arguments_rejected():
    lib = dem.load_hip_library()
    assert lib.dem_create(None, None) == -1
    assert lib.dem_set_gravity(None, None) == -1
    assert lib.dem_destroy(None) == 0


def test_null_arguments_rejected_by_insertion_removal_and_closure_calls():
    """The entry points added for insertion, removal and the closure refuse a
    null context or null arrays before touching a device."""
    lib = dem.load_hip_library()
    from mercurydpm_fork_amd import closure
    closure._lib()
    null = ctypes.c_void_p()
    assert lib.dem_add_particles(None, 1, *([None] * 9)) == -1
    assert lib.dem_remove_particles(None, 1, None) == -1
    assert lib.dem_check_particles(None, 1, *([None] * 5)) == -1
    assert lib.dem_mlp_create(None, 0) == -1
    assert lib.dem_mlp_forward(null, 1, None, None) == -1
    assert lib.dem_mlp_gradient(null, 0, None, None, None, None) == -1
    assert lib.dem_mlp_train(null, 0, None, None, 1, 1e-3, None) == -1
    assert lib.dem_mlp_destroy(null) == 0
