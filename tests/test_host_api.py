"""The C++ host API (mercurydpm_fork_amd/host/Mercury3D.h): it compiles, its
species setters give the same bits as the reference's formulas restated in
mercurydpm_fork_amd/species.py, and drivers written against it reproduce the
reference's unit-test answers, on the oracle (CPU) and on the GPU.  The
drivers that write output files are checked in test_output_files.py."""
import os
import subprocess
import sys

import pytest

from mercurydpm_fork_amd import species as sp

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import output_check as oc  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mercurydpm_fork_amd", "host")
LIB = os.path.join(ROOT, "mercurydpm_fork_amd", "lib")
BIN = os.path.join(ROOT, "tests", "drivers", "bin")
# drivers that check the reference's answers themselves and print "ok":
# name -> (executable, arguments)
GOLDEN = os.path.join(ROOT, "tests", "golden")
SETUPS = os.path.join(ROOT, "tests", "drivers", "setups")
DRIVERS = {"TangentialSpringUnitTest": ("TangentialSpringUnitTest", [GOLDEN]),
           "InclinedPlaneUnitTest": ("SetupDriver", [os.path.join(SETUPS, "InclinedPlaneUnitTest.setup")]),
           "HertzContactRestitutionUnitTest": ("HertzContactRestitutionUnitTest", [GOLDEN]),
           "PeriodicWallsUnitTest": ("PeriodicWallsUnitTest", [GOLDEN])}

SETTERS_SRC = r'''
#include <cstdio>
#include "Mercury3D.h"
int main() {
    LinearViscoelasticSlidingFrictionSpecies a;
    a.setCollisionTimeAndRestitutionCoefficient(1e-3, .2, 1. / 8);
    LinearViscoelasticSpecies b;
    b.setCollisionTimeAndRestitutionCoefficient(0.005, 0.88, 1.0);
    HertzianViscoelasticMindlinSpecies c;
    c.setEffectiveElasticModulusAndRestitutionCoefficient(20000, 0.8);
    SpeciesHandler h;
    h.copyAndAddObject(a);
    h.copyAndAddObject(b);
    const dem_pair_params* m = h.getMixedObject(0, 1);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", a.getStiffness(), a.getDissipation(), b.getStiffness(),
                b.getDissipation(), c.getDissipation(), m->stiffness);
    return 0;
}
'''


def test_host_api_compiles_and_setters_match(tmp_path):
    src = tmp_path / "setters.cpp"
    src.write_text(SETTERS_SRC)
    exe = tmp_path / "setters"
    # the setters need no device: link the host implementation statically
    # against the ABI declarations only
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I" + HOST, "-o", str(exe), str(src),
                    *[os.path.join(HOST, f) for f in oc.HOST_SOURCES], "-L" + LIB,
                    "-lmercurydpm_dem", "-Wl,-rpath," + LIB], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [float(x) for x in out]
    a = sp.Species().set_collision_time_and_restitution(1e-3, .2, 1. / 8)
    b = sp.Species().set_collision_time_and_restitution(0.005, 0.88, 1.0)
    c = sp.Species().set_elastic_modulus_and_restitution(20000, 0.8)
    want = [a.stiffness, a.dissipation, b.stiffness, b.dissipation, c.dissipation,
            sp.average(a.stiffness, b.stiffness)]
    assert got == want


def test_driver_sources_compile(tmp_path):
    for f in sorted(os.listdir(oc.DRIVERS)):
        if f.endswith(".cpp"):
            subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + HOST, os.path.join(oc.DRIVERS, f)],
                           check=True)


@pytest.fixture(scope="module")
def oracle_drivers(tmp_path_factory):
    return oc.build_on_oracle(sorted({exe for exe, _ in DRIVERS.values()}),
                              str(tmp_path_factory.mktemp("oracle_drivers")))


@pytest.mark.parametrize("driver", sorted(DRIVERS))
def test_driver_reproduces_reference_answer_on_oracle(driver, oracle_drivers, tmp_path):
    exe, args = DRIVERS[driver]
    r = subprocess.run([oracle_drivers[exe], *args], cwd=str(tmp_path), capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr, file=sys.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("driver", sorted(DRIVERS))
def test_driver_reproduces_reference_answer(driver, tmp_path):
    name, args = DRIVERS[driver]
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    r = subprocess.run([exe, *args], cwd=str(tmp_path), capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr, file=sys.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


RNG_SRC = r'''
#include <cstdio>
#include <sstream>
#include "Mercury3D.h"
int main() {
    RNG fresh;
    std::ostringstream s;
    fresh.write(s);
    RNG seeded;
    seeded.setRandomSeed(5);
    double a = seeded.getRandomNumber(0.0005, 0.001), b = seeded.getRandomNumber(0.0005, 0.001);
    std::printf("%s|%.17g %.17g\n", s.str().c_str(), a, b);
    return 0;
}
'''


def test_rng_matches_reference_state(tmp_path):
    """RNG (Kernel/Math/RNG.cc): a fresh generator writes the state the
    reference's restart files hold (seed 0 after seeding the 607 Fibonacci
    lags: 784607267, e.g. tests/golden/free_fall_self_test.restart); the
    first radii of ForceSelfTest's pair case (seed 5) are the ones in the
    gold ForceSelfTest7.data."""
    src = tmp_path / "rng.cpp"
    src.write_text(RNG_SRC)
    exe = tmp_path / "rng"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I" + HOST, "-o", str(exe), str(src),
                    *[os.path.join(HOST, f) for f in oc.HOST_SOURCES], "-L" + LIB, "-lmercurydpm_dem",
                    "-Wl,-rpath," + LIB], check=True)
    state, radii = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("|")
    gold = open(os.path.join(GOLDEN, "free_fall_self_test.restart")).read()
    assert gold.split(" random ")[1].split()[:7] == state.split()
    # the radii column of the first frame of the pair case's .data file
    frame = [ln.split() for ln in open(os.path.join(GOLDEN, "force_self_test_7.data"))][1:3]
    want = [float(r[4]) for r in frame]
    got = [float(x) for x in radii.split()]
    assert [float(f"{g:.10g}") for g in got] == want


LOGGER_SRC = r'''
#include <cstdio>
#include <cstring>
#include "Mercury3D.h"
int main(int argc, char** argv) {
    const char* mode = argc > 1 ? argv[1] : "";
    if (!std::strcmp(mode, "info")) {
        logger(INFO, "a % b % c \\% d", 1, 2.5);
        logger(VERBOSE, "dropped at the default level");
        return 0;
    }
    if (!std::strcmp(mode, "solve")) {  // no time step: the default ERROR handler ends the program
        Mercury3D p;
        p.solve();
        std::puts("not reached");
        return 0;
    }
    if (!std::strcmp(mode, "throwing")) {  // a driver's own handler that throws
        LoggerOutput mine = loggerOutputDefault;
        mine.onError = [](std::string, std::string msg, Flusher) { throw std::invalid_argument(msg); };
        loggerOutput = &mine;
        Mercury3D p;
        try {
            p.solve();
        } catch (const std::invalid_argument& e) {
            std::printf("caught: %s\n", e.what());
        }
        loggerOutput = &loggerOutputDefault;
        return 0;
    }
    if (!std::strcmp(mode, "returning")) {  // a handler that returns: the operation still stops
        LoggerOutput mine = loggerOutputDefault;
        mine.onError = [](std::string, std::string, Flusher) {};
        loggerOutput = &mine;
        try {
            logger(ERROR, "value % out of range", 7);
        } catch (const std::runtime_error& e) {
            std::printf("stopped: %s\n", e.what());
        }
        return 0;
    }
    return 1;
}
'''


def test_logger_levels_and_error_handlers(tmp_path):
    """The host's logger: `%` placeholders, level filtering, ERROR ending the
    program with exit status (signed char)-15 by default (the reference's
    Logger.cc), a driver's own handler taking over, and a handler that
    returns still stopping the failing operation."""
    src = tmp_path / "logger.cpp"
    src.write_text(LOGGER_SRC)
    exe = tmp_path / "logger"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + HOST, "-o", str(exe), str(src),
                    *[os.path.join(HOST, f) for f in oc.HOST_SOURCES], "-L" + LIB,
                    "-lmercurydpm_dem", "-Wl,-rpath," + LIB], check=True)
    r = subprocess.run([str(exe), "info"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == "a 1 b 2.5 c % d\n", r
    r = subprocess.run([str(exe), "solve"], capture_output=True, text=True)
    assert r.returncode == (-15) % 256, r
    assert "ERROR in MercuryKernel: DPMBase::solve: time step not set" in r.stderr
    assert "not reached" not in r.stdout
    r = subprocess.run([str(exe), "throwing"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == "caught: DPMBase::solve: time step not set\n", r
    r = subprocess.run([str(exe), "returning"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout == "stopped: MercuryKernel: value 7 out of range\n", r
