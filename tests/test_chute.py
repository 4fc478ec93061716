"""The host API's Chute class (Kernel/Chute.cc): the rough bottoms of
Chute::createBottom, the inflow boundary and cleanChute, through the driver
tests/drivers/ChuteInflowTest.cpp.  On the CPU the driver runs on the oracle
(oracle/dem_abi_on_oracle.c); on the GPU the engine's run is compared with
the oracle's.  The reference holds no gold output for these set-ups, so the
bottoms are checked against their construction rules: the closed-form
lattices of MONOLAYER_ORDERED and MONOLAYER_TRIANGULAR, and for
MONOLAYER_DISORDERED (random sequential placement until 500 candidates in a
row fail) that no two base spheres overlap and all lie in the placement
range; which random positions it keeps is not pinned."""
import os
import subprocess
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import output_check as oc  # noqa: E402

R, L, W = 0.5, 10.0, 5.0  # fixed particle radius, chute length and width of the driver


def first_frame(path):
    """Positions and radii of the first XBalls frame of a .data file."""
    lines = open(path).read().split("\n")
    n = int(lines[0].split()[0])
    if n == 0:
        return np.zeros((0, 3)), np.zeros(0)
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + n]])
    return rows[:, 0:3], rows[:, 6]


def run(exe, bottom, cwd, *extra):
    r = subprocess.run([exe, bottom, *extra], cwd=cwd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout + r.stderr
    summary = dict(zip(*[iter(r.stdout.split()[-8:])] * 2))
    return {k: float(v) for k, v in summary.items()}


@pytest.fixture(scope="module")
def oracle_driver(tmp_path_factory):
    return oc.build_on_oracle(["ChuteInflowTest"], str(tmp_path_factory.mktemp("chute")))["ChuteInflowTest"]


def check_bottom(bottom, x, r):
    fixed = np.isclose(r, R) & (x[:, 2] == 0.0)
    b = x[fixed]
    if bottom == "ordered":
        nx, ny = int(np.floor(L / (2 * R))), int(np.floor(W / (2 * R)))
        dx, dy = L / nx, W / ny
        want = np.array([[R + dx * i, R + dy * j, 0.0] for i in range(nx) for j in range(ny)])
        np.testing.assert_array_equal(b, want)
    elif bottom == "triangular":
        s = 0.1
        dx, dy = (1 + s) * 2 * R, np.sqrt(1 + s) * 2 * R
        nx, ny = int(np.floor(L / dx)) + 1, int(np.floor(W / dy)) + 1
        want = []
        for i in range(nx):
            for j in range(ny):
                x1, y1 = R + dx * i, R + dy * i  # y advances with i, as in Chute.cc
                x2, y2 = x1 + dx / 2, y1 + dy / 2
                if x1 < L and y1 < W:
                    want.append([x1, y1, 0.0])
                if x2 < L and y2 < W:
                    want.append([x2, y2, 0.0])
        np.testing.assert_allclose(b, np.array(want), rtol=0, atol=1e-12)
    elif bottom == "disordered":
        assert len(b) > 20
        d = np.linalg.norm(b[:, None, :2] - b[None, :, :2], axis=2) + np.eye(len(b)) * 1e9
        assert d.min() >= 2 * R
        assert np.all((b[:, 0] >= R) & (b[:, 0] <= L - R) & (b[:, 1] >= R) & (b[:, 1] <= W - R))
    else:
        assert len(b) == 0
    return len(b)


@pytest.mark.parametrize("bottom", ["ordered", "disordered", "triangular", "flat"])
def test_chute_on_oracle(bottom, oracle_driver, tmp_path):
    s = run(oracle_driver, bottom, str(tmp_path))
    x, r = first_frame(os.path.join(str(tmp_path), "ChuteInflowTest.data"))
    assert check_bottom(bottom, x, r) == s["bottom"]
    assert s["inserted"] > 10 and s["flow"] <= s["inserted"]


def test_chute_outflow_removed_on_oracle(oracle_driver, tmp_path):
    """Spheres entering at speed 10 reach the end of the 10-long chute within
    the run: cleanChute (every 101st step) removes those beyond x = 10, so
    fewer remain than were inserted, and none is left further out than one
    cleaning interval of travel."""
    s = run(oracle_driver, "ordered", str(tmp_path), "outflow")
    assert s["flow"] < s["inserted"], s
    assert s["xmax"] < L + 101 * 1e-4 * 12.0, s


@pytest.mark.gpu
@pytest.mark.parametrize("bottom,extra", [("ordered", ()), ("disordered", ()), ("ordered", ("outflow",))])
def test_chute_engine_vs_oracle(bottom, extra, oracle_driver, tmp_path):
    """The same chute on the engine and on the oracle: the same bottom, the
    same number of inserted (and, with outflow, removed) particles, and
    energies and centre of mass within 1e-6 (force sums differ in order, so
    the trajectories part at round-off level)."""
    exe = os.path.join(oc.DRIVERS, "bin", "ChuteInflowTest")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    (tmp_path / "g").mkdir()
    (tmp_path / "o").mkdir()
    sg = run(exe, bottom, str(tmp_path / "g"), *extra)
    so = run(oracle_driver, bottom, str(tmp_path / "o"), *extra)
    if extra:
        assert so["flow"] < so["inserted"]
    assert sg["bottom"] == so["bottom"] and sg["inserted"] == so["inserted"] and sg["flow"] == so["flow"]
    a = np.array(oc.numeric_rows(str(tmp_path / "g" / "ChuteInflowTest.ene")))
    b = np.array(oc.numeric_rows(str(tmp_path / "o" / "ChuteInflowTest.ene")))
    assert a.shape == b.shape
    # the centre of mass is 0/0 while only fixed particles exist
    assert np.array_equal(np.isnan(a), np.isnan(b))
    a, b = np.nan_to_num(a), np.nan_to_num(b)
    scale = np.max(np.abs(b), axis=0)
    assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b) + 1e-9 * scale)


def check_hook_run(summary):
    """ChuteHookTest: the driver's own actionsBeforeTimeStep holds the bed at
    rest on every one of its 500 steps (the speed left is one closing
    half-kick, g dt / 2 = 5e-5), and cleanChute removed the three spheres
    beyond the chute's ends."""
    assert summary["steps"] == 500
    assert summary["particles"] == 10
    assert summary["vmax"] < 1e-4


def run_hook(exe, cwd):
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    w = r.stdout.split()
    return {w[k]: float(w[k + 1]) for k in range(0, len(w) - 1, 2)}


def test_chute_subclass_hook_runs_every_step_on_oracle(tmp_path):
    exe = oc.build_on_oracle(["ChuteHookTest"], str(tmp_path))["ChuteHookTest"]
    check_hook_run(run_hook(exe, str(tmp_path)))


@pytest.mark.gpu
def test_chute_subclass_hook_runs_every_step(tmp_path):
    exe = os.path.join(oc.DRIVERS, "bin", "ChuteHookTest")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    check_hook_run(run_hook(exe, str(tmp_path)))
