"""dem_check_particles_static (the entry point that allocates nothing: its
candidates are staged in the library's static per-GPU array) against
dem_check_particles on the same engine state, through the C ABI: two calls
in a row, the second with fewer candidates, so stale flags of the first call
would show.  GPU."""
import ctypes

import numpy as np
import pytest

from mercurydpm_fork_amd import dem
from mercurydpm_fork_amd import scenarios as S

pytestmark = pytest.mark.gpu


def candidates(sc, k, seed):
    """k candidate spheres spread over the box of the scenario's particles
    (some overlap particles or walls, some are free), with image ranges that
    are sometimes negative (no images)."""
    rng = np.random.default_rng(seed)
    x = sc.state.x
    lo, hi = x.min(axis=0) - 0.5, x.max(axis=0) + 0.5
    cx = lo + (hi - lo) * rng.random((k, 3))
    r = rng.uniform(0.05, 0.4, k)
    sp = rng.integers(0, 2, k).astype(np.int32)  # the mixture has two species
    md = np.where(rng.random(k) < 0.2, -1.0, r + 0.5)
    return cx, r, sp, md


@pytest.mark.parametrize("periodic", [(True, True), (False, False)])
def test_static_equals_allocating_entry_point(periodic):
    sc = S.mixture(n=3000, seed=5, lx=16.0, ly=16.0, lz=16.0, laws="hvm", periodic=periodic, walls=("bottom",))
    g = dem.Engine(device=0)
    S.apply(g, sc)
    g.compute_forces()
    g.step(3)
    for call, k in enumerate((dem.CHECK_STATIC_MAX, 700)):
        x, r, sp, md = candidates(sc, k, seed=11 + call)
        a = g.check_particles(x, r, sp, md)
        b = g.check_particles(x, r, sp, md, static=True)
        assert a.shape == b.shape == (k,)
        np.testing.assert_array_equal(a, b)
        assert 0 < a.sum() < k  # both outcomes occur
    g.close()


def test_static_refuses_more_than_its_maximum():
    sc = S.mixture(n=200, seed=2, laws="lvf", periodic=(False, False), walls=("bottom",))
    g = dem.Engine(device=0)
    S.apply(g, sc)
    k = dem.CHECK_STATIC_MAX + 1
    x = np.zeros((k, 3)); r = np.full(k, 0.1); sp = np.zeros(k, np.int32); md = np.full(k, -1.0)
    lib = dem.load_hip_library()
    out = np.zeros(k, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = lib.dem_check_particles_static(g.h, k, p(x), p(r), p(sp), p(md), p(out))
    assert rc == -4  # DEM_ERR_CAPACITY
    assert g.check_particles(x[:5] + 1e3, r[:5], sp[:5], md[:5], static=True).all()
    g.close()
