"""Batched ensembles (mercurydpm_fork_amd/ensemble.py): independent heaps laid
side by side in one engine context run exactly as they run alone, because
no interaction range reaches across the gap between their tiles."""
import numpy as np
import pytest

import oracle_lib
from mercurydpm_fork_amd import dem, ensemble as E, scenarios as S


def members(sizes, seed=3):
    sp = E.calibration_species()
    return [E.heap(n, seed + k, sp) for k, n in enumerate(sizes)]


def test_pack_layout_and_checks():
    ms = members([50, 300, 120, 800, 20])
    p = E.pack(ms)
    st = p.scenario.state
    assert st.n == sum(m.state.n for m in ms)
    assert len(np.unique(st.id)) == st.n
    # the tiles are apart by at least the gap, so no two members can touch
    for a in range(len(ms)):
        for b in range(a + 1, len(ms)):
            xa, xb = p.member(a, st.x) + p.offsets[a], p.member(b, st.x) + p.offsets[b]
            lo_a, hi_a = xa[:, :2].min(0), xa[:, :2].max(0)
            lo_b, hi_b = xb[:, :2].min(0), xb[:, :2].max(0)
            sep = np.maximum(lo_b - hi_a, lo_a - hi_b).max()
            assert sep >= p.gap
    for k, m in enumerate(ms):  # shifting there and back rounds in the last bit
        np.testing.assert_allclose(p.member(k, st.x), m.state.x, rtol=0, atol=1e-12)
    assert p.check_separated(st.x)
    bad = members([30])[0]
    bad.periodic = [S.periodic_boundary(0, 0.0, 5.0)]
    with pytest.raises(ValueError):
        E.pack([ms[0], bad])


def _run(sc, steps, engine):
    e = engine()
    S.apply(e, sc)
    e.compute_forces()
    e.step(steps)
    return e


def test_ensemble_equals_members_on_oracle():
    """On the oracle: three heaps packed, 150 steps, each member's positions
    equal its own run's to round-off (the packed run sums a particle's
    forces in another order only through the ghost-free pair loop)."""
    ms = members([60, 150, 90])
    p = E.pack(ms)
    x = _run(p.scenario, 150, oracle_lib.engine).download()["x"]
    assert p.check_separated(x)
    for k, m in enumerate(ms):
        xk = _run(m, 150, oracle_lib.engine).download()["x"]
        np.testing.assert_allclose(p.member(k, x), xk, rtol=0, atol=1e-10)


@pytest.mark.gpu
def test_ensemble_equals_members_on_engine():
    """On the engine: six heaps of 100 to 1500 spheres packed, 400 steps:
    every member's contacts (ids mapped back) and positions equal its own
    engine run's (positions to round-off of differently ordered force sums)."""
    sizes = [100, 1500, 300, 800, 150, 600]
    ms = members(sizes, seed=11)
    p = E.pack(ms)
    g = _run(p.scenario, 400, dem.Engine)
    x = g.download()["x"]
    assert p.check_separated(x)
    cp = g.contacts()
    for k, m in enumerate(ms):
        e = _run(m, 400, dem.Engine)
        np.testing.assert_allclose(p.member(k, x), e.download()["x"], rtol=0, atol=1e-9)
        lo, hi = p.first[k], p.first[k + 1]
        own = {(int(a) - lo, int(b) if b & dem.DEM_WALL_KEY else int(b) - lo)
               for a, b in zip(cp.p_id, cp.i_id) if lo <= a < hi}
        assert own == e.contacts().keys()
