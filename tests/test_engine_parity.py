"""Engine (HIP, through the C ABI) versus the CPU oracle on identical seeded
inputs, and the engine against the reference's known answers.

Bar: the stored contact set and every spring are bit-identical; positions
after a step are bit-identical (they depend on forces of the previous pass
only through uploaded data); forces and torques agree within 1e-12 of the
per-particle sum of the magnitudes of the terms summed into them (the
engine gathers a particle's contacts in list order, the reference in handler
order, so only round-off of the sum may differ).
"""
import numpy as np
import pytest

import oracle_lib
from mercurydpm_fork_amd import dem
from mercurydpm_fork_amd import scenarios as S

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-12


def make_pair(sc, contacts=None, **kw):
    g = dem.Engine(**kw)
    S.apply(g, sc, contacts)
    o = oracle_lib.engine()
    S.apply(o, sc, contacts)
    return g, o


def assert_contacts_equal(g, o):
    cg, co = g.contacts(), o.contacts()
    assert cg.keys() == co.keys(), (len(cg.keys() ^ co.keys()), sorted(cg.keys() ^ co.keys())[:10])
    np.testing.assert_array_equal(cg.p_id, co.p_id)
    np.testing.assert_array_equal(cg.i_id, co.i_id)
    np.testing.assert_array_equal(cg.ts, co.ts)
    np.testing.assert_array_equal(cg.spring, co.spring)
    return co.n


def assert_contacts_close(g, o, rtol=1e-6):
    """After many steps the force summation order has perturbed the state at
    round-off level: same contact set and time stamps, springs close."""
    cg, co = g.contacts(), o.contacts()
    assert cg.keys() == co.keys(), (len(cg.keys() ^ co.keys()), sorted(cg.keys() ^ co.keys())[:10])
    np.testing.assert_array_equal(cg.ts, co.ts)
    scale = np.max(np.abs(co.spring)) if co.n else 1.0
    np.testing.assert_allclose(cg.spring, co.spring, rtol=rtol, atol=rtol * scale)
    return co.n


def assert_forces_close(g, o, tol=FORCE_TOL):
    a, b = g.download(), o.download()
    fs, ts = oracle_lib.force_scale(o)
    df = np.linalg.norm(a["f"] - b["f"], axis=1)
    dt = np.linalg.norm(a["t"] - b["t"], axis=1)
    assert np.all(df <= tol * np.maximum(fs, 1e-300) + 1e-300), float(np.max(df / np.maximum(fs, 1e-300)))
    assert np.all(dt <= tol * np.maximum(ts, 1e-300) + 1e-300), float(np.max(dt / np.maximum(ts, 1e-300)))
    return a, b


CASES = [
    dict(laws="lvf", periodic=(True, True), walls=("bottom", "top")),
    dict(laws="lvf", periodic=(False, False), walls=("bottom", "cylinder", "tilted")),
    dict(laws="full", periodic=(True, True), walls=("bottom",)),
    dict(laws="hvm", periodic=(True, True), walls=("bottom", "top")),
    dict(laws="hvm", periodic=(True, False), walls=("bottom", "cylinder")),
    # radius ratio above 2.5: three grid levels, pairs across all of them
    dict(laws="hvm", periodic=(True, True), walls=("bottom",), radii=(0.15, 0.5)),
    # AxisymmetricIntersectionOfWalls: a chamfered pedestal and a casing
    dict(laws="full", periodic=(False, False), walls=("axisymmetric",)),
]


def reseeded(sc, st):
    """The scenario with the particle state replaced by a downloaded one."""
    return S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, sc.gravity, sc.dt,
                      dem.State(id=sc.state.id, species=sc.state.species, radius=sc.state.radius,
                                inv_mass=sc.state.inv_mass, inv_inertia=sc.state.inv_inertia,
                                x=st["x"], v=st["v"], w=st["w"], q=st["q"], f=st["f"], t=st["t"]))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['laws']}-{'P' if c['periodic'][0] else 'np'}-{len(c['walls'])}w"
                         + ("-3lv" if "radii" in c else "") + ("-axi" if "axisymmetric" in c["walls"] else ""))
def test_first_pass_and_one_step(case):
    """The force pass of initialiseSolve, then one computeOneTimeStep from
    the oracle's state after that pass (so both start from identical forces)."""
    sc = S.mixture(n=400, seed=21, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    g, o = make_pair(sc)
    g.compute_forces()
    o.compute_forces()
    n0 = assert_contacts_equal(g, o)
    assert n0 > 100
    assert_forces_close(g, o)
    g, o = make_pair(reseeded(sc, o.download()), o.contacts())
    g.step(1)
    o.step(1)
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    np.testing.assert_array_equal(a["q"], b["q"])
    np.testing.assert_allclose(a["v"], b["v"], rtol=0, atol=1e-12)
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)
    a, b = g.download(), o.download()
    np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", CASES[:4], ids=lambda c: f"{c['laws']}-{'P' if c['periodic'][0] else 'np'}")
def test_reseeded_steps(case):
    """Run the oracle ahead, re-seed both from its state at steps 50 and 200
    (particles + interaction history), advance one step, compare."""
    sc = S.mixture(n=400, seed=33, inflate=1.02, vmax=1.0, **case)
    ref = oracle_lib.engine()
    S.apply(ref, sc)
    ref.compute_forces()
    done = 0
    for target in (50, 200):
        ref.step(target - done)
        done = target
        st = ref.download()
        ct = ref.contacts()
        g, o = make_pair(reseeded(sc, st), ct)
        for e in (g, o):
            e.set_time(done * sc.dt, done)
        g.step(1)
        o.step(1)
        assert_contacts_equal(g, o)
        a, b = assert_forces_close(g, o)
        np.testing.assert_array_equal(a["x"], b["x"])


def test_many_steps_with_frequent_rebuilds():
    """A tiny Verlet skin forces a list rebuild (re-sort + history re-match)
    almost every step; 300 steps must keep the contact set of the oracle."""
    sc = S.mixture(n=300, seed=3, laws="lvf", periodic=(True, True), walls=("bottom",), vmax=2.0)
    g, o = make_pair(sc, skin=0.004)
    g.compute_forces()
    o.compute_forces()
    for _ in range(6):
        g.step(50)
        o.step(50)
        assert_contacts_close(g, o)
        a, b = g.download(), o.download()
        np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-9)
    assert g.stats().n_rebuilds > 100


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['laws']}-{'P' if c['periodic'][0] else 'np'}-{len(c['walls'])}w"
                         + ("-3lv" if "radii" in c else ""))
def test_wavefront_list_sweep_equals_per_particle(case):
    """The Verlet lists built by the wavefronts' shared LDS sweep
    (DEM_SWEEP_WAVEFRONT) are the per-particle sweep's, entry for entry: 200
    steps with a small skin (a rebuild every few steps) give bit-identical
    states and springs either way; the first pass's contacts equal the
    oracle's."""
    sc = S.mixture(n=600, seed=23, lx=8.0, ly=8.0, lz=8.0, inflate=1.12, vmax=1.0, **case)
    outs = []
    for sweep in (dem.SWEEP_PER_PARTICLE, dem.SWEEP_WAVEFRONT):
        g = dem.Engine(skin=0.01, list_sweep=sweep)
        S.apply(g, sc)
        g.compute_forces()
        if sweep == dem.SWEEP_WAVEFRONT:
            o = oracle_lib.engine()
            S.apply(o, sc)
            o.compute_forces()
            assert_contacts_equal(g, o)
        g.step(200)
        assert g.stats().n_rebuilds > 10
        outs.append((g.download(), g.contacts()))
    for k in ("x", "v", "w", "f", "t"):
        np.testing.assert_array_equal(outs[0][0][k], outs[1][0][k])
    np.testing.assert_array_equal(outs[0][1].spring, outs[1][1].spring)


def test_deterministic():
    sc = S.mixture(n=400, seed=9, laws="hvm", periodic=(True, True), walls=("bottom",))
    outs = []
    for _ in range(2):
        g = dem.Engine()
        S.apply(g, sc)
        g.compute_forces()
        g.step(100)
        outs.append((g.download(), g.contacts()))
    for k in ("x", "v", "w", "f", "t"):
        np.testing.assert_array_equal(outs[0][0][k], outs[1][0][k])
    np.testing.assert_array_equal(outs[0][1].spring, outs[1][1].spring)


def run_engine(sc):
    g = dem.Engine()
    S.apply(g, sc)
    g.compute_forces()
    g.step(sc.n_steps())
    return g.download()


def test_engine_tangential_spring_known_answer():
    st = run_engine(S.tangential_spring_unit_test())
    gd = S.TANGENTIAL_SPRING_GOLD
    np.testing.assert_allclose(st["x"][1], gd["x1"], atol=1e-7, rtol=0)
    np.testing.assert_allclose(st["v"][1], gd["v1"], atol=1e-7, rtol=0)
    np.testing.assert_allclose(st["w"][1], gd["w1"], atol=1e-7, rtol=0)
    np.testing.assert_allclose(st["x"][0], gd["x0"], atol=1e-7, rtol=0)


def test_engine_periodic_walls_gold():
    st = run_engine(S.periodic_walls_unit_test())
    for k, (x, v) in enumerate(S.PERIODIC_WALLS_GOLD):
        np.testing.assert_allclose(st["x"][k], x, atol=1e-10, rtol=0, err_msg=f"E{k}")
        np.testing.assert_allclose(st["v"][k], v, atol=1e-10, rtol=0, err_msg=f"E{k}")


def test_engine_inclined_plane_and_hertz():
    st = run_engine(S.inclined_plane_unit_test())
    assert abs(st["v"][0][0] - 1.81167) < 1e-4
    assert abs(st["v"][0][1] - 0.00234107) < 1e-4
    st = run_engine(S.hertz_restitution_unit_test())
    assert abs(st["v"][1][0] - 0.8) < 1e-2 and abs(st["v"][0][0] + 0.8) < 1e-2


@pytest.mark.parametrize("kind", ["rolling", "torsion"])
def test_engine_rolling_and_torsion_friction_self_test_gold(kind):
    """The reference's RollingFrictionSelfTest / TorsionFrictionSelfTest gold
    .torque on the engine: every row to its 6 printed digits."""
    from friction_check import check_torque_rows, squeezed_sphere_rows
    check_torque_rows(squeezed_sphere_rows(dem.Engine(), kind), kind)


def test_engine_periodic_walls_with_friction_unit_test():
    """PeriodicWallsWithSlidingFrictionUnitTest on the engine (full friction
    across periodic boundaries, two removals mid-run): the reference's
    shift invariance within 1e-10, and the oracle's final state to
    round-off."""
    from friction_check import check_periodic_friction, run_periodic_friction
    d = run_periodic_friction(dem.Engine())
    check_periodic_friction(d)
    o = run_periodic_friction(oracle_lib.engine())
    np.testing.assert_allclose(d["x"], o["x"], atol=1e-13, rtol=0)
    np.testing.assert_allclose(d["w"], o["w"], atol=1e-12, rtol=0)


def test_box_lvf_vs_oracle():
    """5000-sphere box (periodic x/y, floor, gravity), 400 steps: same stored
    contacts and state as the oracle."""
    sc = S.box_lvf()
    g, o = make_pair(sc)
    g.compute_forces()
    o.compute_forces()
    g.step(400)
    o.step(400)
    assert_contacts_close(g, o)
    a, b = g.download(), o.download()
    np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-10)


def test_engine_oblique_impact_mindlin_fstat_gold():
    """The reference's ObliqueImpactSelfTest.fstat (Hertz-Mindlin sliding
    against a wall), every saved contact row to its 10 printed digits."""
    from fstat_check import check_oblique_impact
    sc = S.oblique_impact_self_test()
    g = dem.Engine()
    S.apply(g, sc)
    g.compute_forces()
    check_oblique_impact(g)


@pytest.mark.parametrize("which", ["loading", "normal_and_tangential_loading", "frame_independence"])
def test_engine_mindlin_self_test_fstat_gold(which):
    """The reference's MindlinSelfTest*.fstat through dem_step_begin /
    dem_set_kinematics / dem_step_end (motion prescribed where the reference
    calls actionsBeforeTimeStep)."""
    from fstat_check import check_mindlin_self_test
    check_mindlin_self_test(dem.Engine, which)


def test_engine_split_step_equals_step():
    """dem_step_begin + dem_step_end is dem_step(1), bit for bit."""
    sc = S.mixture(n=300, seed=5, laws="hvm", periodic=(True, True), walls=("bottom",))
    a = dem.Engine()
    b = dem.Engine()
    S.apply(a, sc)
    S.apply(b, sc)
    a.compute_forces()
    b.compute_forces()
    for _ in range(20):
        a.step(1)
        b.step_begin()
        b.step_end()
    da, db = a.download(), b.download()
    for k in ("x", "v", "w", "q", "f", "t"):
        assert np.array_equal(da[k], db[k]), k
    assert_contacts_equal(a, b)


def test_engine_cg_basic_self_test_gold():
    """dem_cg_z on CGBasicSelfTest against the reference's
    SelfTestData/CGBasicSelfTest.1.stat (CG<Z, Gauss>, 1000 points)."""
    from test_oracle_golden import check_cg_basic_gold, run_cg_basic
    check_cg_basic_gold(run_cg_basic(dem.Engine))


@pytest.mark.parametrize("laws,periodic,walls", [("hvm", (True, True), ("bottom",)),
                                                 ("lvf", (False, False), ("bottom", "cylinder"))])
def test_engine_cg_vs_oracle(laws, periodic, walls):
    """CG<Z, Gauss> of the engine and the oracle three steps after identical
    inputs with fixed particles (all integral types: I_TO_P, CONTACT_TO_P,
    I_TO_CONTACT, walls): within 1e-10 of each column's largest value (the
    contact forces agree to round-off of their summation).  Right after step
    0 the reference still holds interactions that ended in that step (the
    initial pass and step 0 both stamp 1) with their stale forces; the engine
    does not average those, so the comparison starts later."""
    sc = S.mixture(n=400, seed=17, laws=laws, periodic=periodic, walls=walls, fixed_frac=0.15, inflate=1.1)
    g, o = make_pair(sc)
    g.compute_forces()
    o.compute_forces()
    g.step(3)
    o.step(3)
    x = sc.state.x
    lo = (float(x[:, 0].min()) - 1, float(x[:, 1].min()) - 1, float(x[:, 2].min()) - 1)
    hi = (float(x[:, 0].max()) + 1, float(x[:, 1].max()) + 1, float(x[:, 2].max()) + 1)
    a = g.cg_z(64, 0.3, lo, hi)
    b = o.cg_z(64, 0.3, lo, hi)
    scale = np.max(np.abs(b), axis=0) + 1e-300
    assert np.max(np.abs(b[:, 11:20])) > 0, "no contact stress"
    assert np.max(np.abs(b[:, 20:23])) > 0, "no boundary interaction force"
    err = np.abs(a - b) / scale
    k, col = np.unravel_index(err.argmax(), err.shape)
    assert err.max() < 1e-10, (int(k), dem.CG_COLUMNS[col], a[k, col], b[k, col])


def test_box_lvf_long_run_statistics():
    """The full 20,000-step run of the 5000-sphere box (C1) against the
    oracle's (tests/golden/box_lvf_longrun.txt): the two trajectories part at
    round-off level and diverge, the macroscopic state must not.  Kinetic
    energy (translational and rotational) every 1000 steps within 1% of the
    run's largest value, and the final CG<Z, Gauss> density and contact-stress
    zz profiles within 1% of their peaks."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "box_lvf_longrun.txt")
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    en = np.array([[float(v) for v in r[1:]] for r in rows if r[0] == "E"])
    cg = np.array([[float(v) for v in r[1:]] for r in rows if r[0] == "C"])
    sc = S.box_lvf()
    g = dem.Engine()
    S.apply(g, sc)
    g.compute_forces()
    m = 1.0 / sc.state.inv_mass
    inertia = 1.0 / sc.state.inv_inertia[:, 0]
    got = []
    done = 0
    for step, _, _ in en:
        if step > done:
            g.step(int(step) - done)
            done = int(step)
        st = g.download()
        got.append((0.5 * np.sum(m * np.sum(st["v"] ** 2, axis=1)), 0.5 * np.sum(inertia * np.sum(st["w"] ** 2, axis=1))))
    got = np.array(got)
    for c, name in ((0, "Ekin"), (1, "Erot")):
        err = np.max(np.abs(got[:, c] - en[:, c + 1])) / np.max(np.abs(en[:, c + 1]))
        assert err < 0.01, (name, err)
    out = g.cg_z(len(cg), 0.5, (0.0, 0.0, 0.0), (20.0, 20.0, 40.0))
    for col, k, name in ((1, 1, "density"), (19, 2, "contact stress zz")):
        err = np.max(np.abs(out[:, col] - cg[:, k])) / np.max(np.abs(cg[:, k]))
        assert err < 0.01, (name, err)


@pytest.mark.parametrize("case", [CASES[0], CASES[3], CASES[4]], ids=["lvf-P", "hvm-P", "hvm-Px"])
def test_engine_interactions_vs_oracle(case):
    """dem_download_interactions (what the .fstat / .ene writers read) after
    one step from the oracle's state: the same interactions, flags, springs,
    contact forces and tangential forces, bit for bit; the geometry, which
    both recompute, to 1e-12 of the largest radius."""
    sc = S.mixture(n=400, seed=21, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    ref = oracle_lib.engine()
    S.apply(ref, sc)
    ref.compute_forces()
    ref.step(3)
    # both from the oracle's state, so that only the last pass's summation
    # order differs
    g, o = make_pair(reseeded(sc, ref.download()), ref.contacts())
    for e in (g, o):
        e.set_time(3 * sc.dt, 3)
        e.step(1)
    a, b = g.interactions(), o.interactions()
    assert len(a) == len(b) and len(a) > 0
    for k in ("p_id", "i_id", "time_stamp", "current"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_array_equal(a["spring"], b["spring"])
    assert np.all(b["current"] == 1)
    # a contact's force depends on its pair's state only, which is
    # bit-identical here
    np.testing.assert_array_equal(a["force"], b["force"])
    np.testing.assert_array_equal(a["tangential"], b["tangential"])
    L = float(np.max(sc.state.radius))
    for k in ("contact", "normal", "overlap", "distance"):
        d = np.max(np.abs(a[k] - b[k]))
        assert d <= 1e-12 * (1.0 if k == "normal" else L), (k, d)


def test_engine_rejects_non_finite_input():
    """Non-finite positions, velocities, gravity or prescribed kinematics are
    refused at the ABI (DEM_ERR_ARG) before anything reaches the device; a
    valid state uploaded afterwards steps as usual."""
    sc = S.mixture(n=100, seed=3, laws="lvf", periodic=(True, True), walls=("bottom",))
    bad = reseeded(sc, dict(x=sc.state.x.copy(), v=sc.state.v, w=sc.state.w, q=sc.state.q, f=None, t=None))
    bad.state.x[7, 1] = np.nan
    g = dem.Engine()
    with pytest.raises(dem.DemError, match="finite"):
        S.apply(g, bad)
    S.apply(g, sc)
    with pytest.raises(dem.DemError, match="finite"):
        g.set_gravity((0.0, 0.0, np.inf))
    g.compute_forces()
    x = g.download()["x"]
    x[3, 0] = np.nan
    g.step_begin()
    with pytest.raises(dem.DemError, match="finite"):
        g.set_kinematics(x=x)
    g.step_end()
    g.step(5)
    assert np.all(np.isfinite(g.download()["x"]))


def _contact_force_scale(e, sc):
    """Per particle: the sum of |force| over its interactions."""
    idx = {int(i): k for k, i in enumerate(sc.state.id)}
    scale = np.zeros(sc.state.n)
    for r in e.interactions():
        m = np.linalg.norm(r["force"])
        scale[idx[int(r["p_id"])]] += m
        if not r["i_id"] & dem.DEM_WALL_KEY:
            scale[idx[int(r["i_id"])]] += m
    return scale


@pytest.mark.parametrize("laws,seed", [("lvf", 31), ("hvm", 32), ("full", 33)])
def test_mixed_precision_against_fp64(laws, seed):
    """DEM_PRECISION_MIXED against FP64 on identical inputs (the state of an
    FP64 run after 100 steps): the same stored contacts (the contact test is
    FP64 in both); forces per particle within 1e-6 on average and 2e-4 at
    most of the sum of the magnitudes of its contact forces (FP32 contact
    geometry and forces; measured: mean 1e-7 to 3e-7, max 1e-5 to 9e-5);
    torques likewise per radius; springs to 1e-5 of their scale."""
    sc = S.mixture(n=800, seed=seed, laws=laws, inflate=1.2, lx=8.4, ly=8.4, lz=8.4)
    ref = dem.Engine()
    S.apply(ref, sc)
    ref.compute_forces()
    ref.step(100)
    st, ct = ref.download(), ref.contacts()
    runs = {}
    for prec in (dem.PRECISION_FP64, dem.PRECISION_MIXED):
        e = dem.Engine(precision=prec)
        S.apply(e, reseeded(sc, st), ct)
        e.set_time(100 * sc.dt, 100)
        e.compute_forces()
        runs[prec] = e
    a, b = runs[dem.PRECISION_FP64], runs[dem.PRECISION_MIXED]
    ca, cb = a.contacts(), b.contacts()
    assert ca.keys() == cb.keys() and ca.n > 300
    sscale = max(np.max(np.abs(ca.spring)), 1e-300)
    assert np.max(np.abs(ca.spring - cb.spring)) <= 1e-5 * sscale
    fa, fb = a.download(), b.download()
    scale = _contact_force_scale(a, sc)
    m = scale > 0
    err = np.linalg.norm(fb["f"] - fa["f"], axis=1)[m] / scale[m]
    terr = np.linalg.norm(fb["t"] - fa["t"], axis=1)[m] / (scale[m] * sc.state.radius[m])
    assert err.mean() <= 1e-6 and err.max() <= 2e-4, (err.mean(), err.max())
    assert terr.mean() <= 1e-6 and terr.max() <= 2e-4, (terr.mean(), terr.max())


def _subset(sc, idx):
    """The scenario restricted to the particles idx (the rest are added later)."""
    st = sc.state
    part = dem.State(id=st.id[idx], species=st.species[idx], radius=st.radius[idx], inv_mass=st.inv_mass[idx],
                     inv_inertia=st.inv_inertia[idx], x=st.x[idx], v=st.v[idx], w=st.w[idx], q=st.q[idx],
                     f=None if st.f is None else st.f[idx], t=None if st.t is None else st.t[idx])
    return S.Scenario(sc.name, sc.species, sc.walls, sc.periodic, sc.gravity, sc.dt, part)


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[4]], ids=["lvf-P", "lvf-np-3w", "hvm-Px"])
def test_check_particles_against_oracle(case):
    """DPMBase::checkParticleForInteraction on the device (dem_check_particles)
    against the oracle: 6000 candidates spread over the box and across its
    periodic faces and walls, with periodic images (the range of a handler
    holding the largest particle) and without (an empty handler): the same
    decision for every candidate."""
    sc = S.mixture(n=400, seed=41, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    g, o = make_pair(sc)
    g.compute_forces()
    o.compute_forces()
    rng = np.random.default_rng(5)
    k = 6000
    lo, hi = sc.state.x.min(0) - 0.6, sc.state.x.max(0) + 0.6
    x = rng.uniform(lo, hi, (k, 3))
    r = rng.uniform(0.03, 0.3, k)
    sp = rng.integers(0, sc.n_species, k)
    for md in (r + sc.state.radius.max(), -np.ones(k)):
        a, b = g.check_particles(x, r, sp, md), o.check_particles(x, r, sp, md)
        np.testing.assert_array_equal(a, b)
        assert 0.01 < a.mean() < 0.99, a.mean()


@pytest.mark.parametrize("case", [CASES[0], CASES[4]], ids=["lvf-P", "hvm-Px"])
def test_add_particles_mid_step(case):
    """Particles appended between step_begin and step_end, where insertion
    boundaries add them (DPMBase.cc:4348-4351), take part in that step's
    force pass: after the step the contacts and positions equal the oracle's
    with the same append; 20 steps later they still agree to round-off."""
    sc = S.mixture(n=400, seed=43, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    first, later = np.arange(300), np.arange(300, 400)
    base = _subset(sc, first)
    ref = oracle_lib.engine()
    S.apply(ref, base)
    ref.compute_forces()
    g, o = make_pair(reseeded(base, ref.download()), ref.contacts())
    added = _subset(sc, later).state
    for e in (g, o):
        e.step_begin()
        e.add_particles(added)
        e.step_end()
    assert g.n == o.n == 400
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)
    a, b = g.download(), o.download()
    np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", [CASES[0], CASES[4], CASES[5]], ids=["lvf-P", "hvm-Px", "hvm-P-3lv"])
def test_add_particles_in_place(case):
    """Appends that fit the resident arrays and grid are written in place:
    the capacity dem_get_stats reports stays, and only the new particles are
    copied.  Three batches: the first goes through a re-upload that leaves
    room, the next two are appended in place, one between step_begin and
    step_end and one between steps.  After the step that takes the first
    in-place batch, contacts and positions equal the oracle's and forces
    agree to round-off; 20 steps after the last batch the states still agree
    to round-off."""
    sc = S.mixture(n=400, seed=47, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    # the largest particles first, so that each later one is within the
    # largest radius of its grid level; ids in insertion order
    order = np.argsort(-sc.state.radius, kind="stable")
    sc = _subset(sc, order)
    sc.state.id = np.arange(1, 401, dtype=np.uint32)
    batches = [np.arange(0, 200), np.arange(200, 250), np.arange(250, 330), np.arange(330, 400)]
    base = _subset(sc, batches[0])
    ref = oracle_lib.engine()
    S.apply(ref, base)
    ref.compute_forces()
    g, o = make_pair(reseeded(base, ref.download()), ref.contacts())
    assert g.stats().capacity == 200
    for e in (g, o):
        e.add_particles(_subset(sc, batches[1]).state)
    cap = g.stats().capacity
    assert cap == 250 + 1024
    for e in (g, o):
        e.step_begin()
        e.add_particles(_subset(sc, batches[2]).state)
        e.step_end()
    assert g.n == o.n == 330
    assert g.stats().capacity == cap  # in place
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    g.step(5)
    o.step(5)
    # ids above those the periodic images took meanwhile on the oracle
    # (each image of PeriodicBoundary::createPeriodicParticle takes the
    # handler's next id)
    last = _subset(sc, batches[3]).state
    last.id = last.id + np.uint32(100000)
    for e in (g, o):
        e.add_particles(last)
    assert g.n == o.n == 400
    assert g.stats().capacity == cap
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)
    a, b = g.download(), o.download()
    np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(a["v"], b["v"], rtol=0, atol=1e-8)


@pytest.mark.parametrize("case", [CASES[3], CASES[5]], ids=["hvm-P", "hvm-P-3lv"])
def test_add_particles_larger_than_the_grid(case):
    """Appended particles larger than the largest of their grid level do not
    fit the resident cells: they go through the re-upload that derives the
    grid again, although the arrays have room (the capacity grows).  The
    states after it agree with the oracle's to round-off."""
    sc = S.mixture(n=400, seed=55, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    # the smallest particles first: every later one is larger than the
    # resident ones
    sc = _subset(sc, np.argsort(sc.state.radius, kind="stable"))
    sc.state.id = np.arange(1, 401, dtype=np.uint32)
    base = _subset(sc, np.arange(300))
    ref = oracle_lib.engine()
    S.apply(ref, base)
    ref.compute_forces()
    g, o = make_pair(reseeded(base, ref.download()), ref.contacts())
    for e in (g, o):
        e.step_begin()
        e.add_particles(_subset(sc, np.arange(300, 320)).state)
        e.step_end()
    cap = g.stats().capacity
    assert cap == 320 + 1024
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    last = _subset(sc, np.arange(320, 400)).state
    last.id = last.id + np.uint32(100000)  # above the ids periodic images took on the oracle
    for e in (g, o):
        e.add_particles(last)
    assert g.n == o.n == 400
    assert g.stats().capacity == 400 + 1024  # re-uploaded, although 400 <= cap
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)
    np.testing.assert_allclose(g.download()["x"], o.download()["x"], rtol=0, atol=1e-9)


def test_empty_system_then_insertion():
    """A run that starts empty (its insertion boundaries fill it): steps of
    an empty system only advance the clock; particles added later run as on
    the oracle."""
    sc = S.mixture(n=300, seed=45, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, laws="lvf")
    empty = _subset(sc, np.arange(0))
    g, o = make_pair(empty)
    for e in (g, o):
        e.compute_forces()
        e.step(3)
        e.step_begin()
        assert e.download()["x"].shape == (0, 3)
        assert e.check_particles([[1.0, 1.0, 1.0]], 0.3, 0, -1.0).shape == (1,)
        e.add_particles(sc.state)
        e.step_end()
    assert g.stats().n_steps == 4
    a, b = g.download(), o.download()
    np.testing.assert_array_equal(a["x"], b["x"])
    g.step(10)
    o.step(10)
    assert_contacts_close(g, o)
    np.testing.assert_allclose(g.download()["x"], o.download()["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", [CASES[0], CASES[1]], ids=["lvf-P", "lvf-np-3w"])
def test_remove_particles(case):
    """ParticleHandler::removeObject through dem_remove_particles, between
    dem_step_begin and dem_step_end (where deletion boundaries act) and
    between steps: the last particle takes each removed index, the removed
    particles' interactions go; contacts and positions after the step equal
    the oracle's, and 20 steps later agree to round-off.  (After removals
    the reference orients a periodic-image interaction by handler index, the
    engine by id: with periodic boundaries an interaction may then be stored
    as (j, i) by the oracle and as (i, j) by the engine; lvf-P happens to
    have none.)"""
    sc = S.mixture(n=400, seed=47, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    ref = oracle_lib.engine()
    S.apply(ref, sc)
    ref.compute_forces()
    g, o = make_pair(reseeded(sc, ref.download()), ref.contacts())
    gone = [5, 17, 397, 0, 120, 394]  # each index refers to the handler after the previous removals
    for e in (g, o):
        e.step_begin()
        e.remove_particles(gone[:4])
        e.step_end()
    assert g.n == o.n == 396
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    for e in (g, o):
        e.remove_particles(gone[4:])
        e.step(20)
    assert_contacts_close(g, o)
    np.testing.assert_allclose(g.download()["x"], o.download()["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("case", [CASES[1], CASES[6]], ids=["lvf-np-3w", "full-np-axi"])
def test_remove_many_particles(case):
    """Removals of many particles at once (the device compacts the particle
    arrays and rebuilds over the kept ones): 120 random handler indices
    between dem_step_begin and dem_step_end, then 150 between steps, then
    every remaining particle.  Contacts and positions after the first step
    equal the oracle's; 20 steps after the second removal the states agree
    to round-off; with every particle gone, steps only advance the clock."""
    sc = S.mixture(n=400, seed=51, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    ref = oracle_lib.engine()
    S.apply(ref, sc)
    ref.compute_forces()
    g, o = make_pair(reseeded(sc, ref.download()), ref.contacts())
    rng = np.random.default_rng(11)
    first = [int(rng.integers(0, 400 - j)) for j in range(120)]
    for e in (g, o):
        e.step_begin()
        e.remove_particles(first)
        e.step_end()
    assert g.n == o.n == 280
    n0 = assert_contacts_equal(g, o)
    assert n0 > 50
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    np.testing.assert_allclose(a["v"], b["v"], rtol=0, atol=1e-12)
    second = [int(rng.integers(0, 280 - j)) for j in range(150)]
    for e in (g, o):
        e.remove_particles(second)
        e.step(20)
    assert g.n == o.n == 130
    assert_contacts_close(g, o)
    np.testing.assert_allclose(g.download()["x"], o.download()["x"], rtol=0, atol=1e-9)
    for e in (g, o):
        e.remove_particles(np.zeros(130, np.int64))
        e.step(3)
    assert g.n == o.n == 0
    assert g.stats().n_steps == 24


@pytest.mark.parametrize("case", [CASES[1], CASES[6]], ids=["lvf-np-3w", "full-np-axi"])
def test_remove_then_append_in_the_same_step(case):
    """A deletion boundary and an insertion boundary acting in one step:
    removals compact the particle arrays on the device, and the appended
    particles then go in place into the room the removals left (the
    capacity stays).  After the step contacts and positions equal the
    oracle's; 20 steps later the states agree to round-off."""
    sc = S.mixture(n=400, seed=53, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **case)
    # the largest particles first (each appended one is within the largest
    # radius of its grid level); ids in insertion order
    sc = _subset(sc, np.argsort(-sc.state.radius, kind="stable"))
    sc.state.id = np.arange(1, 401, dtype=np.uint32)
    base = _subset(sc, np.arange(350))
    ref = oracle_lib.engine()
    S.apply(ref, base)
    ref.compute_forces()
    g, o = make_pair(reseeded(base, ref.download()), ref.contacts())
    rng = np.random.default_rng(13)
    gone = [int(rng.integers(0, 350 - j)) for j in range(60)]
    added = _subset(sc, np.arange(350, 400)).state
    for e in (g, o):
        e.step_begin()
        e.remove_particles(gone)
        e.add_particles(added)
        e.step_end()
    assert g.n == o.n == 340
    assert g.stats().capacity == 350
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)
    np.testing.assert_allclose(g.download()["x"], o.download()["x"], rtol=0, atol=1e-9)


def test_find_in_region_against_oracle():
    """dem_find_in_region (the deletion boundaries' test on the device)
    against the oracle: half-spaces of several orientations and open boxes,
    mid-step; the same handler indices."""
    sc = S.mixture(n=400, seed=49, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, **CASES[0])
    g, o = make_pair(sc)
    for e in (g, o):
        e.compute_forces()
        e.step(5)
        e.step_begin()
    rng = np.random.default_rng(9)
    for _ in range(6):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        prm = np.append(nrm, rng.uniform(-2.0, 6.0))
        lo = rng.uniform(0.0, 4.0, 3)
        box = np.concatenate([lo, lo + rng.uniform(1.0, 4.0, 3)])
        for kind, params in ((0, prm), (1, box)):
            a, b = g.find_in_region(kind, params), o.find_in_region(kind, params)
            np.testing.assert_array_equal(a, b)
    assert any(len(g.find_in_region(0, np.array([0.0, 0.0, -1.0, -3.5]))) > 0 for _ in range(1))
    for e in (g, o):
        e.step_end()


def test_engine_axisymmetric_wall_closed_forms(monkeypatch):
    """The oracle's closed-form checks of AxisymmetricIntersectionOfWalls
    (pedestal face and rim edge, side face, casing) on the engine."""
    import test_oracle_golden as tg
    monkeypatch.setattr(tg.oracle_lib, "engine", lambda: dem.Engine())
    tg.test_axisymmetric_wall_closed_forms()


def test_engine_walls_removed_mid_run():
    """WallHandler::removeLastObject between steps (the calibration heap
    lifts its casing, CalibrationHeap.cpp:137-143): the casing's interactions
    are deleted, the pedestal's kept; the run continues as on the oracle."""
    sc = S.mixture(n=400, seed=41, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, laws="full", periodic=(False, False),
                   walls=("axisymmetric",))
    g, o = make_pair(sc)
    for e in (g, o):
        e.compute_forces()
        e.step(30)
    casing = [int(i) for i in o.contacts().i_id if int(i) == dem.DEM_WALL_KEY | 1]
    assert casing, "the casing touches spheres before it is lifted"
    for e in (g, o):
        e.set_walls(sc.walls[:1])
    for e in (g, o):
        assert not any(int(i) == dem.DEM_WALL_KEY | 1 for i in e.contacts().i_id)
    # 30 steps in: the summation order has perturbed the springs at round-off
    # level, so the sets and stamps are equal and the springs close
    assert_contacts_close(g, o)
    for e in (g, o):
        e.step(30)
    assert_contacts_close(g, o)
    a, b = g.download(), o.download()
    np.testing.assert_allclose(a["x"], b["x"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("laws", ["full", "hvm"])
def test_step_graph_replays_equal_launches(monkeypatch, laws):
    """Steady steps of dem_step replayed from captured two-step HIP graphs
    (DEM_STEP_GRAPH=1) against every step launched on its own (the
    default): bit-identical states, springs, time stamps and
    rebuild counts after calls of odd and even lengths, with a rebuild
    almost every step, downloads and coarse graining between the calls, and
    a re-upload of the state (new buffers: the graph is captured again)."""
    sc = S.mixture(n=500, seed=31, lx=8.0, ly=8.0, lz=8.0, inflate=1.12, vmax=2.0, laws=laws,
                   periodic=(True, True), walls=("bottom", "top"), radii=(0.15, 0.5))
    outs = []
    for graphs in ("1", ""):
        if graphs:
            monkeypatch.setenv("DEM_STEP_GRAPH", graphs)
        else:
            monkeypatch.delenv("DEM_STEP_GRAPH", raising=False)
        g = dem.Engine(skin=0.004)
        S.apply(g, sc)
        g.compute_forces()
        trace = []
        for k, n in enumerate((7, 40, 5, 64, 3, 101)):
            g.step(n)
            st = g.download()
            trace.append(st["x"].copy())
            if k == 2:
                g.cg_z(16, 0.5, lo=(0.0, 0.0, 0.0), hi=(8.0, 8.0, 8.0))
            if k == 3:
                ct = g.contacts()
                stats = g.stats()
                g.upload(reseeded(sc, st).state)
                g.upload_contacts(ct)
                g.set_time(stats.time, int(stats.n_steps))
                g.compute_forces()
        outs.append((trace, g.download(), g.contacts(), g.stats()))
        g.close()
    (ta, a, ca, sa), (tb, b, cb, sb) = outs
    for xa, xb in zip(ta, tb):
        np.testing.assert_array_equal(xa, xb)
    for k in ("x", "v", "w", "q", "f", "t"):
        np.testing.assert_array_equal(a[k], b[k])
    assert ca.keys() == cb.keys()
    oa, ob = np.lexsort((ca.i_id, ca.p_id)), np.lexsort((cb.i_id, cb.p_id))
    np.testing.assert_array_equal(ca.spring[oa], cb.spring[ob])
    np.testing.assert_array_equal(ca.ts[oa], cb.ts[ob])
    assert sa.n_rebuilds == sb.n_rebuilds and sa.n_rebuilds > 100
    assert sa.n_steps == sb.n_steps and sa.time == sb.time


@pytest.mark.parametrize("offset", [(4096.3, -2048.7, 1024.1), (-3.0e5, 1.7e5, 2.5e4)])
def test_contacts_far_from_the_origin(offset):
    """A mixture translated far from the origin (cell coordinates in the
    thousands and in the hundred-thousands) keeps the oracle's contact set,
    springs and forces bit for bit through the first pass and one step, and
    its contact set over 20 more."""
    sc = S.mixture(n=400, seed=37, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, laws="full", periodic=(False, False),
                   walls=())
    sc.state.x[:] += np.asarray(offset)
    g, o = make_pair(sc)
    g.compute_forces()
    o.compute_forces()
    assert assert_contacts_equal(g, o) > 100
    assert_forces_close(g, o)
    g, o = make_pair(reseeded(sc, o.download()), o.contacts())
    g.step(1)
    o.step(1)
    assert_contacts_equal(g, o)
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
    g.step(20)
    o.step(20)
    assert_contacts_close(g, o)


def test_removal_reorients_periodic_image_interactions_only():
    """The one deviation after removals: the reference orients an interaction
    found through a periodic image by handler index (the lower index is P,
    InteractionHandler.cc:327-368), the engine by id.  The highest-index
    sphere with an image contact is made the last one and then moves to index
    0 (removeObject's swap with the last), so its image interactions change
    orientation in the reference and not in the engine.  After the step every
    stored interaction of the engine is the oracle's, either as stored or
    reversed -- P and I exchanged, sliding and torsion springs negated,
    rolling not (reverseHistory), to round-off -- with at least one
    reversed; forces and positions agree."""
    sc = S.mixture(n=400, seed=53, lx=7.0, ly=7.0, lz=7.0, inflate=1.12, laws="full", periodic=(True, True),
                   walls=("bottom",))
    ref = oracle_lib.engine()
    S.apply(ref, sc)
    ref.compute_forces()
    ref.step(20)
    st, ct = ref.download(), ref.contacts()
    x = st["x"]
    index = {int(v): k for k, v in enumerate(sc.state.id)}
    pair = (ct.i_id & dem.DEM_WALL_KEY) == 0
    pid, iid = ct.p_id[pair], ct.i_id[pair]
    dx = np.abs(x[[index[int(p)] for p in pid]] - x[[index[int(i)] for i in iid]])
    image = (dx[:, 0] > 3.5) | (dx[:, 1] > 3.5)
    involved = {index[int(p)] for p in pid[image]} | {index[int(i)] for i in iid[image]}
    h = max(involved)
    assert 0 < h < 399
    gone = list(range(399, h, -1)) + [0]  # pop everything above h, then h (now last) takes index 0
    g, o = make_pair(reseeded(sc, st), ct)
    for e in (g, o):
        e.set_time(20 * sc.dt, 20)
        e.step_begin()
        e.remove_particles(gone)
        e.step_end()
    assert g.n == o.n == 400 - len(gone)
    cg, co = g.contacts(), o.contacts()
    oracle = {(int(p), int(i)): s for p, i, s in zip(co.p_id, co.i_id, co.spring)}
    flip = np.array([-1, -1, -1, 1, 1, 1, -1, -1, -1], dtype=float)
    scale = np.abs(co.spring).max()
    reversed_pairs = 0
    assert cg.n == co.n
    for p, i, s in zip(cg.p_id, cg.i_id, cg.spring):
        key = (int(p), int(i))
        if key in oracle:
            np.testing.assert_array_equal(s, oracle[key])
        else:
            # evaluated from the other side: the contact point, hence the
            # relative velocity, rounds differently
            np.testing.assert_allclose(s, flip * oracle[(int(i), int(p))], rtol=1e-10, atol=1e-12 * scale)
            reversed_pairs += 1
    assert reversed_pairs > 0
    a, b = assert_forces_close(g, o)
    np.testing.assert_array_equal(a["x"], b["x"])
