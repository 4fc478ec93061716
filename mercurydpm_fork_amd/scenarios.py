"""Simulation set-ups used by the tests and the benchmark.

Each builder returns a ``Scenario``: species table, walls, periodic boundaries,
gravity, time step and the particle state in handler order, exactly as the
corresponding MercuryDPM driver would hand them to the engine.  Builders named
after a reference test restate that driver's set-up (file cited); the large
workloads are generated from a seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import dem
from .species import (PI, Species, axisymmetric_wall, cylindrical_wall, infinite_wall, periodic_boundary,
                      species_table, sphere_inv_mass, sphere_inv_mass_array)


@dataclass
class Scenario:
    name: str
    species: List[Species]
    walls: list
    periodic: list
    gravity: tuple
    dt: float
    state: dem.State
    t_max: Optional[float] = None
    drag: float = 0.0
    rotation: Optional[bool] = None
    mixed_overrides: Optional[dict] = None
    info: dict = field(default_factory=dict)
    # per-species wall groups (first wall, number of walls), or None: every
    # particle meets every wall
    wall_groups: Optional[tuple] = None
    # a ready species table (dem.PairParams array) in place of table()
    pair_table: Optional[object] = None

    @property
    def n_species(self) -> int:
        return len(self.species)

    def table(self):
        if self.pair_table is not None:
            return self.pair_table
        return species_table(self.species, self.mixed_overrides)

    def uses_rotation(self) -> bool:
        if self.rotation is not None:
            return self.rotation
        return any(s.uses_rotation() for s in self.species)

    def n_steps(self) -> int:
        """Number of computeOneTimeStep calls of DPMBase::solve (DPMBase.cc:
        4279-4295): time_ accumulates dt until it reaches timeMax."""
        t = 0.0
        n = 0
        while t < self.t_max:
            t += self.dt
            n += 1
        return n


def apply(engine: dem.Engine, sc: Scenario, contacts: Optional[dem.Contacts] = None):
    engine.set_species(sc.table(), sc.n_species)
    engine.set_walls(sc.walls)
    if sc.wall_groups is not None:
        engine.set_wall_groups(*sc.wall_groups)
    engine.set_periodic(sc.periodic)
    engine.set_gravity(sc.gravity)
    engine.set_background_drag(sc.drag)
    engine.set_timestep(sc.dt)
    engine.set_rotation(sc.uses_rotation())
    engine.set_time(0.0, 0)
    engine.upload(sc.state)
    if contacts is not None and contacts.n:
        engine.upload_contacts(contacts)


def make_state(pos, vel, radius, species_idx, species: List[Species], fixed=None, dims: int = 3,
               omega=None, ids=None) -> dem.State:
    n = len(radius)
    pos = np.asarray(pos, dtype=np.float64).reshape(n, 3)
    vel = np.asarray(vel, dtype=np.float64).reshape(n, 3)
    radius = np.asarray(radius, dtype=np.float64)
    species_idx = np.asarray(species_idx, dtype=np.int32)
    fixed = np.zeros(n, bool) if fixed is None else np.asarray(fixed, bool)
    dens = np.array([sp.density for sp in species], dtype=np.float64)[species_idx]
    im, ii = sphere_inv_mass_array(radius, dens, dims)
    inv_mass = np.where(fixed, 0.0, im)
    inv_i = np.zeros((n, 6))
    for c in (0, 3, 5):
        inv_i[:, c] = np.where(fixed, 0.0, ii)
    vel = vel.copy()
    vel[fixed] = 0.0
    w = np.zeros((n, 3)) if omega is None else np.asarray(omega, dtype=np.float64).reshape(n, 3).copy()
    w[fixed] = 0.0
    q = np.zeros((n, 4))
    q[:, 0] = 1.0
    ids = np.arange(n, dtype=np.uint32) if ids is None else np.asarray(ids, np.uint32)
    return dem.State(id=ids, species=species_idx, radius=radius, inv_mass=inv_mass, inv_inertia=inv_i,
                     x=pos, v=vel, w=w, q=q)


# --------------------------------------------------------------------------
# Set-ups of the reference's unit tests (Drivers/UnitTests)
# --------------------------------------------------------------------------

def tangential_spring_unit_test() -> Scenario:
    """Drivers/UnitTests/TangentialSpringUnitTest.cpp:31-93."""
    s = Species(density=6. / PI, friction_law=dem.FRICTION_SLIDING)
    s.set_collision_time_and_restitution(1e-3, .2, 1. / 8)
    s.sliding_stiffness = s.stiffness * 2 / 7
    s.set_sliding_friction(1e20)
    st = make_state(pos=[[0.5, 0.5, 0.25], [0.5, 0.5, 0.75]], vel=[[0, 0, 0], [0.1, 0.0, 0.0]],
                    radius=[0.25, 0.25], species_idx=[0, 0], species=[s], fixed=[True, False])
    return Scenario("TangentialSpringUnitTest", [s], [], [], (0.0, 0.0, -1.0), 2e-5, st, t_max=2.5)


def periodic_walls_unit_test() -> Scenario:
    """Drivers/UnitTests/PeriodicWallsUnitTest.cpp:31-148 (Mercury2D: discs of
    particle dimension 2; computeExternalForces is overridden to do nothing,
    so gravity does not act)."""
    s = Species(density=2000)
    s.stiffness = 10000
    s.dissipation = 1
    pos = [(0e-3, 2e-3, 0.0), (1e-3, 2e-3, 0.0), (10e-3, 3e-3, 0.0), (1e-3, 3e-3, 0.0), (0e-3, 4e-3, 0.0),
           (9e-3, 4e-3, 0.0), (10e-3, 5e-3, 0.0), (9e-3, 5e-3, 0.0), (1e-3, 6e-3, 0.0), (10e-3, 6e-3, 0.0),
           (10e-3, 10e-3, 0.0), (9e-3, 9e-3, 0.0), (0e-3, 0e-3, 1e-3), (1e-3, 1e-3, 1e-3)]
    vel = [(0, 0, 0), (-1e-2, 0, 0), (0, 0, 0), (-1e-2, 0, 0), (0, 0, 0), (1e-2, 0, 0), (0, 0, 0), (1e-2, 0, 0),
           (0, 0, 0), (1e-2, 0, 0), (0, 0, 0), (1e-2, 1e-2, 0), (0, 0, 0), (-1e-2, -1e-2, 0)]
    rad = [0.45e-3 if k % 2 == 0 else 0.35e-3 for k in range(14)]
    st = make_state(pos, vel, rad, [0] * 14, [s], dims=2)
    per = [periodic_boundary(0, 0.0, 0.01), periodic_boundary(1, 0.0, 0.01)]
    return Scenario("PeriodicWallsUnitTest", [s], [], per, (0.0, 0.0, 0.0), 1.1e-05, st, t_max=0.11)


def periodic_walls_with_friction_unit_test() -> Scenario:
    """Drivers/UnitTests/PeriodicWallsWithSlidingFrictionUnitTest.cpp:38-123
    (Mercury2D, so discs of particle dimension 2): eight discs of radius 0.1
    and density 2500 in a unit box periodic in x and y, no gravity.
    LinearViscoelasticFrictionSpecies with k = 1e5, no dissipation, sliding
    stiffness 2/7 k, rolling and torsion stiffness 2/5 k, all three friction
    coefficients 0.5.  Three colliding pairs: (4, 5) meets directly, (2, 3)
    starts direct and ends across y = 0, (6, 7) starts across x = 0, meets
    directly and parts across the boundary again; 2, 4 and 6 spin at
    (0.3, 0.6, 0.9).  Particles 0 and 1 sit apart and are removed mid-run
    (`removals`: handler index removed in actionsBeforeTimeStep once the
    time exceeds the given value, cpp:127-181), which moves 7 to index 1 and
    then 6 to index 0 with their springs.  The reference then requires the
    three pairs to end in the same state up to their shifts
    (`PERIODIC_FRICTION_SHIFTS`, cpp:195-249, tolerance 1e-10)."""
    k = 1e5
    s = Species(density=2500, friction_law=dem.FRICTION_FULL)
    s.stiffness = k
    s.sliding_stiffness = 2.0 / 7.0 * k
    s.set_sliding_friction(0.5)
    s.rolling_stiffness = 2.0 / 5.0 * k
    s.set_rolling_friction(0.5)
    s.torsion_stiffness = 2.0 / 5.0 * k
    s.set_torsion_friction(0.5)
    pos = [(0.1, 0.1, 0.0), (0.3, 0.3, 0.0), (0.6, 0.041, 0.0), (0.4, 0.001, 0.0), (0.6, 0.82, 0.0),
           (0.4, 0.78, 0.0), (0.02, 0.42, 0.0), (0.82, 0.38, 0.0)]
    vel = [(0, 0, 0), (0, 0, 0), (-1, 0, 0), (1, 0, 0), (-1, 0, 0), (1, 0, 0), (-1, 0, 0), (1, 0, 0)]
    spin = (0.3, 0.6, 0.9)
    omega = [(0, 0, 0), (0, 0, 0), spin, (0, 0, 0), spin, (0, 0, 0), spin, (0, 0, 0)]
    st = make_state(pos, vel, [0.1] * 8, [0] * 8, [s], dims=2, omega=omega)
    per = [periodic_boundary(0, 0.0, 1.0), periodic_boundary(1, 0.0, 1.0)]
    sc = Scenario("PeriodicWallsWithSlidingFrictionUnitTest", [s], [], per, (0.0, 0.0, 0.0), 2.0e-4, st,
                  t_max=8e-2)
    sc.info["removals"] = [(3.5e-2, 1), (4.5e-2, 0)]
    return sc


# (reference index, compared index, shift) after the run:
# PeriodicWallsWithSlidingFrictionUnitTest.cpp:195-249
PERIODIC_FRICTION_SHIFTS = [(4, 0, (-0.58, -0.4, 0.0)), (4, 2, (0.0, -0.779, 0.0)),
                            (5, 1, (1.0 - 0.58, -0.4, 0.0)), (5, 3, (0.0, 1.0 - 0.779, 0.0))]


PERIODIC_WALLS_GOLD = [  # PeriodicWallsUnitTest.cpp:158-361, tolerance 1e-10
    ((0.00950076062215577, 0.002, 0), (-0.005560409816604, 0, 0)),
    ((0.000725163257251641, 0.002, 0), (-0.000808302139899, 0, 0)),
    ((0.00950076062215577, 0.003, 0), (-0.005560409816604, 0, 0)),
    ((0.000725163257251641, 0.003, 0), (-0.000808302139899, 0, 0)),
    ((0.0004992393778432442, 0.004, 0), (0.00556040981661, 0, 0)),
    ((0.00927483674274562, 0.004, 0), (0.0008083021398896, 0, 0)),
    ((0.000499239377843248, 0.005, 0), (0.00556040981661, 0, 0)),
    ((0.00927483674274562, 0.005, 0), (0.0008083021398892, 0, 0)),
    ((0.00149923937784423, 0.006, 0), (0.005560409816606, 0, 0)),
    ((0.000274836742748357, 0.006, 0), (0.0008083021398959, 0, 0)),
    ((0.000368955529803987, 0.000368955529803987, 0), (0.005560480643586, 0.005560480643586, 0)),
    ((0.0094902039201126, 0.0094902039201126, 0), (0.0008081850585628, 0.0008081850585628, 0)),
    ((0.00963104447019111, 0.00963104447019111, 0.001), (-0.005560480643562, -0.005560480643562, 0)),
    ((0.00050979607988006, 0.00050979607988006, 0.001), (-0.0008081850586013, -0.0008081850586013, 0)),
]

TANGENTIAL_SPRING_GOLD = dict(  # TangentialSpringUnitTest.cpp:89-101, tolerance 1e-7
    x0=(0.5, 0.5, 0.25), v0=(0.0, 0.0, 0.0),
    x1=(0.961524052956078, 0.5, 0.450153103106219), v1=(0.325869890236916, 0, -0.623251003973752),
    w1=(0.0, 2.11896618998424, 0.0))


def inclined_plane_unit_test() -> Scenario:
    """Drivers/UnitTests/InclinedPlaneUnitTest.cpp:33-83 (expects v_x(1) =
    1.81167 +- 1e-4, v_y = 0.00234107 +- 1e-4)."""
    s = Species(density=2000, friction_law=dem.FRICTION_SLIDING)
    s.stiffness = 10000
    s.dissipation = 0
    s.sliding_stiffness = 10000
    s.sliding_dissipation = 0
    s.set_sliding_friction(0.5)
    theta = 15.0 * PI / 180.0
    g = (math.sin(theta) * 9.8, -math.cos(theta) * 9.8, 0.0 * 9.8)
    dt = 0.02 * s.collision_time(s.mass_from_radius(0.005))
    w = infinite_wall((0, -1, 0), (0, 0.0, 0))
    r = 0.005
    st = make_state([((0.0 + 0.1) / 2.0, 0.0 + r, 0)], [(0, 0, 0)], [r], [0], [s])
    return Scenario("InclinedPlaneUnitTest", [s], [w], [], g, dt, st, t_max=1.0)


def hertz_restitution_unit_test() -> Scenario:
    """Drivers/UnitTests/HertzContactRestitutionUnitTest.cpp:31-67 (expects
    final velocities -0.8 / +0.8 within 1e-2)."""
    s = Species(density=1.0, friction_law=dem.FRICTION_MINDLIN)
    s.set_elastic_modulus_and_restitution(20000, 0.8)
    st = make_state([(0, 0, 0), (3, 0, 0)], [(1, 0, 0), (-1, 0, 0)], [1.0, 1.0], [0, 0], [s])
    dt = s.hertz_collision_time(1, 1, PI / 6) / 50
    return Scenario("HertzContactRestitutionUnitTest", [s], [], [], (0.0, 0.0, 0.0), dt, st, t_max=1.5)


def oblique_impact_self_test() -> Scenario:
    """Drivers/SelfTests/Interactions/ObliqueImpactSelfTest.cpp:31-78: a
    HertzianViscoelasticMindlin sphere (E* = 1e5, nu = 0.3, mu = 0.1, no
    dissipation) hits a flat wall at 10 m/s tangential, 1 m/s normal; the
    reference's .fstat (SelfTestData/ObliqueImpactSelfTest.fstat) holds the
    contact every 500 steps."""
    s = Species(density=6.0 / PI, friction_law=dem.FRICTION_MINDLIN)
    s.set_elastic_modulus_and_poisson_ratio(1e5, 0.3)
    s.set_sliding_friction(0.1)
    r = 0.005
    w = infinite_wall((0, -1, 0), (0, 0.0, 0))
    st = make_state([(0.0 + r, r + 0.000005, 0.0)], [(10.0, -1.0, 0.0)], [r], [0], [s])
    return Scenario("ObliqueImpactSelfTest", [s], [w], [], (0.0, 0.0, 0.0), 0.25e-8, st, t_max=0.0005)


OBLIQUE_IMPACT_SAVE_COUNT = 500


def squeezed_sphere_self_test(kind: str) -> Scenario:
    """The rolling and torsion friction self tests
    (Drivers/SelfTests/Interactions/RollingFrictionSelfTest.cpp:39-81 and
    TorsionFrictionSelfTest.cpp:39-81): a free unit-mass sphere (r = 0.5,
    density 6/pi) squeezed between two fixed spheres at z = -+0.995 spins at
    omega = e_y (kind "rolling") or e_z ("torsion").  LinearViscoelastic-
    FrictionSpecies with t_c = 0.02 and normal and tangential restitution
    0.7 (mass argument 1), rolling (torsion) stiffness and dissipation 0.4 of
    the normal ones and friction coefficient 0.1; no sliding friction, no
    gravity, dt = 0.02 t_c, 500 steps.  The reference prints, after each
    step's force pass, the torque (y for rolling, z for torsion) and the
    force z of the upper fixed sphere (SelfTestData/*FrictionSelfTest.torque);
    `torque_axis` names the component."""
    if kind not in ("rolling", "torsion"):
        raise ValueError("kind is 'rolling' or 'torsion'")
    tc = 2e-2
    s = Species(density=6.0 / PI, friction_law=dem.FRICTION_FULL)
    s.set_collision_time_and_normal_and_tangential_restitution(tc, 0.7, 0.7, 1)
    if kind == "rolling":
        s.rolling_stiffness = 0.4 * s.stiffness
        s.rolling_dissipation = 0.4 * s.dissipation
        s.set_rolling_friction(0.1)
        spin = (0.0, 1.0, 0.0)
    else:
        s.torsion_stiffness = 0.4 * s.stiffness
        s.torsion_dissipation = 0.4 * s.dissipation
        s.set_torsion_friction(0.1)
        spin = (0.0, 0.0, 1.0)
    r = 0.5
    st = make_state([(0, 0, 0), (0, 0, -1.99 * r), (0, 0, 1.99 * r)], [(0, 0, 0)] * 3, [r] * 3, [0] * 3, [s],
                    fixed=[False, True, True], omega=[spin, (0, 0, 0), (0, 0, 0)])
    sc = Scenario(f"{kind.capitalize()}FrictionSelfTest", [s], [], [], (0.0, 0.0, 0.0), 0.02 * tc, st,
                  t_max=10 * tc)
    sc.info["torque_axis"] = 1 if kind == "rolling" else 2
    return sc


# Drivers/UnitTests/MindlinSelfTest.cpp:36-58 with the set-ups of
# Drivers/UnitTests/ContactLawTestHelpers.h.  The particles' motion is
# prescribed in actionsBeforeTimeStep; `prescribe(t)` returns what those hooks
# set at time t as (x, v, w, q) arrays in handler order (None = not set).
MINDLIN_DISPLACEMENT = 0.015
MINDLIN_TANGENTIAL_DISPLACEMENT = 0.0025
MINDLIN_VELOCITY = 0.01
MINDLIN_RADIUS = 0.25


def mindlin_self_test_species() -> Species:
    """MindlinSelfTest.cpp:44-48."""
    s = Species(density=6. / PI, friction_law=dem.FRICTION_MINDLIN)
    s.set_elastic_modulus_and_poisson_ratio(1e5, 0.3)
    s.dissipation = 2.0
    s.set_sliding_friction(0.1)
    return s


def mindlin_loading_test():
    """loadingTest (ContactLawTestHelpers.h:21-84): a sphere pressed into a
    flat wall and pulled back out; saved every step."""
    d, vel, r = MINDLIN_DISPLACEMENT, MINDLIN_VELOCITY, MINDLIN_RADIUS
    s = mindlin_self_test_species()
    t_max = 2.0 * d / vel
    w = infinite_wall((0, 0, -1), (0.0, 0.0, 0.0))
    st = make_state([(0, 0, r)], [(0, 0, 0)], [r], [0], [s])
    sc = Scenario("MindlinSelfTestLoading", [s], [w], [], (0.0, 0.0, 0.0), 2e-3 * t_max, st, t_max=t_max)

    def prescribe(t):
        if t <= d / vel:
            v, x = (0, 0, vel), (0, 0, r - vel * t)
        else:
            v, x = (0, 0, -vel), (0, 0, r - d - d + vel * t)
        return np.array([x], float), np.array([v], float), np.zeros((1, 3)), None

    return sc, prescribe, 1


def mindlin_normal_and_tangential_loading_test():
    """normalAndTangentialLoadingTest (ContactLawTestHelpers.h:92-162): a
    sphere pressed into a wall and moved back and forth tangentially."""
    d, td, vel, r = MINDLIN_DISPLACEMENT, MINDLIN_TANGENTIAL_DISPLACEMENT, MINDLIN_VELOCITY, MINDLIN_RADIUS
    s = mindlin_self_test_species()
    t_max = 4.0 * td / vel
    w = infinite_wall((0, 0, -1), (0.0, 0.0, 0.0))
    st = make_state([(0, 0, r - d)], [(0, 0, 0)], [r], [0], [s])
    sc = Scenario("MindlinSelfTestNormalAndTangentialLoading", [s], [w], [], (0.0, 0.0, 0.0), 4e-4 * t_max, st,
                  t_max=t_max)

    def prescribe(t):
        move_right = int(t / (2.0 * td / vel) + 0.5) % 2 == 0
        if move_right:
            v, x = (-vel, 0, 0), (td - vel * t, 0, r - d)
        else:
            v, x = (vel, 0, 0), (-2 * td + vel * t, 0, r - d)
        return np.array([x], float), np.array([v], float), np.zeros((1, 3)), None

    return sc, prescribe, 1


def mindlin_frame_independence_test():
    """objectivenessTest (ContactLawTestHelpers.h:170-249): two spheres
    loaded normally and tangentially, then the pair is rotated rigidly; saved
    every 20 steps."""
    d, td, vel, r = MINDLIN_DISPLACEMENT, MINDLIN_TANGENTIAL_DISPLACEMENT, MINDLIN_VELOCITY, MINDLIN_RADIUS
    s = mindlin_self_test_species()
    t_max = (td + 0.5 * PI * r) / vel
    st = make_state([(0, r - d, 0), (0, -r + d, 0)], [(0, 0, 0), (0, 0, 0)], [r, r], [0, 0], [s])
    sc = Scenario("MindlinSelfTestFrameIndependence", [s], [], [], (0.0, 0.0, 0.0), 1e-4 * t_max, st, t_max=t_max)

    def prescribe(t):
        if t <= td / vel:
            x = [(-td + vel * t, r - d, 0), (td - vel * t, -r + d, 0)]
            v = [(vel, 0, 0), (-vel, 0, 0)]
            return np.array(x, float), np.array(v, float), np.zeros((2, 3)), None
        angle = vel / (r - d) * (t - td / vel)
        sn, cs = math.sin(angle), math.cos(angle)
        k = vel / (r - d)
        wp = (0 * k, 0 * k, -1 * k)
        vp = (vel * cs, vel * -sn, vel * 0)
        xp = ((r - d) * sn, (r - d) * cs, (r - d) * 0)
        neg = lambda a: tuple(-c for c in a)
        x = np.array([xp, neg(xp)], float)
        v = np.array([vp, neg(vp)], float)
        w = np.array([wp, neg(wp)], float)
        q = np.array([(1.0, 0.0, 0.0, 0.0), (-1.0, -0.0, -0.0, -0.0)])
        return x, v, w, q

    return sc, prescribe, 20


def load_fstat(path):
    """Rows of a .fstat fixture as a float array (columns as written by
    BaseInteraction::writeToFStat, BaseInteraction.cc:487-531)."""
    return np.loadtxt(path, comments="#", ndmin=2)


def cg_basic_self_test():
    """Drivers/SelfTests/CG/CGBasicSelfTest.cpp:70-108: two spheres stacked on
    a fixed one under gravity (-0.1, 0, -1), linear viscoelastic, until
    t = 0.21; its CG<Z, Gauss> output (n = 1000, width 0.5 / 3.01, over the
    domain below) is SelfTestData/CGBasicSelfTest.1.stat.  Returns the scenario
    and the CG settings (n, width, min, max)."""
    s = Species(density=6.0 / PI)
    s.set_collision_time_and_restitution(.01, .1, 1.)
    r = 0.5 * math.sqrt(1.01)
    st = make_state([(0.5, 0.5, 0.5), (0.6, 0.5, 1.5), (0.4, 0.5, -0.5)], [(0, 0, 0)] * 3, [r, r, r], [0, 0, 0], [s],
                    fixed=[False, False, True])
    sc = Scenario("CGBasicSelfTest", [s], [], [], (-0.1, 0.0, -1.0), s.collision_time(1.0) / 50.0, st, t_max=0.21)
    return sc, dict(n=1000, width=0.5 / 3.01, lo=(-0.1, 0.0, -0.6), hi=(1.1, 1.0, 2.1))


# --------------------------------------------------------------------------
# Seeded workloads
# --------------------------------------------------------------------------

def rsa_positions(n, lo, hi, radii, rng, max_tries=2000, periodic=(False, False, False)):
    """Random sequential addition of non-overlapping spheres in the box
    [lo, hi].  Along non-periodic axes centres keep a radius from the faces;
    along periodic axes they fill [lo, hi) and overlaps are tested with the
    minimum image."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    L = hi - lo
    per = np.asarray(periodic, bool)
    h = 2.0 * float(np.max(radii))
    ncell = np.where(per, np.maximum(1, np.floor(L / h)).astype(int), 0)
    grid = {}
    pos = np.empty((n, 3))

    def cell(p):
        c = np.floor((p - lo) / h).astype(int)
        return np.where(per, c % np.maximum(ncell, 1), c)

    for k in range(n):
        r = float(radii[k])
        for _ in range(max_tries):
            margin = np.where(per, 0.0, r)
            p = lo + margin + rng.random(3) * (L - 2 * margin)
            c = cell(p)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        t = c + (dx, dy, dz)
                        t = np.where(per, t % np.maximum(ncell, 1), t)
                        for j in grid.get(tuple(t), ()):
                            dv = pos[j] - p
                            dv = np.where(per & (dv > 0.5 * L), dv - L, dv)
                            dv = np.where(per & (dv < -0.5 * L), dv + L, dv)
                            if dv @ dv < (radii[j] + r) ** 2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[k] = p
                grid.setdefault(tuple(c), []).append(k)
                break
        else:
            raise RuntimeError("random sequential addition did not converge")
    return pos


def box_lvf(n=5000, seed_pos=1, seed_vel=2, lx=20.0, ly=20.0, lz=40.0) -> Scenario:
    """5000 monodisperse spheres (d = 1, m = 1) in a 20 x 20 x 40 box, periodic
    in x and y, flat InfiniteWall at z = 0, gravity (0, 0, -1).
    LinearViscoelasticFriction: e = 0.88 at t_c = 0.005 (mass 1), sliding
    friction 0.5 with k_t = 2/7 k and gamma_t = 2/7 gamma (the convention of
    PeriodicWallsWithSlidingFrictionUnitTest), dt = t_c / 50."""
    s = Species(density=6. / PI, friction_law=dem.FRICTION_FULL)
    tc = 0.005
    s.set_collision_time_and_restitution(tc, 0.88, 1.0)
    s.sliding_stiffness = 2. / 7. * s.stiffness
    s.sliding_dissipation = 2. / 7. * s.dissipation
    s.set_sliding_friction(0.5)
    r = np.full(n, 0.5)
    pos = rsa_positions(n, (0.0, 0.0, 0.0), (lx, ly, lz), r, np.random.default_rng(seed_pos))
    vel = np.random.default_rng(seed_vel).uniform(-0.1, 0.1, (n, 3))
    st = make_state(pos, vel, r, np.zeros(n, np.int32), [s])
    wall = infinite_wall((0, 0, -1), (0, 0, 0))
    per = [periodic_boundary(0, 0.0, lx), periodic_boundary(1, 0.0, ly)]
    return Scenario(f"box_lvf_{n}", [s], [wall], per, (0.0, 0.0, -1.0), tc / 50, st, t_max=20000 * tc / 50,
                    info=dict(n_flowing=n))


def hertz_mindlin_species(restitution=0.8, mu=0.5, tc=0.005, density=6. / PI, poisson=0.3):
    """HertzianViscoelasticMindlin: E* chosen so that the collision time of
    two unit spheres at relative velocity 1 is t_c
    (HertzianViscoelasticNormalSpecies::getCollisionTime,
    HertzianViscoelasticNormalSpecies.cc:290-294), G* = E*/(2(1+nu)), sliding
    dissipation equal to the normal one."""
    s = Species(density=density, friction_law=dem.FRICTION_MINDLIN)
    # 2.214 ((rho/E)^2 / v)^0.2 d = tc  ->  E = rho / (tc / 2.214)^2.5
    E = density / (tc / 2.214) ** 2.5
    s.set_elastic_modulus_and_restitution(E, restitution)
    s.shear_modulus = E / (2.0 * (1.0 + poisson))
    s.sliding_dissipation = s.dissipation
    s.set_sliding_friction(mu)
    s.info_tc = tc
    return s


def lattice_packing(n, lx, ly, z0, a, radii, rng):
    """Jittered simple-cubic lattice: every sphere stays inside its own cell of
    side a >= its diameter, so no two overlap."""
    nx = int(math.floor(lx / a))
    ny = int(math.floor(ly / a))
    ax = lx / nx
    ay = ly / ny
    per_layer = nx * ny
    k = np.arange(n)
    iz = k // per_layer
    rem = k % per_layer
    iy = rem // nx
    ix = rem % nx
    r = np.asarray(radii)
    jx = (rng.random(n) - 0.5) * np.maximum(ax - 2 * r, 0)
    jy = (rng.random(n) - 0.5) * np.maximum(ay - 2 * r, 0)
    jz = (rng.random(n) - 0.5) * np.maximum(a - 2 * r, 0)
    x = (ix + 0.5) * ax + jx
    y = (iy + 0.5) * ay + jy
    z = z0 + (iz + 0.5) * a + jz
    return np.stack([x, y, z], axis=1)


def rocksalt_packing(d, lx, ly, z0, rng, jitter=0.004):
    """Non-overlapping polydisperse packing on a rock-salt lattice: the
    smaller half of the diameters sits on one sublattice, the larger half on
    the other, so nearest neighbours are always a small and a large sphere and
    the spacing only needs to exceed (max small + max large) / 2.  Returns
    centres in the order of the diameters given."""
    n = len(d)
    order = np.argsort(d, kind="stable")
    small = order[: n // 2]
    large = order[n // 2:]
    a = 0.5 * (float(d[small].max()) + float(d[large].max())) + 2.0 * math.sqrt(3.0) * jitter + 1e-3
    nx = int(math.floor(lx / a)) // 2 * 2
    ny = int(math.floor(ly / a)) // 2 * 2
    ax, ay = lx / nx, ly / ny
    per_layer = nx * ny
    layers = int(math.ceil(2 * max(len(small), len(large)) / per_layer)) + 1
    iz, iy, ix = np.meshgrid(np.arange(layers), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    parity = (ix + iy + iz) % 2
    sites_a = np.flatnonzero(parity == 0)[: len(small)]
    sites_b = np.flatnonzero(parity == 1)[: len(large)]
    pos = np.empty((n, 3))
    for idx, sites in ((small, sites_a), (large, sites_b)):
        perm = rng.permutation(len(idx))
        st = sites[perm]
        pos[idx, 0] = (ix[st] + 0.5) * ax
        pos[idx, 1] = (iy[st] + 0.5) * ay
        pos[idx, 2] = z0 + (iz[st] + 0.5) * a
    pos += rng.uniform(-jitter, jitter, pos.shape)
    return pos


def chute_hvm(n_flow=2_000_000, seed=3, seed_base=4, angle_deg=24.0, l=None, depth=30.0,
              base_d=1.0, d_range=(0.5, 1.5), species=None) -> Scenario:
    """Polydisperse periodic chute: n_flow spheres with d ~ U[d_range]
    (default [0.5, 1.5]), tilted by angle_deg (gravity (sin, 0, -cos), |g| = 1,
    the SilbertPeriodic convention of Chute::setChuteAngleAndMagnitudeOfGravity,
    Chute.cc:789-805), periodic in x and y, a rough base of fixed spheres
    (d = base_d, square lattice with random heights) over an InfiniteWall.
    HertzianViscoelasticMindlin with e = 0.8 and mu = 0.5 (or the given
    species, a hertz_mindlin_species, with its collision time), dt = t_c / 50.
    The flowing spheres start on a rock-salt lattice (loose, non-overlapping);
    the footprint is chosen so that the bed would be `depth` diameters deep at
    packing fraction 0.55."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(d_range[0], d_range[1], n_flow)
    r_flow = 0.5 * d
    if l is None:
        vol = float(np.sum(PI / 6 * d ** 3))
        l = math.sqrt(vol / 0.55 / depth)
    s = hertz_mindlin_species() if species is None else species
    tc = s.info_tc if hasattr(s, "info_tc") else 0.005
    # rough base
    rngb = np.random.default_rng(seed_base)
    nb = int(math.floor(l / base_d))
    ab = l / nb
    bx, by = np.meshgrid((np.arange(nb) + 0.5) * ab, (np.arange(nb) + 0.5) * ab, indexing="ij")
    bx = bx.ravel()
    by = by.ravel()
    bz = rngb.uniform(-0.25, 0.25, bx.size) * base_d
    base = np.stack([bx, by, bz], axis=1)
    r_base = np.full(bx.size, 0.5 * base_d)
    flow = rocksalt_packing(d, l, l, 1.0 * base_d, rng)
    pos = np.concatenate([base, flow])
    rad = np.concatenate([r_base, r_flow])
    fixed = np.concatenate([np.ones(bx.size, bool), np.zeros(n_flow, bool)])
    vel = np.zeros_like(pos)
    st = make_state(pos, vel, rad, np.zeros(rad.size, np.int32), [s], fixed=fixed)
    th = angle_deg * PI / 180.0
    g = (math.sin(th), 0.0, -math.cos(th))
    wall = infinite_wall((0, 0, -1), (0, 0, -0.5 * base_d))
    per = [periodic_boundary(0, 0.0, l), periodic_boundary(1, 0.0, l)]
    return Scenario(f"chute_hvm_{n_flow}", [s], [wall], per, g, tc / 50, st,
                    info=dict(n_flowing=n_flow, n_base=int(bx.size), footprint=l))


def settle_chute(engine: dem.Engine, sc: Scenario, steps: int = 80_000, chunk: int = 10_000, log=None):
    """Turn the loose lattice of chute_hvm into a random loose packing: the
    spheres fall onto the rough base under |g| straight down (no tilt, no
    drag) for `steps` steps, then the chute gravity is restored.  From the
    lattice 30 diameters deep the bed is at rest after about 70,000 steps.
    Returns the engine's stats at the end."""
    gmag = math.sqrt(sum(c * c for c in sc.gravity))
    engine.set_gravity((0.0, 0.0, -gmag))
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        engine.step(k)
        done += k
        if log is not None:
            st = engine.stats()
            log(f"settle {done}/{steps} steps (|g| {gmag:g} straight down): C/N {st.n_contacts / sc.state.n:.3f}")
    engine.set_gravity(sc.gravity)
    return engine.stats()


def mixture(n=400, seed=7, lx=8.0, ly=8.0, lz=8.0, periodic=(True, True), walls=("bottom",), laws="lvf",
            fixed_frac=0.05, vmax=0.5, radii=(0.2, 0.45), omega_max=1.0, n_species=2,
            inflate=1.06) -> Scenario:
    """Small random mixture for engine-versus-oracle parity: polydisperse
    spheres of several species with mixed-pair parameters, random velocities
    and spins, a few fixed spheres, optional periodic x/y, a floor and a
    cylinder."""
    rng = np.random.default_rng(seed)
    spec = []
    for k in range(n_species):
        if laws == "lvf":
            s = Species(density=1.0 + 0.5 * k, friction_law=dem.FRICTION_SLIDING)
            s.set_collision_time_and_restitution(0.01 * (1 + 0.3 * k), 0.7 + 0.1 * k, 0.5)
            s.sliding_stiffness = 2. / 7. * s.stiffness
            s.sliding_dissipation = 2. / 7. * s.dissipation
            s.set_sliding_friction(0.3 + 0.2 * k)
        elif laws == "full":
            s = Species(density=1.0 + 0.5 * k, friction_law=dem.FRICTION_FULL)
            s.set_collision_time_and_restitution(0.01 * (1 + 0.3 * k), 0.7 + 0.1 * k, 0.5)
            s.sliding_stiffness = 2. / 7. * s.stiffness
            s.sliding_dissipation = 2. / 7. * s.dissipation
            s.set_sliding_friction(0.3 + 0.2 * k)
            s.rolling_stiffness = 2. / 5. * s.stiffness
            s.rolling_dissipation = 2. / 5. * s.dissipation
            s.set_rolling_friction(0.05 + 0.05 * k)
            s.torsion_stiffness = 2. / 5. * s.stiffness
            s.torsion_dissipation = 2. / 5. * s.dissipation
            s.set_torsion_friction(0.02 + 0.02 * k)
        else:
            s = hertz_mindlin_species(restitution=0.7 + 0.1 * k, mu=0.3 + 0.2 * k, tc=0.01,
                                      density=1.0 + 0.5 * k)
        spec.append(s)
    r = rng.uniform(radii[0], radii[1], n)
    pos = rsa_positions(n, (0, 0, 0), (lx, ly, lz), r, rng, periodic=(periodic[0], periodic[1], False))
    # grow the spheres after placement so that many pairs (also across the
    # periodic faces) and sphere-wall pairs start in contact
    r = r * inflate
    vel = rng.uniform(-vmax, vmax, (n, 3))
    om = rng.uniform(-omega_max, omega_max, (n, 3))
    fixed = rng.random(n) < fixed_frac
    sp_idx = rng.integers(0, n_species, n).astype(np.int32)
    st = make_state(pos, vel, r, sp_idx, spec, fixed=fixed, omega=om)
    wl = []
    if "bottom" in walls:
        wl.append(infinite_wall((0, 0, -1), (0, 0, 0.0), species=0, wall_id=0))
    if "top" in walls:
        wl.append(infinite_wall((0, 0, 1), (0, 0, lz), species=min(1, n_species - 1), wall_id=1))
    if "tilted" in walls:
        wl.append(infinite_wall((1, 0.3, -0.2), (lx, 0, 0), species=0, wall_id=2))
    if "cylinder" in walls:
        wl.append(cylindrical_wall(0.5 * min(lx, ly) + 0.3, species=0, wall_id=3))
    if "axisymmetric" in walls:
        # AxisymmetricIntersectionOfWalls about a vertical axis through the
        # box centre: a pedestal with a chamfered rim (three sub-walls) and a
        # casing around it, as the calibration heap uses them
        c = (0.5 * lx, 0.5 * ly, 0.0)
        rb = 0.3 * min(lx, ly)
        s3 = 1.0 / math.sqrt(2.0)
        wl.append(axisymmetric_wall((0, 0, 1), [((-1, 0, 0), (rb, 0, 0)), ((0, 0, -1), (0, 0, 0.3 * lz)),
                                                ((-s3, 0, -s3), (rb - 0.4, 0, 0.3 * lz))],
                                    position=c, species=0, wall_id=4))
        wl.append(axisymmetric_wall((0, 0, 1), [((1, 0, 0), (0.5 * min(lx, ly) - 0.2, 0, 0))], position=c,
                                    species=min(1, n_species - 1), wall_id=5))
    per = []
    if periodic[0]:
        per.append(periodic_boundary(0, 0.0, lx))
    if periodic[1]:
        per.append(periodic_boundary(1, 0.0, ly))
    tc = 0.01
    return Scenario(f"mixture_{laws}_{n}", spec, wl, per, (0.1, -0.05, -1.0), tc / 50, st)
