"""Batched ensembles: many independent simulations in one engine context.

Calibration studies run hundreds of small simulations that differ in size and
random seed but share species, walls and time step (a heap poured onto a
floor, Drivers/Calibration/CalibrationHeap.cpp).  One such simulation cannot
fill the GPU: at a few thousand particles every step is a handful of launches
of a few microseconds.  Here the simulations are laid side by side in the
x-y plane, each in its own tile, with a gap wider than any interaction range,
so that they never interact; the engine then steps all of them with the same
launches.  Each simulation's particles keep their own ids (offset per
simulation) and can be taken apart again after the run.

Requirements: the same species table, time step and gravity for every member;
walls that are invariant under translations in x and y (planes with a normal
along z, such as the floor); no periodic boundaries.  Members must stay
within their tile plus half the gap for the duration of the run (the packer
reports the gap; `check_separated` verifies it).
"""
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import dem
from . import scenarios as S


@dataclass
class Packed:
    scenario: S.Scenario        # the ensemble as one scenario
    offsets: np.ndarray         # (m, 3) shift applied to each member
    first: np.ndarray           # (m + 1,) handler-index ranges of the members
    gap: float

    def member(self, k: int, x: np.ndarray) -> np.ndarray:
        """Positions of member k, shifted back to its own frame."""
        return x[self.first[k]:self.first[k + 1]] - self.offsets[k]

    def check_separated(self, x: np.ndarray) -> bool:
        """No member has moved into another member's half of the gap."""
        for k in range(len(self.offsets)):
            p = x[self.first[k]:self.first[k + 1]]
            r = self.scenario.state.radius[self.first[k]:self.first[k + 1]]
            lo, hi = self._tiles[k]
            if np.any(p[:, :2] - r[:, None] < lo - 0.5 * self.gap) or np.any(p[:, :2] + r[:, None] > hi + 0.5 * self.gap):
                return False
        return True


def _xy_invariant(w) -> bool:
    return w.type == dem.WALL_INFINITE and w.axis[0] == 0.0 and w.axis[1] == 0.0


def pack(members: Sequence[S.Scenario], gap: float = None) -> Packed:
    """Lay the members out on a square grid of tiles in x-y."""
    m = len(members)
    base = members[0]
    tab = base.table()
    for sc in members[1:]:
        if sc.dt != base.dt or tuple(sc.gravity) != tuple(base.gravity):
            raise ValueError("ensemble members must share the time step and gravity")
        if [bytes(t) for t in sc.table()] != [bytes(t) for t in tab]:
            raise ValueError("ensemble members must share the species table")
    for sc in members:
        if sc.periodic:
            raise ValueError("ensemble members cannot have periodic boundaries")
        if not all(_xy_invariant(w) for w in sc.walls):
            raise ValueError("ensemble walls must be planes with a normal along z")
    rmax = max(float(sc.state.radius.max()) for sc in members)
    if gap is None:
        gap = 8.0 * rmax
    lo = np.array([sc.state.x[:, :2].min(0) - sc.state.radius.max() for sc in members])
    hi = np.array([sc.state.x[:, :2].max(0) + sc.state.radius.max() for sc in members])
    tile = float(np.max(hi - lo)) + gap
    side = int(np.ceil(np.sqrt(m)))
    offsets = np.zeros((m, 3))
    tiles = []
    for k in range(m):
        cx, cy = k % side, k // side
        offsets[k, :2] = np.array([cx * tile, cy * tile]) - lo[k]
        tiles.append((lo[k] + offsets[k, :2], hi[k] + offsets[k, :2]))
    first = np.zeros(m + 1, np.int64)
    parts = []
    id0 = 0
    for k, sc in enumerate(members):
        st = sc.state
        first[k + 1] = first[k] + st.n
        parts.append(dem.State(id=(np.arange(st.n) + id0).astype(np.uint32), species=st.species, radius=st.radius,
                               inv_mass=st.inv_mass, inv_inertia=st.inv_inertia, x=st.x + offsets[k], v=st.v,
                               w=st.w, q=st.q, f=None, t=None))
        id0 += st.n
    cat = lambda f: np.concatenate([getattr(p, f) for p in parts])  # noqa: E731
    state = dem.State(id=cat("id"), species=cat("species"), radius=cat("radius"), inv_mass=cat("inv_mass"),
                      inv_inertia=cat("inv_inertia"), x=cat("x"), v=cat("v"), w=cat("w"), q=cat("q"))
    sc = S.Scenario("ensemble", base.species, list(base.walls), [], base.gravity, base.dt, state,
                    mixed_overrides=base.mixed_overrides, rotation=base.rotation)
    out = Packed(sc, offsets, first, gap)
    out._tiles = tiles
    return out


def heap(n: int, seed: int, species: List[S.Species], radii=(0.5, 1.5), dt=None) -> S.Scenario:
    """A column of n polydisperse spheres above a floor (the pouring stage of
    a calibration heap): a jittered cubic lattice of square footprint and
    height about twice its width, at rest, under gravity -1 in z."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(radii[0] / 2, radii[1] / 2, n)
    a = 2.0 * r.max() * 1.02
    nx = max(1, int(np.ceil((n / 2.0) ** (1.0 / 3.0))))
    ij = np.arange(n)
    pos = np.stack([(ij % nx) * a, ((ij // nx) % nx) * a, (ij // (nx * nx)) * a + r.max()], axis=1)
    pos[:, :2] += rng.uniform(-0.005, 0.005, (n, 2)) * a
    sp = rng.integers(0, len(species), n).astype(np.int32)
    st = S.make_state(pos, np.zeros((n, 3)), r, sp, species)
    floor = S.infinite_wall((0, 0, -1), (0, 0, 0.0), species=0, wall_id=0)
    if dt is None:
        dt = 0.02 * min(s.collision_time(1.0) for s in species)
    return S.Scenario(f"heap{n}", species, [floor], [], (0.0, 0.0, -1.0), dt, st)


def calibration_species() -> List[S.Species]:
    """LinearViscoelasticFriction with rolling and torsion friction (the
    calibration heaps need rolling resistance), t_c = 0.005, e = 0.5."""
    s = S.Species(density=6.0 / np.pi, friction_law=dem.FRICTION_FULL)
    s.set_collision_time_and_restitution(0.005, 0.5, 1.0)
    s.sliding_stiffness = 2. / 7. * s.stiffness
    s.sliding_dissipation = 2. / 7. * s.dissipation
    s.set_sliding_friction(0.5)
    s.rolling_stiffness = 2. / 5. * s.stiffness
    s.rolling_dissipation = 2. / 5. * s.dissipation
    s.set_rolling_friction(0.1)
    s.torsion_stiffness = 2. / 5. * s.stiffness
    s.torsion_dissipation = 2. / 5. * s.dissipation
    s.set_torsion_friction(0.05)
    return [s]
