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
import dataclasses
import time
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


# --------------------------------------------------------------------------
# Calibration heaps with their own parameters and walls
# (Drivers/Calibration/CalibrationHeap.cpp)
# --------------------------------------------------------------------------

HEAP_GRAVITY = 9.8          # CalibrationHeap.cpp:55
HEAP_FILL_FRACTION = 0.3    # initialVolumeFraction, CalibrationHeap.cpp:47
HEAP_TC = 0.02              # collision time of a d = 1 sphere (mass 1)


@dataclass
class HeapParams:
    """Per-simulation parameters of a calibration heap: sliding and rolling
    friction of the particles (and of the base), restitution, and the seed of
    its packing."""
    n: int
    sliding_mu: float
    rolling_mu: float
    restitution: float
    seed: int


def heap_species(hp: HeapParams) -> List[S.Species]:
    """The heap's particle species (LinearViscoelasticFriction, k_t = 2/7 k,
    rolling k_r = 2/5 k, at t_c = HEAP_TC for mass 1) and the frictionless
    species of its casing (frictionlessWallSpecies, CalibrationHeap.cpp:66)."""
    s = S.Species(density=6.0 / np.pi, friction_law=dem.FRICTION_FULL)
    s.set_collision_time_and_restitution(HEAP_TC, hp.restitution, 1.0)
    s.sliding_stiffness = 2. / 7. * s.stiffness
    s.sliding_dissipation = 2. / 7. * s.dissipation
    s.set_sliding_friction(hp.sliding_mu)
    s.rolling_stiffness = 2. / 5. * s.stiffness
    s.rolling_dissipation = 2. / 5. * s.dissipation
    s.set_rolling_friction(hp.rolling_mu)
    casing = S.Species(density=6.0 / np.pi, friction_law=dem.FRICTION_FULL)
    casing.set_collision_time_and_restitution(HEAP_TC, hp.restitution, 1.0)
    return [s, casing]


def heap_base_radius(n: int, radii=(0.4, 0.6)) -> float:
    """baseRadius of a heap of n particles: the fill of CalibrationHeap.cpp:
    47-88 puts particles of volume 0.3 x the cylinder (radius R, height 3R)
    into it, so R = (n v / (0.9 pi))^(1/3), v the mean particle volume."""
    a, b = radii
    mean_r3 = (b ** 4 - a ** 4) / (4.0 * (b - a))
    v = 4.0 / 3.0 * np.pi * mean_r3
    return float((n * v / (0.9 * np.pi)) ** (1.0 / 3.0))


def calibration_heap(hp: HeapParams, radii=(0.4, 0.6)) -> S.Scenario:
    """One heap of CalibrationHeap.cpp:36-92: particles with d ~ U[0.8, 1.2]
    at volume fraction 0.3 in a cylinder of radius R and height 3R (a seeded
    loose packing: a cubic lattice of spacing (v / 0.3)^(1/3) >= the largest
    diameter, cut to the cylinder, filled from the bottom, each sphere
    jittered within its lattice cell), standing on the base -- an
    AxisymmetricIntersectionOfWalls pedestal r < R, z < 0 -- inside the
    casing, the cylinder r = R (the last wall, lifted once the fill has
    settled).  Gravity -9.8 z."""
    rng = np.random.default_rng(hp.seed)
    n = hp.n
    r = rng.uniform(radii[0], radii[1], n)
    R = heap_base_radius(n, radii)
    a, b = radii
    v = 4.0 / 3.0 * np.pi * (b ** 4 - a ** 4) / (4.0 * (b - a))
    h = max((v / HEAP_FILL_FRACTION) ** (1.0 / 3.0), 2.0 * b * 1.001)
    m = int(np.ceil((R - b) / h))
    g = np.arange(-m, m + 1) * h
    gx, gy = np.meshgrid(g, g, indexing="ij")
    keep = gx ** 2 + gy ** 2 <= (R - b) ** 2
    gx, gy = gx[keep], gy[keep]
    layers = int(np.ceil(n / gx.size)) + 1
    z = b + h * np.repeat(np.arange(layers), gx.size)
    pos = np.stack([np.tile(gx, layers), np.tile(gy, layers), z], axis=1)[:n]
    jit = 0.5 * (h - 2.0 * r)[:, None]
    pos = pos + rng.uniform(-1.0, 1.0, (n, 3)) * jit
    # stay inside the casing and above the base
    rad = np.hypot(pos[:, 0], pos[:, 1])
    over = rad > R - r
    pos[over, :2] *= ((R - r[over]) / rad[over])[:, None]
    pos[:, 2] = np.maximum(pos[:, 2], r)
    sp = heap_species(hp)
    st = S.make_state(pos, np.zeros((n, 3)), r, np.zeros(n, np.int32), sp)
    base = S.axisymmetric_wall((0, 0, 1), [((-1, 0, 0), (R, 0, 0)), ((0, 0, -1), (0, 0, 0))], species=0, wall_id=0)
    casing = S.axisymmetric_wall((0, 0, 1), [((1, 0, 0), (R, 0, 0))], species=1, wall_id=1)
    dt = HEAP_TC / 50.0
    return S.Scenario(f"heap{n}", sp, [base, casing], [], (0.0, 0.0, -HEAP_GRAVITY), dt, st,
                      info=dict(base_radius=R, params=hp))


def draw_heap_params(m: int, seed: int, n_range=(500, 20000)) -> List[HeapParams]:
    """m heaps: particle counts log-uniform in n_range, sliding friction
    U[0.1, 0.8], rolling friction U[0, 0.1], restitution U[0.5, 0.95], one
    seed each."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), m)).astype(int)
    return [HeapParams(int(ns[k]), float(rng.uniform(0.1, 0.8)), float(rng.uniform(0.0, 0.1)),
                       float(rng.uniform(0.5, 0.95)), int(seed * 100003 + k)) for k in range(m)]


def _set_pos(w, x, y):
    w.position[0] = x
    w.position[1] = y


@dataclass
class HeapEnsemble:
    scenario: S.Scenario
    offsets: np.ndarray         # (m, 3) shift of each heap
    id_first: np.ndarray        # (m + 1,) id ranges of the heaps
    base_radius: np.ndarray     # (m,)
    params: List[HeapParams]

    def member_of(self, ids: np.ndarray) -> np.ndarray:
        return np.searchsorted(self.id_first, ids, side="right") - 1


def heap_layout(R: np.ndarray, gap: float, schedule: str = "shelf") -> np.ndarray:
    """Centres (x, y) of heaps of base radii R in the x-y plane, at least `gap`
    apart (edge to edge of their bases).  "grid": a square grid of equal tiles
    of the largest heap; "shelf": the heaps by decreasing radius, left to
    right in rows about as wide as the square grid, each row as deep as its
    largest heap (a smaller footprint for mixed sizes)."""
    m = len(R)
    side = int(np.ceil(np.sqrt(m)))
    out = np.zeros((m, 2))
    if schedule == "grid":
        tile = 2.0 * float(R.max()) + gap
        for k in range(m):
            out[k] = ((k % side) * tile, (k // side) * tile)
        return out
    if schedule != "shelf":
        raise ValueError(f"unknown layout schedule {schedule!r}")
    width = side * (2.0 * float(np.mean(R)) + gap)
    order = np.argsort(-R, kind="stable")
    x = y = 0.0
    row_depth = 0.0
    for k in order:
        d = 2.0 * float(R[k]) + gap
        if x > 0.0 and x + d > width:
            y += row_depth
            x = 0.0
            row_depth = 0.0
        out[k] = (x + 0.5 * d, y + 0.5 * d)
        x += d
        row_depth = max(row_depth, d)
    return out


def pack_heaps(members: Sequence[S.Scenario], gap: float = None, schedule: str = "shelf") -> HeapEnsemble:
    """Heaps with their own species and walls in one scenario: heap k's
    particles get species 2k (its casing's species 2k + 1) and meet only
    walls 2k (base) and 2k + 1 (casing) through per-species wall groups;
    the heaps stand side by side in x-y (heap_layout), a gap of 8 largest
    radii apart."""
    m = len(members)
    dt = members[0].dt
    for sc in members:
        if sc.dt != dt or tuple(sc.gravity) != tuple(members[0].gravity):
            raise ValueError("ensemble members must share the time step and gravity")
    rmax = max(float(sc.state.radius.max()) for sc in members)
    if gap is None:
        gap = 8.0 * rmax
    R = np.array([sc.info["base_radius"] for sc in members])
    offsets = np.zeros((m, 3))
    offsets[:, :2] = heap_layout(R, gap, schedule)
    species, walls, parts = [], [], []
    id_first = np.zeros(m + 1, np.int64)
    for k, sc in enumerate(members):
        species += sc.species
        for w0 in sc.walls:
            w = dem.Wall.from_buffer_copy(w0)
            w.species = 2 * k + w0.species
            w.id = len(walls)
            _set_pos(w, w0.position[0] + offsets[k, 0], w0.position[1] + offsets[k, 1])
            walls.append(w)
        st = sc.state
        id_first[k + 1] = id_first[k] + st.n
        parts.append(dem.State(id=(np.arange(st.n) + id_first[k]).astype(np.uint32),
                               species=(st.species + 2 * k).astype(np.int32), radius=st.radius,
                               inv_mass=st.inv_mass, inv_inertia=st.inv_inertia, x=st.x + offsets[k], v=st.v,
                               w=st.w, q=st.q, f=None, t=None))
    cat = lambda f: np.concatenate([getattr(p, f) for p in parts])  # noqa: E731
    state = dem.State(id=cat("id"), species=cat("species"), radius=cat("radius"), inv_mass=cat("inv_mass"),
                      inv_inertia=cat("inv_inertia"), x=cat("x"), v=cat("v"), w=cat("w"), q=cat("q"))
    # the species table: each heap's own 2 x 2 block (particles, casing and
    # their mixed pair); heaps never meet, so the entries between heaps only
    # fill the table
    ns = 2 * m
    table = (dem.PairParams * (ns * ns))()
    blocks = [S.species_table(sc.species) for sc in members]
    for a in range(ns):
        ka = a // 2
        for b in range(ns):
            kb = b // 2
            table[a * ns + b] = blocks[ka][(a % 2) * 2 + (b % 2)] if ka == kb else blocks[ka][(a % 2) * 3]
    first = np.repeat(2 * np.arange(m), 2).astype(np.int32)
    count = np.tile([2, 0], m).astype(np.int32)
    sc = S.Scenario("heap_ensemble", species, walls, [], members[0].gravity, dt, state,
                    wall_groups=(first, count), pair_table=table)
    return HeapEnsemble(sc, offsets, id_first, R, [sc_.info["params"] for sc_ in members])


@dataclass
class HeapRun:
    fill_time: np.ndarray       # time the casing was lifted (nan: not yet)
    angle_of_repose: np.ndarray  # degrees, measured once settled after twice the fill time (nan: not reached)
    angle_at_end: np.ndarray    # degrees, atan(4 z_com / R) of the last state (nan: casing still on)
    removed: int                # particles deleted below the base
    steps: int
    seconds: dict = None        # wall time in stepping, the energy checks and the deletions


def deletion_sequence(n: int, doomed) -> tuple:
    """DeletionBoundary::checkBoundaryAfterParticlesMove (DeletionBoundary.cc:
    197-219) over a handler of n particles of which the indices `doomed` go:
    the scan in index order removes a particle with removeObject (the last
    particle takes its index, ParticleHandler::removeObject) and checks that
    index again.  Returns the indices in the order they are removed and the
    handler order afterwards (as indices into the old order)."""
    order = np.arange(n)
    dead = np.zeros(n, bool)
    dead[np.asarray(doomed, np.int64)] = True
    seq = []
    i = 0
    while i < n:
        if dead[i]:
            seq.append(i)
            dead[i] = dead[n - 1]
            order[i] = order[n - 1]
            n -= 1
        else:
            i += 1
    return np.array(seq, np.int64), order[:n]


def run_heap_protocol(e: dem.Engine, ens: HeapEnsemble, steps: int, check_every: int = 1000,
                      delete_every: int = 5, log=None) -> HeapRun:
    """CalibrationHeap::actionsAfterTimeStep (CalibrationHeap.cpp:103-141) for
    every heap of the ensemble, every check_every steps: during the fill, once
    the heap's kinetic energy is below 1e-2 of its elastic energy, its casing
    is lifted (wallHandler.removeLastObject: the group drops its last wall)
    and a DeletionBoundary removes its particles below z = 0; after twice the
    fill time, whenever the criterion holds again, the angle of repose is
    atan(4 z_com / R) (a conical heap).  The same angle of the last state is
    kept for heaps whose run ends before that.  Energies are the device's per-particle
    ones summed per heap.  The spheres below the base (which meet nothing
    once they have left it) are deleted every delete_every checks, since a
    deletion re-uploads the whole ensemble."""
    m = len(ens.params)
    ids = ens.scenario.state.id.copy()
    inv_mass = ens.scenario.state.inv_mass.copy()
    fill = np.ones(m, bool)
    fill_time = np.full(m, np.nan)
    aor = np.full(m, np.nan)
    first, count = (np.array(a) for a in ens.scenario.wall_groups)
    removed = 0
    done = 0
    checks = 0
    dt = ens.scenario.dt
    st0 = e.stats()
    t0 = st0.time
    secs = {"step": 0.0, "check": 0.0, "delete": 0.0}
    clock = time.perf_counter
    while done < steps:
        k = min(check_every, steps - done)
        c0 = clock()
        e.step(k)
        e.stats()  # waits for the steps
        c1 = clock()
        secs["step"] += c1 - c0
        done += k
        t = t0 + done * dt
        en = e.particle_energies()
        mem = ens.member_of(ids)
        mass = np.where(inv_mass > 0, 1.0 / np.where(inv_mass > 0, inv_mass, 1.0), 0.0)
        # spheres below the base are the deletion boundary's (deleted at once
        # in the reference): they count for no heap
        on = (en[:, 2] >= 0.0).astype(float)
        ke = np.bincount(mem, on * en[:, 0], minlength=m)
        ee = np.bincount(mem, on * en[:, 1], minlength=m)
        mz = np.bincount(mem, on * mass * en[:, 2], minlength=m)
        ms = np.bincount(mem, on * mass, minlength=m)
        settled = ke < 1e-2 * ee
        lift = fill & settled
        if lift.any():
            count[2 * np.flatnonzero(lift)] = 1
            e.set_wall_groups(first, count)
            fill &= ~lift
            fill_time[lift] = t
        measure = (~fill) & settled & ~lift & (t > 2.0 * fill_time)
        com = mz / np.maximum(ms, 1e-300)
        aor[measure] = np.degrees(np.arctan(4.0 * com[measure] / ens.base_radius[measure]))
        at_end = np.where(fill, np.nan, np.degrees(np.arctan(4.0 * com / ens.base_radius)))
        checks += 1
        c2 = clock()
        secs["check"] += c2 - c1
        if not fill.all() and (checks % delete_every == 0 or done >= steps):
            idx = e.find_in_region(0, (0.0, 0.0, -1.0, 0.0))  # DEM_REGION_HALF_SPACE
            if len(idx):
                seq, order = deletion_sequence(ids.size, idx)
                e.remove_particles(seq)
                ids, inv_mass = ids[order], inv_mass[order]
                removed += len(idx)
            secs["delete"] += clock() - c2
        if log is not None:
            log(f"heaps: {done}/{steps} steps, {int((~fill).sum())}/{m} casings lifted, "
                f"{int(np.isfinite(aor).sum())} angles, {removed} removed")
    return HeapRun(fill_time, aor, at_end, removed, done, {k: round(v, 3) for k, v in secs.items()})


def autotune(members: Sequence[S.Scenario], skins=(0.2, 0.4, 0.8, 1.6, 2.4), schedules=("shelf", "grid"),
             steps: int = 300, warm_steps: int = 3000, device: int = 0, log=None):
    """The layout schedule (heap_layout) and Verlet-list skin (a multiple of
    the smallest radius: the cell size of every grid level is twice the
    level's largest radius plus the skin) that step this ensemble fastest.
    Per schedule the packed ensemble is first run warm_steps steps (the fill:
    spheres falling into the casings, the phase whose displacements trigger
    most list rebuilds); every skin is then timed over `steps` steps from that
    same state (state and interactions uploaded again, after 20 untimed
    steps).  A skin whose lists overflow the neighbour capacity
    (DEM_ERR_CAPACITY) is left out.  Returns (schedule, skin, table of
    candidates with their ms per step)."""
    rmin = min(float(m.state.radius.min()) for m in members)
    table = []
    best = None
    for sch in schedules:
        ens = pack_heaps(members, schedule=sch)
        w = dem.Engine(device=device)
        S.apply(w, ens.scenario)
        w.compute_forces()
        w.step(warm_steps)
        st, ct, stats = w.download(), w.contacts(), w.stats()
        w.close()
        s0 = ens.scenario.state
        warm = dataclasses.replace(ens.scenario, state=dem.State(
            id=s0.id, species=s0.species, radius=s0.radius, inv_mass=s0.inv_mass, inv_inertia=s0.inv_inertia,
            x=st["x"], v=st["v"], w=st["w"], q=st["q"], f=st["f"], t=st["t"]))
        for f in skins:
            e = dem.Engine(device=device, skin=f * rmin)
            try:
                S.apply(e, warm, ct)
                e.set_time(stats.time, int(stats.n_steps))
                e.compute_forces()
                e.step(20)
                e.stats()
                t0 = time.perf_counter()
                e.step(steps)
                e.stats()
                ms = 1e3 * (time.perf_counter() - t0) / steps
            except dem.DemError as err:
                if err.code != dem.ERR_CAPACITY:
                    raise
                if log is not None:
                    log(f"autotune: {sch} layout, skin {f * rmin:.3g}: lists overflow, left out")
                continue
            finally:
                e.close()
            table.append({"schedule": sch, "skin": f * rmin, "ms_per_step": round(ms, 4)})
            if log is not None:
                log(f"autotune: {sch} layout, skin {f * rmin:.3g}: {ms:.3f} ms per step")
            if best is None or ms < best[2]:
                best = (sch, f * rmin, ms)
    return best[0], best[1], table
