"""ctypes binding of the DEM engine's C ABI (``include/mercurydpm_dem.h``).

``Engine`` wraps one ``dem_ctx``.  The product library is
``mercurydpm_fork_amd/lib/libmercurydpm_dem.so`` (HIP, gfx950); loading fails
loudly when it has not been built -- there is no CPU fallback.

``Engine`` accepts any library that exports the ABI's call sequence under a
symbol prefix, so test code can drive a checker with the same calls; the
package itself only ever loads the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libmercurydpm_dem.so")

DEM_WALL_KEY = 0x80000000

NORMAL_LINEAR_VISCOELASTIC = 0
NORMAL_HERTZIAN_VISCOELASTIC = 1
FRICTION_NONE = 0
FRICTION_SLIDING = 1
FRICTION_FULL = 2
FRICTION_MINDLIN = 3
WALL_INFINITE = 0
WALL_CYLINDER = 1
WALL_AXISYMMETRIC = 2
AXI_MAX_SUB = 4

ERRORS = {-1: "DEM_ERR_ARG", -2: "DEM_ERR_STATE", -3: "DEM_ERR_DEVICE", -4: "DEM_ERR_CAPACITY",
          -5: "DEM_ERR_UNSUPPORTED"}
ERR_ARG, ERR_STATE, ERR_DEVICE, ERR_CAPACITY = -1, -2, -3, -4


class DemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class PairParams(C.Structure):
    _fields_ = [("normal_law", C.c_int32), ("friction_law", C.c_int32),
                ("constant_restitution", C.c_int32), ("reserved", C.c_int32),
                ("stiffness", C.c_double), ("dissipation", C.c_double),
                ("elastic_modulus", C.c_double), ("sliding_stiffness", C.c_double),
                ("sliding_dissipation", C.c_double), ("sliding_mu", C.c_double),
                ("sliding_mu_static", C.c_double), ("shear_modulus", C.c_double),
                ("rolling_stiffness", C.c_double), ("rolling_dissipation", C.c_double),
                ("rolling_mu", C.c_double), ("rolling_mu_static", C.c_double),
                ("torsion_stiffness", C.c_double), ("torsion_dissipation", C.c_double),
                ("torsion_mu", C.c_double), ("torsion_mu_static", C.c_double),
                ("interaction_distance", C.c_double)]


class Wall(C.Structure):
    """dem_wall (include/mercurydpm_dem.h)."""
    _fields_ = [("type", C.c_int32), ("species", C.c_int32), ("id", C.c_uint32),
                ("n_sub", C.c_int32), ("position", C.c_double * 3),
                ("dist_coef", C.c_double * 3), ("axis", C.c_double * 3), ("radius", C.c_double),
                ("sub_position", (C.c_double * 3) * 4), ("sub_coef", (C.c_double * 3) * 4),
                ("sub_axis", (C.c_double * 3) * 4)]


class Periodic(C.Structure):
    _fields_ = [("axis", C.c_int32), ("reserved", C.c_int32), ("left", C.c_double),
                ("right", C.c_double), ("shift", C.c_double)]


class Config(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_neighbors", C.c_int32), ("skin", C.c_double),
                ("grid_levels", C.c_int32), ("precision", C.c_int32), ("list_sweep", C.c_int32),
                ("reserved", C.c_int32)]


SWEEP_PER_PARTICLE = 0
SWEEP_WAVEFRONT = 1


# candidates per call of dem_check_particles_static (DEM_CHECK_STATIC_MAX)
CHECK_STATIC_MAX = 4096

PRECISION_FP64 = 0
PRECISION_MIXED = 1


class Stats(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("n_steps", C.c_int64), ("n_rebuilds", C.c_int64),
                ("n_contacts", C.c_int64), ("n_wall_contacts", C.c_int64),
                ("n_list_entries", C.c_int64), ("max_list_length", C.c_int32),
                ("grid_levels", C.c_int32), ("time", C.c_double), ("skin", C.c_double),
                ("cell_size", C.c_double * 4), ("n_timed_steps", C.c_int64), ("t_force_ms", C.c_double),
                ("t_kick_ms", C.c_double), ("t_rebuild_ms", C.c_double), ("capacity", C.c_int64)]


_PTR = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32


class CGConfig(C.Structure):
    """dem_cg_config (include/mercurydpm_dem.h)."""
    _fields_ = [("n", C.c_int32), ("particle_dimensions", C.c_int32), ("width", C.c_double),
                ("min", C.c_double * 3), ("max", C.c_double * 3)]


# dem_interaction (include/mercurydpm_dem.h): one stored interaction with the
# force on P and the geometry of the last force pass
INTERACTION_DTYPE = np.dtype([("p_id", np.uint32), ("i_id", np.uint32), ("time_stamp", np.uint32),
                              ("current", np.uint32), ("force", np.float64, 3), ("tangential", np.float64, 3),
                              ("contact", np.float64, 3),
                              ("normal", np.float64, 3), ("overlap", np.float64), ("distance", np.float64),
                              ("spring", np.float64, 9)], align=True)
assert INTERACTION_DTYPE.itemsize == 200

CG_FIELDS = 29
CG_COLUMNS = (["VolumeFraction", "Density"] + [f"Momentum{a}" for a in "XYZ"] +
              [f"MomentumFlux{a}" for a in ("XX", "XY", "XZ", "YY", "YZ", "ZZ")] +
              [f"ContactStress{a}{b}" for a in "XYZ" for b in "XYZ"] +
              [f"InteractionForce{a}" for a in "XYZ"] + [f"ParticleSize{k}" for k in range(6)])


def _declare(lib, prefix: str, has_config: bool):
    def f(name, res, *args):
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = list(args)
        return fn

    f("create", _I32, C.POINTER(_PTR), *( [C.POINTER(Config)] if has_config else []))
    f("destroy", _I32, _PTR)
    f("last_error", C.c_char_p, _PTR)
    f("set_species", _I32, _PTR, _I32, _PTR)
    f("set_walls", _I32, _PTR, _I32, _PTR)
    f("set_wall_groups", _I32, _PTR, _I32, _PTR, _PTR)
    f("set_periodic", _I32, _PTR, _I32, _PTR)
    f("set_gravity", _I32, _PTR, _PTR)
    f("set_background_drag", _I32, _PTR, C.c_double)
    f("set_timestep", _I32, _PTR, C.c_double)
    f("set_rotation", _I32, _PTR, _I32)
    f("set_time", _I32, _PTR, C.c_double, _I64)
    f("upload_particles", _I32, _PTR, _I64, *([_PTR] * 11))
    f("upload_contacts", _I32, _PTR, _I64, _PTR, _PTR, _PTR, _PTR)
    f("compute_forces", _I32, _PTR)
    f("step", _I32, _PTR, _I64)
    f("step_begin", _I32, _PTR)
    f("step_end", _I32, _PTR)
    f("set_kinematics", _I32, _PTR, _I64, _PTR, _PTR, _PTR, _PTR)
    f("add_particles", _I32, _PTR, _I64, *([_PTR] * 9))
    f("check_particles", _I32, _PTR, _I64, *([_PTR] * 5))
    if hasattr(lib, prefix + "check_particles_static"):  # the HIP library only
        f("check_particles_static", _I32, _PTR, _I64, *([_PTR] * 5))
    f("remove_particles", _I32, _PTR, _I64, _PTR)
    f("find_in_region", _I32, _PTR, _I32, _PTR, _I64, C.POINTER(_I64), _PTR)
    f("cg_z", _I32, _PTR, C.POINTER(CGConfig), _PTR)
    f("download_particles", _I32, _PTR, *([_PTR] * 6))
    f("contact_count", _I32, _PTR, C.POINTER(_I64))
    f("download_contacts", _I32, _PTR, _I64, C.POINTER(_I64), _PTR, _PTR, _PTR, _PTR)
    f("download_interactions", _I32, _PTR, _I64, C.POINTER(_I64), _PTR)
    if hasattr(lib, prefix + "particle_energies"):  # the HIP library only
        f("particle_energies", _I32, _PTR, _PTR)


_hip_lib = None


def load_hip_library(path: Optional[str] = None):
    """Load the HIP engine; raises if it is missing (no fallback)."""
    global _hip_lib
    if _hip_lib is not None and path is None:
        return _hip_lib
    p = path or HIP_LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP engine library not found at {p}; run __graft_entry__.build() (hipcc, gfx950)")
    lib = C.CDLL(p)
    _declare(lib, "dem_", True)
    lib.dem_get_stats.restype = _I32
    lib.dem_get_stats.argtypes = [_PTR, C.POINTER(Stats)]
    lib.dem_abi_version.restype = _I32
    lib.dem_abi_version.argtypes = []
    lib.dem_set_profiling.restype = _I32
    lib.dem_set_profiling.argtypes = [_PTR, _I32]
    if path is None:
        _hip_lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class State:
    """Per-particle state in handler order (numpy arrays)."""
    id: np.ndarray
    species: np.ndarray
    radius: np.ndarray
    inv_mass: np.ndarray
    inv_inertia: np.ndarray
    x: np.ndarray
    v: np.ndarray
    w: np.ndarray
    q: np.ndarray
    f: Optional[np.ndarray] = None
    t: Optional[np.ndarray] = None

    @property
    def n(self) -> int:
        return int(self.id.shape[0])


@dataclass
class Contacts:
    p_id: np.ndarray
    i_id: np.ndarray
    spring: np.ndarray  # (n, 9)
    ts: np.ndarray

    @property
    def n(self) -> int:
        return int(self.p_id.shape[0])

    def keys(self):
        return set(zip(self.p_id.tolist(), self.i_id.tolist()))


class Engine:
    """One simulation context behind the C ABI."""

    def __init__(self, lib=None, prefix: str = "dem_", device: int = 0, max_neighbors: int = 0,
                 skin: float = 0.0, grid_levels: int = 0, precision: int = 0, list_sweep: int = 0):
        self.lib = lib if lib is not None else load_hip_library()
        self.prefix = prefix
        self._has_stats = hasattr(self.lib, prefix + "get_stats") and prefix == "dem_"
        h = _PTR()
        if prefix == "dem_":
            cfg = Config(device, max_neighbors, skin, grid_levels, precision, list_sweep, 0)
            rc = self._fn("create")(C.byref(h), C.byref(cfg))
        else:
            rc = self._fn("create")(C.byref(h))
        if rc != 0:
            raise DemError(rc, "create failed")
        self.h = h
        self.n = 0

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc: int):
        if rc != 0:
            msg = self._fn("last_error")(self.h)
            raise DemError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self._fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ----------------------------------------------------------------- setup
    def set_species(self, table: Sequence[PairParams], n_species: int):
        arr = table if isinstance(table, C.Array) else (PairParams * len(table))(*table)
        self._check(self._fn("set_species")(self.h, n_species, C.cast(arr, C.c_void_p)))

    def set_walls(self, walls: Sequence[Wall]):
        arr = (Wall * max(1, len(walls)))(*walls)
        self._check(self._fn("set_walls")(self.h, len(walls), C.cast(arr, C.c_void_p)))

    def set_wall_groups(self, first, count):
        """Per-species wall groups: particles of species s meet walls
        first[s] .. first[s] + count[s] - 1 only (dem_set_wall_groups)."""
        f = np.ascontiguousarray(first, np.int32)
        c = np.ascontiguousarray(count, np.int32)
        self._check(self._fn("set_wall_groups")(self.h, len(f), _ptr(f), _ptr(c)))

    def set_periodic(self, bnds: Sequence[Periodic]):
        arr = (Periodic * max(1, len(bnds)))(*bnds)
        self._check(self._fn("set_periodic")(self.h, len(bnds), C.cast(arr, C.c_void_p)))

    def set_gravity(self, g):
        a = np.ascontiguousarray(g, dtype=np.float64)
        self._check(self._fn("set_gravity")(self.h, _ptr(a)))

    def set_background_drag(self, d: float):
        self._check(self._fn("set_background_drag")(self.h, float(d)))

    def set_timestep(self, dt: float):
        self._check(self._fn("set_timestep")(self.h, float(dt)))

    def set_rotation(self, on: bool):
        self._check(self._fn("set_rotation")(self.h, 1 if on else 0))

    def set_time(self, t: float, nt: int):
        self._check(self._fn("set_time")(self.h, float(t), int(nt)))

    def upload(self, s: State):
        def c(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt)
        arrs = [c(s.id, np.uint32), c(s.species, np.int32), c(s.radius, np.float64),
                c(s.inv_mass, np.float64), c(s.inv_inertia, np.float64), c(s.x, np.float64),
                c(s.v, np.float64), c(s.w, np.float64), c(s.q, np.float64), c(s.f, np.float64),
                c(s.t, np.float64)]
        self._keep = arrs
        self._check(self._fn("upload_particles")(self.h, s.n, *[_ptr(a) for a in arrs]))
        self.n = s.n

    def upload_contacts(self, ct: Contacts):
        p = np.ascontiguousarray(ct.p_id, dtype=np.uint32)
        i = np.ascontiguousarray(ct.i_id, dtype=np.uint32)
        sp = np.ascontiguousarray(ct.spring, dtype=np.float64)
        ts = np.ascontiguousarray(ct.ts, dtype=np.uint32)
        self._check(self._fn("upload_contacts")(self.h, ct.n, _ptr(p), _ptr(i), _ptr(sp), _ptr(ts)))

    # ------------------------------------------------------------------ run
    def compute_forces(self):
        self._check(self._fn("compute_forces")(self.h))

    def step(self, n: int = 1):
        self._check(self._fn("step")(self.h, int(n)))

    def cg_z(self, n: int, width: float, lo, hi, particle_dimensions: int = 3) -> np.ndarray:
        """CG<Z, Gauss, StandardFields> of the current state: an (n, 29)
        array, columns as CG_COLUMNS, at z = lo[2] + (k + 1/2) (hi[2] - lo[2]) / n."""
        cfg = CGConfig(int(n), int(particle_dimensions), float(width), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi))
        out = np.zeros((int(n), CG_FIELDS))
        self._check(self._fn("cg_z")(self.h, C.byref(cfg), _ptr(out)))
        return out

    def step_begin(self):
        """First part of computeOneTimeStep, up to actionsBeforeTimeStep."""
        self._check(self._fn("step_begin")(self.h))

    def step_end(self):
        """The rest of the step (forces, closing half-kick, erase, clock)."""
        self._check(self._fn("step_end")(self.h))

    def set_kinematics(self, x=None, v=None, w=None, q=None):
        """Prescribed positions / velocities / angular velocities /
        orientations of all particles (handler order); None = unchanged."""
        arrs = []
        for a, m in ((x, 3), (v, 3), (w, 3), (q, 4)):
            arrs.append(None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).reshape(self.n, m)))
        self._check(self._fn("set_kinematics")(self.h, self.n, *[_ptr(a) for a in arrs]))

    def add_particles(self, s: State):
        """ParticleHandler::copyAndAddObject during a run (between step_begin
        and step_end, as insertion boundaries do): append the particles of
        `s` (ids above every resident id) after the resident ones."""
        def c(a, dt):
            return None if a is None else np.ascontiguousarray(a, dtype=dt)
        arrs = [c(s.id, np.uint32), c(s.species, np.int32), c(s.radius, np.float64),
                c(s.inv_mass, np.float64), c(s.inv_inertia, np.float64), c(s.x, np.float64),
                c(s.v, np.float64), c(s.w, np.float64), c(s.q, np.float64)]
        self._check(self._fn("add_particles")(self.h, s.n, *[_ptr(a) for a in arrs]))
        self.n += s.n

    def remove_particles(self, index):
        """ParticleHandler::removeObject for each handler index in turn (the
        last particle takes the removed one's index)."""
        idx = np.ascontiguousarray(np.atleast_1d(np.asarray(index, np.int64)))
        self._check(self._fn("remove_particles")(self.h, len(idx), _ptr(idx)))
        self.n -= len(idx)

    def find_in_region(self, kind: int, params) -> np.ndarray:
        """Handler indices of the particles a deletion boundary removes
        (kind 0: half-space normal, distance; kind 1: open box min, max)."""
        prm = np.ascontiguousarray(params, np.float64)
        n = _I64()
        self._check(self._fn("find_in_region")(self.h, int(kind), _ptr(prm), 0, C.byref(n), None))
        out = np.zeros(max(n.value, 1), np.int64)
        self._check(self._fn("find_in_region")(self.h, int(kind), _ptr(prm), n.value, C.byref(n), _ptr(out)))
        return out[:n.value]

    def check_particles(self, x, radius, species, max_distance, static: bool = False) -> np.ndarray:
        """DPMBase::checkParticleForInteraction for candidate spheres against
        the resident particles and the walls: True where a candidate (and its
        periodic images within max_distance of a boundary; none where
        max_distance < 0) touches nothing.  static: through
        dem_check_particles_static, which stages the candidates in the
        library's static per-GPU array (at most CHECK_STATIC_MAX) and
        allocates nothing."""
        x = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 3))
        k = len(x)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, np.float64), (k,)))
        sp = np.ascontiguousarray(np.broadcast_to(np.asarray(species, np.int32), (k,)))
        md = np.ascontiguousarray(np.broadcast_to(np.asarray(max_distance, np.float64), (k,)))
        out = np.zeros(k, np.int32)
        fn = self._fn("check_particles_static" if static else "check_particles")
        self._check(fn(self.h, k, _ptr(x), _ptr(r), _ptr(sp), _ptr(md), _ptr(out)))
        return out.astype(bool)

    # ------------------------------------------------------------- download
    def download(self):
        n = self.n
        x = np.empty((n, 3)); v = np.empty((n, 3)); w = np.empty((n, 3))
        q = np.empty((n, 4)); f = np.empty((n, 3)); t = np.empty((n, 3))
        self._check(self._fn("download_particles")(self.h, _ptr(x), _ptr(v), _ptr(w), _ptr(q),
                                                   _ptr(f), _ptr(t)))
        return dict(x=x, v=v, w=w, q=q, f=f, t=t)

    def particle_energies(self) -> np.ndarray:
        """(n, 3) in handler order: kinetic energy, elastic energy of the
        interactions held as P, height z (dem_particle_energies)."""
        out = np.empty((self.n, 3))
        self._check(self._fn("particle_energies")(self.h, _ptr(out)))
        return out

    def contact_count(self) -> int:
        n = _I64()
        self._check(self._fn("contact_count")(self.h, C.byref(n)))
        return int(n.value)

    def contacts(self) -> Contacts:
        cap = max(1, self.contact_count())
        p = np.empty(cap, np.uint32); i = np.empty(cap, np.uint32)
        sp = np.empty((cap, 9)); ts = np.empty(cap, np.uint32)
        n = _I64()
        self._check(self._fn("download_contacts")(self.h, cap, C.byref(n), _ptr(p), _ptr(i),
                                                  _ptr(sp), _ptr(ts)))
        k = int(n.value)
        return Contacts(p[:k].copy(), i[:k].copy(), sp[:k].copy(), ts[:k].copy())

    def interactions(self) -> np.ndarray:
        """The stored interactions as an INTERACTION_DTYPE array sorted by
        (p_id, i_id): what the .fstat writer reads."""
        cap = max(1, self.contact_count())
        out = np.zeros(cap, INTERACTION_DTYPE)
        n = _I64()
        self._check(self._fn("download_interactions")(self.h, cap, C.byref(n), _ptr(out)))
        return out[:int(n.value)].copy()

    def set_profiling(self, on: bool):
        self._check(self.lib.dem_set_profiling(self.h, 1 if on else 0))

    def stats(self) -> Stats:
        s = Stats()
        if self._has_stats:
            self._check(self.lib.dem_get_stats(self.h, C.byref(s)))
        return s
