"""Species, particle-mass and wall-orientation setters, restated from the
reference so the engine receives the parameters a MercuryDPM driver would
produce.  Python floats are IEEE binary64 and every expression keeps the
reference's operation order, so the values are bit-identical to the C++
setters.  (The C++ host in mercurydpm_fork_amd/host implements the same
setters for C++ drivers.)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

from . import dem

PI = 3.1415926535897932384626433832795  # constants::pi (ExtendedMath.h:44)
SQR_PI = 9.8696044010893586188344909998762  # constants::sqr_pi (ExtendedMath.h:46)


def maths_log(power: float) -> float:
    """mathsFunc::log (Kernel/Math/ExtendedMath.cc:104-129): a Newton-type
    iteration used by setCollisionTimeAndRestitutionCoefficient."""
    E = 2.71828182845905
    P = power
    N = 0.0
    while P >= E:
        P /= E
        N += 1
    N += P / E
    P = power
    while True:
        A = N
        L = P / math.exp(N - 1.0)
        R = (N - 1.0) * E
        N = (L + R) / E
        if not (N < A):
            break
    return N


def average(a: float, b: float) -> float:
    """BaseSpecies::average (BaseSpecies.cc:110-115): harmonic mean."""
    return 0.0 if (a == 0 or b == 0) else 2.0 / (1.0 / a + 1.0 / b)


def sphere_inv_mass(radius: float, density: float, dims: int = 3):
    """BaseParticle::computeMass (BaseParticle.cc:874-883): (invMass, invInertia XX)."""
    if dims == 3:
        inv_mass = 1.0 / (4.0 / 3.0 * PI * radius * radius * radius * density)
        mass = 1.0 / inv_mass
        inv_i = 1.0 / (.4 * mass * (radius * radius))
    else:
        inv_mass = 1.0 / (PI * radius * radius * density)
        mass = 1.0 / inv_mass
        inv_i = 1.0 / (.5 * mass * (radius * radius))
    return inv_mass, inv_i


def sphere_inv_mass_array(radius, density, dims: int = 3):
    """Element-wise sphere_inv_mass for numpy arrays (same operation order)."""
    import numpy as np
    r = np.asarray(radius, dtype=np.float64)
    rho = np.asarray(density, dtype=np.float64)
    if dims == 3:
        inv_mass = 1.0 / (4.0 / 3.0 * PI * r * r * r * rho)
        inv_i = 1.0 / (.4 * (1.0 / inv_mass) * (r * r))
    else:
        inv_mass = 1.0 / (PI * r * r * rho)
        inv_i = 1.0 / (.5 * (1.0 / inv_mass) * (r * r))
    return inv_mass, inv_i


def volume_from_radius(radius: float, dims: int = 3) -> float:
    """BaseParticle::getVolume (BaseParticle.cc:147-166)."""
    return 4.0 / 3.0 * PI * radius * radius * radius if dims == 3 else PI * radius * radius


@dataclass
class Species:
    """A ParticleSpecies with its normal and friction parameters
    (Species<Normal, Friction> in the reference)."""
    density: float = 1.0
    normal_law: int = dem.NORMAL_LINEAR_VISCOELASTIC
    friction_law: int = dem.FRICTION_NONE
    constant_restitution: bool = False
    stiffness: float = 0.0
    dissipation: float = 0.0
    elastic_modulus: float = 0.0
    sliding_stiffness: float = 0.0
    sliding_dissipation: float = 0.0
    sliding_mu: float = 0.0
    sliding_mu_static: float = 0.0
    shear_modulus: float = 0.0
    rolling_stiffness: float = 0.0
    rolling_dissipation: float = 0.0
    rolling_mu: float = 0.0
    rolling_mu_static: float = 0.0
    torsion_stiffness: float = 0.0
    torsion_dissipation: float = 0.0
    torsion_mu: float = 0.0
    torsion_mu_static: float = 0.0

    # --- LinearViscoelasticNormalSpecies (LinearViscoelasticNormalSpecies.cc)
    def set_collision_time_and_restitution(self, tc: float, eps: float, mass: float):
        """setCollisionTimeAndRestitutionCoefficient (cc:225-245)."""
        if self.constant_restitution:
            mass = 1
        if eps == 0.0:
            self.stiffness = .5 * mass * ((PI / tc) * (PI / tc))
            self.dissipation = math.sqrt(2.0 * mass * self.stiffness)
        else:
            self.dissipation = -mass / tc * maths_log(eps)
            self.stiffness = .5 * mass * ((PI / tc) * (PI / tc) + (self.dissipation / mass) * (self.dissipation / mass))
        return self

    def set_collision_time_and_normal_and_tangential_restitution(self, tc: float, eps: float, beta: float,
                                                                   mass: float):
        """SlidingFrictionSpecies::setCollisionTimeAndNormalAndTangential-
        RestitutionCoefficient (SlidingFrictionSpecies.cc:207-246) for a
        linear viscoelastic normal law: k and gamma from (tc, eps), then the
        tangential stiffness and dissipation of Deen et al. (2007), eqs. 30
        and 43.  The friction coefficient is left as it is."""
        self.set_collision_time_and_restitution(tc, eps, mass)
        stiffness = self.stiffness
        lb = math.log(beta) if beta != 0.0 else -math.inf
        le = math.log(eps)
        self.sliding_stiffness = 2.0 / 7.0 * stiffness * (PI * PI + lb * lb) / (PI * PI + le * le)
        if beta != 0.0:
            self.sliding_dissipation = -2 * lb * math.sqrt(1.0 / 7.0 * mass * self.sliding_stiffness /
                                                           (PI * PI + lb * lb))
        else:
            self.sliding_dissipation = 2. * math.sqrt(1.0 / 7.0 * mass * self.sliding_stiffness)
        return self

    def set_stiffness_and_restitution(self, k: float, eps: float, mass: float):
        """setStiffnessAndRestitutionCoefficient / setRestitutionCoefficient (cc:207-222)."""
        if self.constant_restitution:
            mass = 1
        self.stiffness = k
        if eps == 0.0:
            self.dissipation = math.sqrt(2.0 * mass * k)
        else:
            log_eps = math.log(eps)
            self.dissipation = -math.sqrt(2.0 * mass * k / (SQR_PI + log_eps * log_eps)) * log_eps
        return self

    def collision_time(self, mass: float) -> float:
        """LinearViscoelasticNormalSpecies::getCollisionTime (cc:137-166)."""
        if self.constant_restitution:
            mass = 1
        tosqrt = self.stiffness / (.5 * mass) - (self.dissipation / mass) * (self.dissipation / mass)
        return PI / math.sqrt(tosqrt)

    def restitution(self, mass: float) -> float:
        if self.constant_restitution:
            mass = 1
        return math.exp(-self.dissipation / mass * self.collision_time(mass))

    # --- HertzianViscoelasticNormalSpecies (cc:109-135, 290-294)
    def set_elastic_modulus_and_restitution(self, E: float, rest: float):
        self.normal_law = dem.NORMAL_HERTZIAN_VISCOELASTIC
        self.elastic_modulus = E
        if rest > 0.0:
            lr2 = math.log(rest) * math.log(rest)
            self.dissipation = math.sqrt(5.0 * lr2 / (lr2 + SQR_PI))
        else:
            self.dissipation = math.sqrt(5.0)
        return self

    def set_elastic_modulus_and_poisson_ratio(self, E: float, nu: float):
        """setEffectiveElasticModulusAndPoissonRatio
        (HertzianViscoelasticNormalSpecies.cc:142-159).  The shear modulus is
        the reference's expression as written, E / 2 * (1 + nu)."""
        self.normal_law = dem.NORMAL_HERTZIAN_VISCOELASTIC
        self.elastic_modulus = E
        self.shear_modulus = E / 2 * (1 + nu)
        return self

    def hertz_collision_time(self, diameter: float, density: float, rel_velocity: float) -> float:
        r = density / self.elastic_modulus
        return 2.214 * math.pow(r * r / rel_velocity, 0.2) * diameter

    # --- friction setters (SlidingFrictionSpecies / MindlinSpecies)
    def set_sliding_friction(self, mu: float):
        self.sliding_mu = mu
        self.sliding_mu_static = mu
        return self

    def set_rolling_friction(self, mu: float):
        self.rolling_mu = mu
        self.rolling_mu_static = mu
        return self

    def set_torsion_friction(self, mu: float):
        self.torsion_mu = mu
        self.torsion_mu_static = mu
        return self

    def mass_from_radius(self, radius: float, dims: int = 3) -> float:
        return self.density * volume_from_radius(radius, dims)

    def uses_rotation(self) -> bool:
        """getUseAngularDOFs: true for every friction law."""
        return self.friction_law != dem.FRICTION_NONE

    def pair_params(self) -> "dem.PairParams":
        p = dem.PairParams()
        p.normal_law = self.normal_law
        p.friction_law = self.friction_law
        p.constant_restitution = 1 if self.constant_restitution else 0
        for name in ("stiffness", "dissipation", "elastic_modulus", "sliding_stiffness",
                     "sliding_dissipation", "sliding_mu", "sliding_mu_static", "shear_modulus",
                     "rolling_stiffness", "rolling_dissipation", "rolling_mu", "rolling_mu_static",
                     "torsion_stiffness", "torsion_dissipation", "torsion_mu", "torsion_mu_static"):
            setattr(p, name, getattr(self, name))
        p.interaction_distance = 0.0
        return p


def mix(a: Species, b: Species) -> Species:
    """MixedSpecies::mixAll: every parameter is the harmonic mean
    (LinearViscoelasticNormalSpecies::mix cc:267, SlidingFrictionSpecies::mix,
    MindlinSpecies::mix, HertzianViscoelasticNormalSpecies::mix)."""
    m = replace(a)
    for name in ("stiffness", "dissipation", "elastic_modulus", "sliding_stiffness",
                 "sliding_dissipation", "sliding_mu", "sliding_mu_static", "shear_modulus",
                 "rolling_stiffness", "rolling_dissipation", "rolling_mu", "rolling_mu_static",
                 "torsion_stiffness", "torsion_dissipation", "torsion_mu", "torsion_mu_static"):
        setattr(m, name, average(getattr(a, name), getattr(b, name)))
    return m


def species_table(species, mixed_overrides=None):
    """SpeciesHandler objects + mixed objects as the n x n table the ABI takes."""
    n = len(species)
    table = []
    for i in range(n):
        for j in range(n):
            if i == j:
                table.append(species[i].pair_params())
            else:
                key = (min(i, j), max(i, j))
                if mixed_overrides and key in mixed_overrides:
                    table.append(mixed_overrides[key].pair_params())
                else:
                    table.append(mix(species[max(i, j)], species[min(i, j)]).pair_params())
    return table


def _normalise(v):
    l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
    l = math.sqrt(l2)
    return (v[0] / l, v[1] / l, v[2] / l)


def orientation_via_normal(normal):
    """Quaternion::setOrientationViaNormal (Quaternion.cc:543-566)."""
    n = _normalise(normal)
    if n[0] <= -1 + 1e-14:
        return (0.0, 0.0, 1.0, 0.0)
    half = (1.0 + n[0], 0.0 + n[1], 0.0 + n[2])
    q = [half[0], 0.0, -half[2], half[1]]
    l2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]
    l = math.sqrt(l2)
    return tuple(x / l for x in q)


def _infinite_wall_fields(normal, point):
    """InfiniteWall::set(normal, point) (InfiniteWall.cc:118-122): the
    distance coefficients of Quaternion::getDistance (Quaternion.cc:658-669)
    and the axis of Quaternion::getAxis (Quaternion.cc:501-512) of the
    orientation via the normal."""
    q0, q1, q2, q3 = orientation_via_normal(normal)
    q00 = q0 * q0
    q02 = 2 * q0 * q2
    q03 = 2 * q0 * q3
    q11 = q1 * q1
    q12 = 2 * q1 * q2
    q13 = 2 * q1 * q3
    q22 = q2 * q2
    q33 = q3 * q3
    coef = (q00 + q11 - q22 - q33, q12 + q03, q13 - q02)
    a12 = q1 * q2
    a30 = q3 * q0
    a13 = q1 * q3
    a20 = q2 * q0
    axis = (q00 - q33 + q11 - q22, 2.0 * (a12 + a30), 2.0 * (a13 - a20))
    return tuple(float(c) for c in point), coef, axis


def infinite_wall(normal, point, species: int = 0, wall_id: int = 0) -> "dem.Wall":
    """InfiniteWall::set(normal, point) (InfiniteWall.cc:118-122)."""
    pos, coef, axis = _infinite_wall_fields(normal, point)
    w = dem.Wall()
    w.type = dem.WALL_INFINITE
    w.species = species
    w.id = wall_id
    for k in range(3):
        w.position[k] = pos[k]
        w.dist_coef[k] = coef[k]
        w.axis[k] = axis[k]
    w.radius = 0.0
    return w


def axisymmetric_wall(axis, sub_walls, position=(0.0, 0.0, 0.0), species: int = 0, wall_id: int = 0) -> "dem.Wall":
    """AxisymmetricIntersectionOfWalls (AxisymmetricIntersectionOfWalls.cc):
    setAxis(axis) (setOrientationViaNormal; the engine uses getOrientation()
    .getAxis()), setPosition(position), then IntersectionOfWalls::addObject(
    normal, point) for each (normal, point) of `sub_walls` in the (radial, 0,
    axial) plane (IntersectionOfWalls.cc:138-148: the normal is normalised,
    then InfiniteWall::set)."""
    if not 1 <= len(sub_walls) <= dem.AXI_MAX_SUB:
        raise ValueError("an axisymmetric wall has 1 to 4 sub-walls")
    _, _, ax = _infinite_wall_fields(axis, (0.0, 0.0, 0.0))
    w = dem.Wall()
    w.type = dem.WALL_AXISYMMETRIC
    w.species = species
    w.id = wall_id
    w.n_sub = len(sub_walls)
    for k in range(3):
        w.position[k] = float(position[k])
        w.axis[k] = ax[k]
    for j, (normal, point) in enumerate(sub_walls):
        pos, coef, a = _infinite_wall_fields(_normalise(normal), point)
        for k in range(3):
            w.sub_position[j][k] = pos[k]
            w.sub_coef[j][k] = coef[k]
            w.sub_axis[j][k] = a[k]
    return w


def cylindrical_wall(radius: float, species: int = 0, wall_id: int = 0) -> "dem.Wall":
    """CylindricalWall::set(radius) (CylindricalWall.cc:66-70)."""
    w = dem.Wall()
    w.type = dem.WALL_CYLINDER
    w.species = species
    w.id = wall_id
    w.radius = float(radius)
    return w


def periodic_boundary(axis: int, left: float, right: float) -> "dem.Periodic":
    """PeriodicBoundary::set(normal e_axis, left, right) (PeriodicBoundary.cc:84-94)."""
    scale = 1.0 / math.sqrt(1.0)
    b = dem.Periodic()
    b.axis = axis
    b.left = left * scale
    b.right = right * scale
    b.shift = 1.0 * (b.right - b.left)
    return b
