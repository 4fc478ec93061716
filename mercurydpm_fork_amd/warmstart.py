"""Warm starts of chute flows from the learned closure.

A periodic chute at angle theta is steady when the shear stress balances
gravity at every depth: sigma_xz / sigma_zz = tan(theta), with the pressure
the weight of the material above, P(z) = g cos(theta) int_z^H rho.  The
closure network maps the local state (inertial number I, packing fraction,
pressure, granular temperature) to the effective friction mu; solving
mu = tan(theta) bin by bin gives I(z), hence the shear rate
gamma_dot = I sqrt(P / rho_p) / d and, integrated up from the base, the
steady velocity profile u(z).  A DEM run started from the settled bed with
that profile (instead of at rest) reaches the steady flow in fewer steps;
the profile is measured against DEM itself, run to steady state from rest.

The granular temperature enters the network as an input; in steady flow it
follows the shear rate, T ~ (gamma_dot d)^2, and its law is fitted to the
training states (log T linear in log I and log P)."""
from dataclasses import dataclass

import numpy as np

from . import closure_data as CD
from . import dem
from . import scenarios as S

RHO_P = 6.0 / np.pi  # particle density: mass 1 for d = 1


def chute_case(angle: float, mu: float = 0.45, n: int = 6000, depth: float = 10.0, tc: float = 0.02,
               seed: int = 5) -> S.Scenario:
    """A small periodic chute of Hertz-Mindlin spheres, d ~ U[0.8, 1.2], e =
    0.8, sliding friction mu, collision time tc (dt = tc / 50), over a rough
    base of fixed d = 1 spheres."""
    sp = S.hertz_mindlin_species(mu=mu, tc=tc)
    return S.chute_hvm(n_flow=n, seed=seed, seed_base=seed + 1, angle_deg=angle, depth=depth, d_range=(0.8, 1.2),
                       species=sp)


def settled_engine(sc: S.Scenario, settle_steps: int = 25000, **kw) -> dem.Engine:
    """The engine holding the chute's bed settled at rest under |g| straight
    down (the common start of a cold and a warm run); the chute gravity is
    set afterwards."""
    e = dem.Engine(**kw)
    S.apply(e, sc)
    e.compute_forces()
    S.settle_chute(e, sc, settle_steps)
    return e


def cg_profile(e: dem.Engine, sc: S.Scenario, zmax: float, nz: int = 64, width: float = 0.5, n_avg: int = 1,
               every: int = 1000) -> np.ndarray:
    """CG<Z, Gauss> fields on nz points over [0, zmax] (width 0.5 d),
    averaged over n_avg states `every` steps apart (the last one is the
    current state when n_avg = 1)."""
    l = sc.info["footprint"]
    acc = 0.0
    for k in range(n_avg):
        if k:
            e.step(every)
        acc = acc + e.cg_z(nz, width, lo=(0.0, 0.0, 0.0), hi=(l, l, zmax))
    return acc / n_avg


def profile_z(zmax: float, nz: int) -> np.ndarray:
    """Heights of the CG points of cg_profile: bin centres of [0, zmax]."""
    return (np.arange(nz) + 0.5) * zmax / nz


def temperature_law(states: np.ndarray) -> np.ndarray:
    """Coefficients (a, b, c) of log10 T = a + b log10 I + c log10 P over
    training states (rows of I, phi, P, T, mu)."""
    I, _, P, T, _ = states.T
    A = np.stack([np.ones_like(I), np.log10(I), np.log10(P)], axis=1)
    coef, *_ = np.linalg.lstsq(A, np.log10(np.abs(T) + 1e-12), rcond=None)
    return coef


@dataclass
class Prediction:
    z: np.ndarray        # heights of the bins
    inertial: np.ndarray  # I(z) where the bed is (nan elsewhere)
    pressure: np.ndarray
    velocity: np.ndarray  # u(z)
    shear_rate: np.ndarray
    packing: np.ndarray   # the network's phi(z) in the bed


def predict_profile(net, tlaw: np.ndarray, angle: float, fields: np.ndarray, z: np.ndarray, z_base: float,
                    d: float = 1.0, phi_min: float = 0.3) -> Prediction:
    """Steady velocity profile of a chute at `angle` degrees from the closure
    network `net` (a closure.Closure or closure.MLP) and the settled bed's CG
    fields (density and volume fraction per height): in every bin with packing
    fraction above phi_min, the inertial number where the network's friction
    equals tan(angle) (bisection in log I between 1e-4 and 1; phi and T
    consistent with it), then the shear rate, integrated from z_base up
    (the plane of the base spheres' centres: the first layer above it slips
    over the base)."""
    th = np.radians(angle)
    rho = fields[:, CD.COL["Density"]]
    phi0 = fields[:, CD.COL["VolumeFraction"]]
    dz = z[1] - z[0]
    # overburden: the weight per area of the material above each height
    above = np.cumsum((rho * dz)[::-1])[::-1] - 0.5 * rho * dz
    P = np.cos(th) * above
    bed = (phi0 > phi_min) & (z > z_base) & (P > 0)
    lo = np.full(z.size, -4.0)
    hi = np.zeros(z.size)
    phi = phi0.copy()
    target = np.tan(th)
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        I = 10.0 ** mid
        T = 10.0 ** (tlaw[0] + tlaw[1] * mid + tlaw[2] * np.log10(np.where(bed, P, 1.0)))
        st = np.stack([I, phi, np.where(bed, P, 1.0), T, np.zeros_like(I)], axis=1)
        x, _ = CD.standardise(st)
        y = np.asarray(net(x), np.float64)
        mu = 0.5 + 0.1 * y[:, 0]
        phi = np.where(bed, np.clip(0.55 + 0.05 * y[:, 1], 0.3, 0.7), phi)
        up = mu > target  # friction grows with I: the root lies below
        hi = np.where(up, mid, hi)
        lo = np.where(up, lo, mid)
    I = np.where(bed, 10.0 ** (0.5 * (lo + hi)), np.nan)
    gdot = np.where(bed, I * np.sqrt(np.where(bed, P, 1.0) / RHO_P) / d, 0.0)
    u = np.cumsum(gdot) * dz
    u = u - np.interp(z_base, z, u)
    u = np.where(z > z_base, u, 0.0)
    # above the bed the free surface moves with the top of the bed
    top = np.max(np.flatnonzero(bed)) if bed.any() else 0
    u[top + 1:] = u[top]
    return Prediction(z, I, P, u, gdot, np.where(bed, phi, np.nan))


def warm_start(e: dem.Engine, sc: S.Scenario, pred: Prediction, settled_phi: np.ndarray, z_base: float):
    """Start the settled bed in the predicted steady flow: every free sphere
    above the base moves with u(z) along x and spins with half the local
    vorticity (omega_y = gamma_dot / 2), and the bed is dilated vertically
    from the predicted packing fraction of the flow (heights above z_base
    scaled by the settled over the predicted mean packing fraction, at most
    by 10%)."""
    st = e.download()
    free = sc.state.inv_mass > 0
    bed = np.isfinite(pred.packing)
    scale = float(np.mean(settled_phi[bed]) / np.mean(pred.packing[bed])) if bed.any() else 1.0
    x = st["x"].copy()
    zf = x[free, 2]
    scale = min(max(scale, 1.0), 1.1)  # a flowing bed dilates by a few per cent
    x[free, 2] = np.where(zf > z_base, z_base + (zf - z_base) * scale, zf)
    v = np.zeros_like(st["v"])
    w = np.zeros_like(st["w"])
    v[free, 0] = np.interp(x[free, 2], pred.z, pred.velocity)
    w[free, 1] = 0.5 * np.interp(x[free, 2], pred.z, pred.shear_rate)
    e.set_kinematics(x=x, v=v, w=w)
    return scale


def velocity_of(fields: np.ndarray) -> np.ndarray:
    """Mean flow velocity per bin, MomentumX / Density."""
    rho = fields[:, CD.COL["Density"]]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(rho > 0, fields[:, CD.COL["MomentumX"]] / rho, 0.0)


def profile_error(u: np.ndarray, ref: np.ndarray, mask: np.ndarray) -> float:
    """Relative L2 error of a velocity profile over the bins of mask."""
    return float(np.linalg.norm((u - ref)[mask]) / np.linalg.norm(ref[mask]))
