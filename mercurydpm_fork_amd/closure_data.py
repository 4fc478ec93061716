"""Training data for the learned closure from the DEM itself: coarse-grained
depth profiles of periodic chute flows (CG<Z, Gauss, StandardFields> through
dem_cg_z, averaged over snapshots) turned into local states (I, phi, P, T)
and their effective friction mu = -sigma_xz / sigma_zz.

Stress = contact stress + kinetic stress (MomentumFlux - rho v v); pressure
P = sigma_zz with compression positive; shear rate from the depth derivative
of the mean velocity v_x = MomentumX / Density; granular temperature
T = (trace of the kinetic stress) / (3 rho); I = gamma_dot d / sqrt(P / rho_p).
Only bins in the bulk (away from the base and the free surface, phi above
0.4) are kept."""
import numpy as np

from . import dem
from . import scenarios as S

COL = {name: k for k, name in enumerate(dem.CG_COLUMNS)}


def profile_states(fields, z, d=1.0, rho_p=6.0 / np.pi):
    """Local states of a CG depth profile: (I, phi, P, T, mu) per bin."""
    f = lambda name: fields[:, COL[name]]  # noqa: E731
    rho = f("Density")
    with np.errstate(divide="ignore", invalid="ignore"):
        vx = f("MomentumX") / rho
        kin = lambda a, b: f(f"MomentumFlux{a}{b}") - rho * (f(f"Momentum{a}") / rho) * (f(f"Momentum{b}") / rho)  # noqa: E731
        sxz = f("ContactStressXZ") + kin("X", "Z")
        szz = f("ContactStressZZ") + kin("Z", "Z")
        T = (kin("X", "X") + kin("Y", "Y") + kin("Z", "Z")) / (3.0 * rho)
        gdot = np.abs(np.gradient(vx, z))
        P = szz
        I = gdot * d / np.sqrt(P / rho_p)
        mu = -sxz / szz
    return np.stack([I, f("VolumeFraction"), P, T, mu], axis=1)


def chute_samples(angle, n_flow=4000, depth=12.0, steps=150_000, n_avg=20, avg_every=1000, nz=40, width=0.5,
                  seed=3, engine=None):
    """Run a periodic chute at `angle` degrees for `steps` steps, average the
    CG profile over n_avg snapshots avg_every steps apart, and return the bulk
    states (rows of I, phi, P, T, mu)."""
    sc = S.chute_hvm(n_flow=n_flow, seed=seed, angle_deg=angle, depth=depth)
    e = engine() if engine else dem.Engine()
    S.apply(e, sc)
    e.compute_forces()
    e.step(steps)
    l = sc.info["footprint"]
    zmax = float(e.download()["x"][:, 2].max()) + 1.0
    acc = 0.0
    for _ in range(n_avg):
        e.step(avg_every)
        acc = acc + e.cg_z(nz, width, lo=(0.0, 0.0, 0.0), hi=(l, l, zmax))
    fields = acc / n_avg
    z = (np.arange(nz) + 0.5) * zmax / nz
    st = profile_states(fields, z)
    keep = (z > 2.0) & (z < zmax - 3.0) & (st[:, 1] > 0.4) & np.all(np.isfinite(st), axis=1) & (st[:, 2] > 0) & \
        (st[:, 0] > 0)
    return st[keep]


def standardise(states):
    """Network inputs (log10 I, phi, log10 P, log10 T), centred and scaled
    with fixed constants, and targets ((mu - 0.5) / 0.1, (phi - 0.55) / 0.05)."""
    I, phi, P, T, mu = states.T
    x = np.stack([np.log10(I), phi, np.log10(P), np.log10(np.abs(T) + 1e-12)], axis=1)
    x = (x - np.array([-1.6, 0.55, 0.6, -3.5])) / np.array([0.8, 0.05, 0.4, 1.5])
    t = np.stack([(mu - 0.5) / 0.1, (phi - 0.55) / 0.05], axis=1)
    return x, t
