"""Warm starts of chute flows from the learned closure
(mercurydpm_fork_amd/warmstart.py).

* CPU: the profile prediction against its closed form.  With the local
  rheology mu(I) of Jop, Forterre & Pouliquen standing in for the network and
  a bed of uniform density, mu = tan(theta) gives one inertial number I* at
  every depth and the Bagnold profile
  u(z) = (2/3) I* sqrt(g cos(theta) phi / d^2) (H^(3/2) - (H - z)^(3/2)).
* GPU: the configured network (4 -> 3 x 128 GELU -> 2, BF16 weights, matrix
  cores) trained on the engine's own chute flows at 21-27 degrees
  (tests/golden/closure_chute128.npz, 24 degrees held out) predicts the steady
  flow at 24 degrees; DEM started from the settled bed in that flow reaches
  the steady CG velocity profile (within 5%, the profile of the same DEM run
  from rest to steady state being the reference) in at least 3x fewer steps than
  DEM started at rest, and stays within 5% of it.
"""
import os

import numpy as np
import pytest

from mercurydpm_fork_amd import closure, dem
from mercurydpm_fork_amd import closure_data as CD
from mercurydpm_fork_amd import warmstart as W

HERE = os.path.dirname(os.path.abspath(__file__))


class RheologyNet:
    """mu(I) and a constant phi in the network's standardised form."""

    def __init__(self, phi):
        self.phi = phi

    def __call__(self, x):
        I = 10.0 ** (x[:, 0] * 0.8 - 1.6)
        return np.stack([(closure.mu_of_I(I) - 0.5) / 0.1, np.full(len(I), (self.phi - 0.55) / 0.05)], axis=1)


def test_prediction_is_the_bagnold_profile():
    theta, phi, H, nz = 24.0, 0.58, 10.0, 256
    zmax = 12.0
    z = W.profile_z(zmax, nz)
    fields = np.zeros((nz, dem.CG_FIELDS))
    inside = z < H
    fields[:, CD.COL["Density"]] = np.where(inside, phi * W.RHO_P, 0.0)
    fields[:, CD.COL["VolumeFraction"]] = np.where(inside, phi, 0.0)
    tlaw = np.array([-1.0, 2.0, 1.0])
    p = W.predict_profile(RheologyNet(phi), tlaw, theta, fields, z, z_base=0.0)
    t = np.tan(np.radians(theta))
    mu1, mu2, I0 = np.tan(np.radians(20.9)), np.tan(np.radians(32.8)), 0.279
    istar = I0 * (t - mu1) / (mu2 - t)
    bed = inside & (z > 0.5) & (z < H - 0.5)
    np.testing.assert_allclose(p.inertial[bed], istar, rtol=1e-6)
    want = 2.0 / 3.0 * istar * np.sqrt(np.cos(np.radians(theta)) * phi) * (H ** 1.5 - np.clip(H - z, 0, None) ** 1.5)
    err = np.abs(p.velocity - want)[inside] / want[inside].max()
    assert err.max() < 0.01, err.max()


def steady_steps(errs, tol=0.05, run=3):
    """The step at which the mean error of `run` consecutive checkpoints
    first falls to tol."""
    for k in range(run - 1, len(errs)):
        if np.mean([e for _, e in errs[k - run + 1:k + 1]]) <= tol:
            return errs[k][0]
    return None


@pytest.mark.gpu
def test_warm_start_reaches_steady_flow_in_fewer_steps():
    d = np.load(os.path.join(HERE, "golden", "closure_chute128.npz"))
    states = np.concatenate([d[k] for k in d.files]).astype(np.float64)
    x, t = CD.standardise(states)
    net = closure.Closure(seed=3)
    net.train(x, t, epochs=2000, lr=1e-3)
    tlaw = W.temperature_law(states)
    zmax, nz, angle = 16.0, 64, 24.0
    z = W.profile_z(zmax, nz)
    sc = W.chute_case(angle)
    cold = W.settled_engine(sc)
    warm = W.settled_engine(sc)
    settled = W.cg_profile(cold, sc, zmax, nz)
    pred = W.predict_profile(net, tlaw, angle, settled, z, z_base=0.0)
    for e in (cold, warm):
        e.set_gravity(sc.gravity)
    W.warm_start(warm, sc, pred, settled[:, CD.COL["VolumeFraction"]], 0.75)

    def run(e, steps, every=10000):
        out = []
        for done in range(every, steps + 1, every):
            e.step(every - 4000)
            f = W.cg_profile(e, sc, zmax, nz, n_avg=5, every=1000)
            out.append((done, W.velocity_of(f), f[:, CD.COL["VolumeFraction"]]))
        return out

    cr = run(cold, 600_000)
    wr = run(warm, 200_000)
    ref = np.mean([u for _, u, _ in cr[-10:]], axis=0)
    phi = np.mean([p for _, _, p in cr[-10:]], axis=0)
    mask = (z > 1.5) & (z < z[phi > 0.3].max() - 1.0)
    ce = [(s, W.profile_error(u, ref, mask)) for s, u, _ in cr]
    we = [(s, W.profile_error(u, ref, mask)) for s, u, _ in wr]
    n_cold, n_warm = steady_steps(ce), steady_steps(we)
    late = np.mean([e for _, e in we[-10:]])
    print(f"warm start at {angle} deg: surrogate profile error {W.profile_error(pred.velocity, ref, mask):.3f}; "
          f"steady (5%) after {n_cold} steps from rest, {n_warm} warm; warm error over its last 100,000 steps "
          f"{late:.3f}")
    assert n_warm is not None and n_cold is not None
    assert n_cold >= 3 * n_warm, (n_cold, n_warm)
    assert late <= 0.05, late
