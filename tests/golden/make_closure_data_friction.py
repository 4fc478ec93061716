"""Writes tests/golden/closure_chute_friction.npz: training and test states of
the configured closure network over chute angle and particle friction, from
the engine's own chute flows.  Periodic chutes of 6000
HertzianViscoelasticMindlin spheres (d ~ U[0.8, 1.2], e = 0.8, t_c = 0.02, 10
diameters deep) at sliding friction 0.3, 0.45 and 0.6 and angles 20 to 28
degrees: each bed settled, tilted and run 400,000 steps, then 100
CG<Z, Gauss> profiles (64 points, width 0.5 d) 1000 steps apart, whose bulk
bins give states (I, phi, P, T, mu) (mercurydpm_fork_amd.closure_data.
profile_states); 600 states per run are kept (seeded draw).  Array names:
"mu0.30_angle21" and so on.  Needs a GPU (about 30 s per run):

    python tests/golden/make_closure_data_friction.py [out.npz]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mercurydpm_fork_amd import closure_data as CD  # noqa: E402
from mercurydpm_fork_amd import warmstart as W  # noqa: E402

FRICTIONS = (0.3, 0.45, 0.6)
# 24 degrees at 0.3 and 0.6, and friction 0.45 at 22 and 26 degrees, are the
# held-out cases of tests/test_closure.py; the others train
RUNS = [(mu, a) for mu in (0.3, 0.6) for a in (20.0, 21.0, 22.0, 23.0, 24.0, 25.0, 26.0, 27.0, 28.0)] + \
       [(0.45, 22.0), (0.45, 26.0)]
NZ, WIDTH, ZMAX = 64, 0.5, 16.0
KEEP = 600


def run(mu, angle, seed):
    sc = W.chute_case(angle, mu=mu)
    e = W.settled_engine(sc)
    e.set_gravity(sc.gravity)
    e.step(400_000)
    z = W.profile_z(ZMAX, NZ)
    out = []
    for _ in range(100):
        e.step(1000)
        f = W.cg_profile(e, sc, ZMAX, NZ, WIDTH)
        st = CD.profile_states(f, z)
        top = z[f[:, CD.COL["VolumeFraction"]] > 0.3].max()
        keep = (z > 2.0) & (z < top - 2.0) & (st[:, 1] > 0.4) & np.all(np.isfinite(st), axis=1) & \
            (st[:, 2] > 0) & (st[:, 0] > 0)
        out.append(st[keep])
    e.close()
    st = np.concatenate(out)
    rng = np.random.default_rng(seed)
    if len(st) > KEEP:
        st = st[np.sort(rng.choice(len(st), KEEP, replace=False))]
    return st.astype(np.float32)


def main(path):
    data = {}
    for k, (mu, a) in enumerate(RUNS):
        t = time.time()
        st = run(mu, a, 1000 + k)
        data[f"mu{mu:.2f}_angle{a:g}"] = st
        print(f"mu {mu:.2f} angle {a:g}: {len(st)} bulk states, mu_eff {np.median(st[:, 4]):.3f}, "
              f"I {np.median(st[:, 0]):.3g} ({time.time() - t:.1f} s)", flush=True)
    np.savez_compressed(path, **data)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "closure_chute_friction.npz"))
