"""Writes tests/golden/closure_chute128.npz: training states of the
configured closure network from the engine's own chute flows.  Six periodic
chutes of 6000 HertzianViscoelasticMindlin spheres (d ~ U[0.8, 1.2], e = 0.8,
sliding friction 0.45, t_c = 0.02, 10 diameters deep) at 21, 22, 23, 25, 26
and 27 degrees (24 is held out for the warm-start test): each bed settled,
tilted and run 400,000 steps to steady flow, then 100 CG<Z, Gauss> profiles
(64 points, width 0.5 d) 1000 steps apart; every profile's bulk bins give
states (I, phi, P, T, mu) (mercurydpm_fork_amd.closure_data.profile_states).
Needs a GPU:

    python tests/golden/make_closure_data128.py [out.npz]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mercurydpm_fork_amd import closure_data as CD  # noqa: E402
from mercurydpm_fork_amd import warmstart as W  # noqa: E402

ANGLES = (21.0, 22.0, 23.0, 25.0, 26.0, 27.0)
NZ, WIDTH, ZMAX = 64, 0.5, 16.0


def run(angle):
    sc = W.chute_case(angle)
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
    return np.concatenate(out).astype(np.float32)


def main(path):
    data = {}
    for a in ANGLES:
        t = time.time()
        st = run(a)
        data[f"angle_{a:g}"] = st
        print(f"angle {a:g}: {len(st)} bulk states, mu {np.median(st[:, 4]):.3f}, I {np.median(st[:, 0]):.3g} "
              f"({time.time() - t:.1f} s)", flush=True)
    np.savez_compressed(path, **data)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "closure_chute128.npz"))
