"""Writes tests/golden/closure_chute_cg.npz: bulk states (I, phi, P, T, mu)
of periodic HertzianViscoelasticMindlin chute flows at 21 to 27 degrees,
from CG depth profiles of the engine's own runs
(mercurydpm_fork_amd.closure_data.chute_samples: 4000 flowing spheres, 12
diameters deep, 150,000 steps, then 20 profiles 1000 steps apart).  Needs a
GPU:

    python tests/golden/make_closure_data.py [out.npz]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mercurydpm_fork_amd import closure_data as CD  # noqa: E402

ANGLES = (21.0, 22.0, 23.0, 24.0, 25.0, 26.0, 27.0)


def main(out):
    data = {}
    for a in ANGLES:
        t = time.time()
        st = CD.chute_samples(a)
        data[f"angle_{a:g}"] = st
        print(f"angle {a:g}: {len(st)} bulk states, mu {np.median(st[:, 4]):.3f}, I {np.median(st[:, 0]):.3g} "
              f"({time.time() - t:.1f} s)", flush=True)
    np.savez(out, **data)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                "closure_chute_cg.npz"))
