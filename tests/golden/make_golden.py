"""Write the golden fixtures: FP64 oracle probe series and DFTs of small
seeded scenes.  They pin the oracle against regressions (the reference has no
fixtures of its own) and give the GPU tests a fixed target.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fdtd_amd import scenes  # noqa: E402


def fixture_scenes():
    """Name -> scene of every committed fixture."""
    return {
        "dielectric40": scenes.dielectric(nx=40, ny=36, nz=32, steps=120, L=6, n_boxes=5, n_src=2, n_probes=6,
                                          n_freqs=4, seed=7, name="dielectric40"),
        "yee32": scenes.yee64(n=32, steps=100, L=6, n_probes=4),
    }


def main():
    from oracle.oracle import Oracle
    for name, sc in fixture_scenes().items():
        o = Oracle(sc, 64)
        o.steps(sc.steps)
        path = os.path.join(HERE, name + ".npz")
        np.savez(path, probes=o.probes(), dft=o.dft(), freqs=np.array(sc.freqs, dtype=np.float64),
                 steps=np.int64(sc.steps))
        print("wrote", path)


if __name__ == "__main__":
    main()
