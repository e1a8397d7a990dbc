"""Write the golden fixtures (FP64 oracle results).  They pin the oracle
against regressions (the reference has no fixtures of its own) and give the
GPU tests and the benchmark a fixed FP64 target at full size.

Run from the repository root:
  python tests/golden/make_golden.py            small scenes (seconds)
  python tests/golden/make_golden.py adjoint128 FP64 adjoint gradient and FP64
                                                central differences (h = 1e-3,
                                                16 seeded cells), minutes
  python tests/golden/make_golden.py slab512    FP64 probe series and DFTs of the
                                                3000-step 512x512x256 run (about
                                                36 GB of host memory, tens of
                                                minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fdtd_amd import scenes  # noqa: E402

FD_H = 1e-3          # central-difference step in eps_r (the adjoint check's h)
FD_CELLS = 16


def fixture_scenes():
    """Name -> scene of every small committed fixture."""
    return {
        "dielectric40": scenes.dielectric(nx=40, ny=36, nz=32, steps=120, L=6, n_boxes=5, n_src=2, n_probes=6,
                                          n_freqs=4, seed=7, name="dielectric40"),
        "yee32": scenes.yee64(n=32, steps=100, L=6, n_probes=4),
    }


def adjoint_terms(sc):
    """The objective of the adjoint workloads: sum over probes and
    frequencies of |DFT_Ez|^2, unit weights."""
    return [(p, f, 1.0) for p in range(len(sc.probes)) for f in range(len(sc.freqs))]


def fd_cells(sc, n=FD_CELLS):
    """The seeded design cells of the central-difference check."""
    return np.random.default_rng(sc.seed).choice(sc.design_eps.size, n, replace=False)


def _log(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def small():
    from oracle.oracle import Oracle
    for name, sc in fixture_scenes().items():
        o = Oracle(sc, 64)
        o.steps(sc.steps)
        path = os.path.join(HERE, name + ".npz")
        np.savez(path, probes=o.probes(), dft=o.dft(), freqs=np.array(sc.freqs, dtype=np.float64),
                 steps=np.int64(sc.steps))
        _log(f"wrote {path}")


def adjoint128():
    """FP64 oracle adjoint of the adjoint128 workload (J, dJ/d eps_r on all
    16384 design cells) and FP64 central differences on 16 seeded cells:
    eps_r -> float32(eps_r +- h), fd = (J+ - J-) / (eps+ - eps-)."""
    from oracle.oracle import Oracle
    sc = scenes.adjoint128()
    terms = adjoint_terms(sc)
    _log("adjoint128: FP64 adjoint")
    J, g = Oracle(sc, 64).adjoint(terms)
    cells = fd_cells(sc)
    eps = sc.design_eps.astype(np.float32)
    fd, lo, hi = [], [], []
    for q in cells:
        vals, ev = [], []
        for sgn in (1, -1):
            e2 = eps.copy()
            e2[q] = np.float32(eps[q] + sgn * FD_H)
            ev.append(float(e2[q]))
            s2 = scenes.adjoint128()
            s2.design_eps = e2
            o = Oracle(s2, 64)
            o.steps(s2.steps)
            X = o.dft()
            o.close()
            vals.append(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
        fd.append((vals[0] - vals[1]) / (ev[0] - ev[1]))
        hi.append(ev[0])
        lo.append(ev[1])
        _log(f"adjoint128: cell {q}: adjoint {g[q]:.6e} fd {fd[-1]:.6e}")
    path = os.path.join(HERE, "adjoint128.npz")
    np.savez_compressed(path, J=np.float64(J), grad=g, fd_cells=cells.astype(np.int64), fd=np.array(fd),
                        eps_plus=np.array(hi), eps_minus=np.array(lo), h=np.float64(FD_H),
                        steps=np.int64(sc.steps), seed=np.int64(sc.seed))
    _log(f"wrote {path}")


def slab512():
    """FP64 oracle probe series [probe][step] and DFTs [probe][freq] of the
    full slab512 workload (512x512x256, CPML 16, 256 boxes, 8 sources, 128
    probes, 16 frequencies, 3000 steps)."""
    from oracle.oracle import Oracle
    sc = scenes.slab512()
    t0 = time.perf_counter()
    o = Oracle(sc, 64)
    for s in range(0, sc.steps, 100):
        o.steps(min(100, sc.steps - s))
        _log(f"slab512: FP64 oracle at step {min(s + 100, sc.steps)} of {sc.steps}")
    path = os.path.join(HERE, "slab512_fp64.npz")
    np.savez_compressed(path, probes=o.probes(), dft=o.dft(), freqs=np.array(sc.freqs, dtype=np.float64),
                        steps=np.int64(sc.steps), seed=np.int64(sc.seed), host_s=np.float64(time.perf_counter() - t0))
    _log(f"wrote {path}")


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "small"
    {"small": small, "adjoint128": adjoint128, "slab512": slab512}[which]()
