"""The two adjoint workloads at full size on the GPU.

adjoint128 (128x128x64, 1 mm, CPML 10, one source, design 32x32x16 with
eps_r ~ U[1,4), J = |DFT_Ez(probe, 10 GHz)|^2, FP32, 1000 steps,
checkpoints every 64 steps): the gradient against the committed FP64 oracle
fixture tests/golden/adjoint128.npz -- the FP64 CPU adjoint over all 16384
design cells and FP64 central differences (h = 1e-3) on 16 seeded cells.

adjoint768 (768x768x384, 0.5 mm, CPML 16, 4 sources, design 128x128x64 with
eps_r ~ U[1,6), 8 probes x {5,10,15,20} GHz, FP16 storage, 4000 steps,
checkpoints every 64 steps): the FP16-storage gradient against the FP32
engine's stored-history gradient, and both against a central difference at
h = 1e-3 of FP32 forward runs along sign(g) over the whole design region.  A
single cell's difference at h = 1e-3 is below the FP32 forwards' rounding of
J (relative 1e-5); tests/reports.py adjoint --workload adjoint768 --fd-h ...
states that floor.
"""
import os
import sys

import numpy as np
import pytest

from fdtd_amd import scenes

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_golden import adjoint_terms  # noqa: E402


def test_adjoint128_full_size_against_fp64_fixture(engine):
    fx = np.load(os.path.join(HERE, "golden", "adjoint128.npz"))
    sc = scenes.adjoint128()
    assert int(fx["steps"]) == sc.steps and int(fx["seed"]) == sc.seed
    terms = adjoint_terms(sc)
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    fwd = sim.info()["last_run_ms"]
    res = {}
    for K in (64, 0):
        sim.reset()
        J, g = sim.adjoint_gradient(terms, sc.steps, K)
        inf = sim.info()
        res[K] = (J, g)
        print(f"[adjoint128] checkpoint_every={K}: adjoint {inf['adjoint_total_ms']:.1f} ms, forward {fwd:.1f} ms, "
              f"ratio {inf['adjoint_total_ms'] / fwd:.2f}, segments {inf['adjoint_segments']}", flush=True)
    sim.close()
    J, g = res[64]
    assert res[0][0] == J
    np.testing.assert_array_equal(res[0][1], g)
    g64 = fx["grad"]
    rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    gmax = np.abs(g64).max()
    fd_err = np.abs(g[fx["fd_cells"]] - fx["fd"]) / gmax
    print(f"[adjoint128] J rel {abs(J - fx['J']) / fx['J']:.2e}, gradient rel L2 vs FP64 adjoint {rel:.2e}, "
          f"FD (h=1e-3, 16 cells) max err {fd_err.max():.2e} of max|g|", flush=True)
    assert abs(J - float(fx["J"])) <= 1e-4 * float(fx["J"])
    assert rel <= 1e-3
    assert fd_err.max() <= 1e-3


def test_adjoint768_fp16_against_fp32(engine):
    sc = scenes.adjoint768()
    terms = adjoint_terms(sc)
    s16 = engine.Simulation.from_scene(sc, "f16")
    s16.run(sc.steps)
    fwd = s16.info()["last_run_ms"]
    s16.reset()
    J16, g16 = s16.adjoint_gradient(terms, sc.steps, 64)
    inf = s16.info()
    print(f"[adjoint768] FP16, checkpoints every 64: adjoint {inf['adjoint_total_ms']:.0f} ms, forward {fwd:.0f} ms, "
          f"ratio {inf['adjoint_total_ms'] / fwd:.2f}, peak {inf['adjoint_device_bytes'] / 1e9:.1f} GB", flush=True)
    s16.close()
    s32 = engine.Simulation.from_scene(sc, "f32")
    J32, g32 = s32.adjoint_gradient(terms, sc.steps, 0)
    rel = np.linalg.norm(g16 - g32) / np.linalg.norm(g32)
    print(f"[adjoint768] J FP16 {J16:.6e} FP32 {J32:.6e}; gradient rel L2 FP16 vs FP32 {rel:.2e}", flush=True)
    assert abs(J16 - J32) <= 1e-2 * J32
    assert rel <= 1e-3
    # Directional central difference at h = 1e-3 of FP32 forwards along
    # sign(g) over all 1,048,576 design cells: the change in J is a sum over
    # the region, far above the FP32 forwards' rounding of J (a single cell's
    # difference at h = 1e-3 is not).
    eps = sc.design_eps.astype(np.float32).ravel()
    v = np.sign(g32).astype(np.float32)
    vals, es = [], []
    for sgn in (1, -1):
        e2 = (eps + np.float32(sgn * 1e-3) * v).astype(np.float32)
        es.append(e2.astype(np.float64))
        s32.set_design_eps(e2)
        s32.run(sc.steps)
        X = s32.dft()
        vals.append(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
    s32.close()
    fd = vals[0] - vals[1]
    e32 = abs(fd - np.dot(g32, es[0] - es[1])) / abs(fd)
    e16 = abs(fd - np.dot(g16, es[0] - es[1])) / abs(fd)
    print(f"[adjoint768] directional FD (h = 1e-3, sign(g)): FP32 adjoint {e32:.2e}, FP16 adjoint {e16:.2e}", flush=True)
    assert e32 <= 1e-3 and e16 <= 1e-3
