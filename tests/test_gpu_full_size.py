"""GPU parity at the workloads' full grid sizes.

The small scenes of test_gpu_parity.py reach every code path; these runs put
the same kernels at the real shapes (64^3, 256^3 and 512x512x256 with CPML 12
and 16), where the interior, x-edge and general tile variants split the grid
as they do in the benchmark.  FP32: bit for bit against the oracle's FP32
restatement on probes, DFT and every field, for enough steps that the
wavefronts of the seeded sources reach the CPML slabs (c dt / dx = 0.57 cells
per step).  Storage forms: the full 3000-step slab512 run of FP32, FP16 and
block-scaled int16 storage against the committed FP64 fixture, each inside its
regression band, and the 1e-3 target of the 2-byte forms as an expected
failure (the gap stays visible in every GPU test report).
"""
import gc

import numpy as np
import pytest

from fdtd_amd import scenes

pytestmark = pytest.mark.gpu

FIELDS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


@pytest.mark.parametrize("workload,steps", [("yee64", 500), ("dielectric256", 400), ("slab512", 250)])
def test_full_size_fp32_bitwise(workload, steps, engine):
    from oracle.oracle import Oracle
    sc = scenes.WORKLOADS[workload](steps=steps)
    print(f"[full-size] workload={workload} grid={sc.nx}x{sc.ny}x{sc.nz} steps={steps}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(steps)
    gp, gd = sim.probes(), (sim.dft() if sc.freqs else None)
    o32 = Oracle(sc, 32)
    o32.steps(steps)
    op = o32.probes()
    assert gp.shape == op.shape
    mism = np.flatnonzero(gp.astype(np.float64).ravel() != op.ravel())
    assert mism.size == 0, f"{mism.size} probe samples differ, first at {mism[:5]}"
    assert np.count_nonzero(gp) > 0
    if gd is not None:
        np.testing.assert_array_equal(gd, o32.dft())
    for c, nm in enumerate(FIELDS):
        g = sim.field(c)
        o = o32.field(nm).astype(np.float32)
        bad = np.flatnonzero(g.ravel() != o.ravel())
        assert bad.size == 0, f"{workload} {nm}: {bad.size} cells differ, max |d| = {np.abs(g - o).max()}"
        # the wavefronts reached the x-, y- and z-CPML slabs
        L = sc.cpml.thick
        if c == 2:
            assert np.abs(g[:, :, :L[0]]).max() > 0 and np.abs(g[:, :L[2], :]).max() > 0
            assert np.abs(g[:L[4], :, :]).max() > 0
        del g, o
        gc.collect()
    sim.close()
    o32.close()
    if workload == "yee64":
        # and the FP64 restatement within 1e-5 on every probe series
        o64 = Oracle(sc, 64)
        o64.steps(steps)
        p64 = o64.probes()
        o64.close()
        r = np.linalg.norm(gp - p64, axis=1) / np.linalg.norm(p64, axis=1)
        assert r.max() <= 1e-5, r.max()


_ERRORS = {}


def _errors_vs_fp64(engine):
    """slab512 at full length (3000 steps) in every storage form against the
    committed FP64 oracle fixture: per-probe relative L2 over the probes that
    carry signal, and the DFT error over each probe's spectral peak.  Computed
    once per session."""
    if _ERRORS:
        return _ERRORS
    import os
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slab512_fp64.npz"))
    ref, X64 = fx["probes"], fx["dft"]
    sc = scenes.slab512()
    assert ref.shape == (len(sc.probes), sc.steps)
    n = np.linalg.norm(ref, axis=1)
    sig = n >= 1e-3 * n.max()
    for storage in ("f32", "f16", "bq16"):
        print(f"[full-size] slab512 {storage}, {sc.steps} steps against the FP64 fixture", flush=True)
        sim = engine.Simulation.from_scene(sc, storage)
        sim.run(sc.steps)
        p = sim.probes()[:, :sc.steps].astype(np.float64)
        X = sim.dft()
        sim.close()
        r = np.linalg.norm(p - ref, axis=1)[sig] / n[sig]
        d = (np.abs(X - X64).max(axis=1) / np.abs(X64).max(axis=1))[sig]
        print(f"[full-size] {storage}: probe rel L2 max {r.max():.3e} median {np.median(r):.3e}, DFT {d.max():.3e}",
              flush=True)
        _ERRORS[storage] = (r, d)
    return _ERRORS


# Regression bands (what each form measures at 3000 steps, with margin), not
# the 1e-3 accuracy target of the 2-byte forms: FP32 2.7e-5 max; FP16 2.1e-2 max / 1.2e-2 median;
# block-scaled int16 2.6e-3 / 1.5e-3.  A lost scale or a missed rounding is
# orders of magnitude off.
BANDS = {"f32": (1e-4, 5e-5, 2e-4), "f16": (5e-2, 2.5e-2, 4e-2), "bq16": (5e-3, 3e-3, 5e-3)}


@pytest.mark.parametrize("storage", ["f32", "f16", "bq16"])
def test_full_size_storage_error_band(engine, storage):
    """slab512 (512x512x256, CPML 16, 256 boxes, 8 sources, 128 probes, 16
    DFT frequencies) for its full 3000 steps, each storage form against FP64:
    max and median probe error and the DFT error inside the form's
    regression band."""
    r, d = _errors_vs_fp64(engine)[storage]
    assert r.size >= 64
    mx, med, dft = BANDS[storage]
    assert r.max() < mx and np.median(r) < med and d.max() < dft


@pytest.mark.parametrize("storage", ["f16", "bq16"])
@pytest.mark.xfail(reason="2-byte storage misses the 1e-3 probe-error target at 3000 steps: each stored value is "
                          "rounded once per step and the roundings accumulate as a random walk (worst probe: FP16 "
                          "2.1e-2, block-scaled int16 2.6e-3)", strict=False)
def test_16bit_storage_meets_the_1e3_target(engine, storage):
    """The accuracy target of the 2-byte storage forms: every probe within
    1e-3 relative L2 of FP64 at slab512's 3000 steps."""
    r, _ = _errors_vs_fp64(engine)[storage]
    assert r.max() <= 1e-3
