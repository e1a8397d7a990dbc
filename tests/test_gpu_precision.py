"""Scaled FP16/BF16 storage on the GPU against the oracle's storage-rounding
model (the FP32 oracle with every stored E, H and psi value rounded to the
storage type at the engine's scales) and against the FP64 oracle.

The engine folds the scales into its coefficients, so its FP32 arithmetic
differs from the model's in the last bit; an occasional storage rounding then
goes the other way and starts a different random walk.  Engine and model are
therefore two realisations of the same rounding process: their errors against
FP64 must agree in size probe by probe, which shows the engine's error is the
storage type's and not a defect, and that size is what is reported."""
import math

import numpy as np
import pytest

from fdtd_amd import scenes

pytestmark = pytest.mark.gpu


def storage_scales(sc):
    """The engine's stored = SI * scale for E and H (E as E/eta0 with a
    common gain that puts the source-gap field V/d at 2^11)."""
    vmax = max([abs(s.amp) for s in sc.sources] or [1.0]) or 1.0
    dmin = min(sc.dx, sc.dy, sc.dz)
    eta0 = math.sqrt(sc.mu0 / sc.eps0)
    gain = 2048.0 / ((vmax / dmin) / eta0)
    return float(np.float32(gain / eta0)), float(np.float32(gain))


def rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


@pytest.mark.parametrize("storage,mode", [("f16", 16), ("bf16", 17)])
def test_scaled_storage_matches_rounding_model(engine, storage, mode):
    from oracle.oracle import Oracle
    sc = scenes.dielectric(nx=40, ny=36, nz=32, steps=200, L=6, n_boxes=6, n_src=2, n_probes=8, n_freqs=3,
                           seed=21, name="precision-small")
    print(f"[precision] storage={storage} grid={sc.nx}x{sc.ny}x{sc.nz} steps={sc.steps}", flush=True)
    sim = engine.Simulation.from_scene(sc, storage)
    sim.run(sc.steps)
    got = sim.probes()[:, :sc.steps].astype(np.float64)
    sim.close()
    o64 = Oracle(sc, 64)
    o64.steps(sc.steps)
    ref = o64.probes()
    model = Oracle(sc, 32)
    sE, sH = storage_scales(sc)
    model.set_quant(mode, sE, sH)
    model.steps(sc.steps)
    mod = model.probes()
    sig = np.linalg.norm(ref, axis=-1) >= 1e-3 * np.abs(ref).max() * math.sqrt(ref.shape[1])
    e_store = rel(got, ref)[sig]
    e_model = rel(got, mod)[sig]
    e_model_store = rel(mod, ref)[sig]
    print(f"[precision] {storage}: engine vs FP64 max {e_store.max():.3e}; model vs FP64 max "
          f"{e_model_store.max():.3e}; engine vs model max {e_model.max():.3e}", flush=True)
    # same-sized storage error, probe by probe and overall ...
    ratio = e_store / e_model_store
    assert np.all((ratio > 0.4) & (ratio < 2.5)), ratio
    assert 0.6 < np.median(ratio) < 1.6, ratio
    # ... of the size the storage type gives over 200 steps
    assert e_store.max() < (6e-3 if storage == "f16" else 5e-2)
    # a defect (a wrong scale, a missed rounding) would be far off: the
    # engine is much closer to the FP64 result than to a run with the H
    # scale off by 2x
    assert e_model.max() < 4 * e_store.max()
