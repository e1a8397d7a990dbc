"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle still reproduces them (regression pin of the checker).
GPU: the engine's FP32 run matches them within the FP32 parity tolerance
(1e-5 relative L2 on every significant probe series, DFTs likewise)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_golden import fixture_scenes  # noqa: E402

NAMES = sorted(fixture_scenes())


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def significant(ref):
    """Probe rows carrying signal: the rest are compared in absolute terms."""
    scale = np.abs(ref).max()
    return np.linalg.norm(ref, axis=-1) >= 1e-3 * scale * np.sqrt(ref.shape[-1])


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    from oracle.oracle import Oracle
    sc = fixture_scenes()[name]
    g = load(name)
    o = Oracle(sc, 64)
    o.steps(sc.steps)
    assert int(g["steps"]) == sc.steps
    np.testing.assert_allclose(o.probes(), g["probes"], rtol=1e-12, atol=1e-12 * np.abs(g["probes"]).max())
    if len(sc.freqs):
        np.testing.assert_allclose(o.dft(), g["dft"], rtol=1e-12, atol=1e-12 * np.abs(g["dft"]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_fixture(name, engine):
    sc = fixture_scenes()[name]
    g = load(name)
    print(f"[golden] fixture={name} grid={sc.nx}x{sc.ny}x{sc.nz} steps={sc.steps}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    got = sim.probes()[:, :sc.steps].astype(np.float64)
    ref = g["probes"]
    sig = significant(ref)
    err = np.linalg.norm(got - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-300)
    assert np.all(err[sig] <= 1e-5), err
    assert np.abs(got[~sig] - ref[~sig]).max(initial=0.0) <= 1e-5 * np.abs(ref).max()
    if len(sc.freqs):
        X, R = sim.dft(), g["dft"]
        derr = np.abs(X - R) / np.abs(R).max()
        assert derr.max() <= 1e-5, derr.max()
    sim.close()
