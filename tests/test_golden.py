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


def test_adjoint128_fixture_is_self_consistent():
    """The committed adjoint128 fixture (FP64 oracle): its adjoint gradient
    agrees with its own FP64 central differences on the 16 seeded cells to
    1e-5 of max |g| (the two are computed independently: a reverse sweep of
    the transposed step, and two forward runs per cell), the cells are the
    seeded ones, and every perturbed eps_r differs from the design value by
    h rounded to float32."""
    from fdtd_amd import scenes
    from make_golden import FD_H, fd_cells
    g = load("adjoint128")
    sc = scenes.adjoint128()
    assert int(g["steps"]) == sc.steps and int(g["seed"]) == sc.seed
    np.testing.assert_array_equal(g["fd_cells"], fd_cells(sc))
    gmax = np.abs(g["grad"]).max()
    assert np.abs(g["grad"][g["fd_cells"]] - g["fd"]).max() <= 1e-5 * gmax
    eps = sc.design_eps.astype(np.float32)[g["fd_cells"]]
    np.testing.assert_array_equal(g["eps_plus"], (eps + np.float32(FD_H)).astype(np.float64))
    np.testing.assert_array_equal(g["eps_minus"], (eps - np.float32(FD_H)).astype(np.float64))


def test_slab512_fixture_shape():
    """The committed FP64 fixture of the full slab512 run (the benchmark's
    accuracy target) covers every probe for the workload's 3000 steps and its
    16 DFT frequencies, and every probe carries the pulse."""
    from fdtd_amd import scenes
    g = load("slab512_fp64")
    sc = scenes.slab512()
    assert g["probes"].shape == (len(sc.probes), sc.steps)
    assert g["dft"].shape == (len(sc.probes), len(sc.freqs))
    np.testing.assert_array_equal(g["freqs"], np.array(sc.freqs))
    assert int(g["seed"]) == sc.seed
    assert np.all(np.isfinite(g["probes"])) and np.all(np.linalg.norm(g["probes"], axis=1) > 0)
