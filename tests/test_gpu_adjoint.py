"""GPU discrete adjoint: against the FP64 CPU adjoint, between checkpoint
intervals, and against central finite differences of GPU forward runs."""
import numpy as np
import pytest

from fdtd_amd import scenes

pytestmark = pytest.mark.gpu


def small_design(**kw):
    base = dict(nx=28, ny=24, nz=22, steps=160, L=4, d=1e-3, design=(6, 5, 4), eps_range=(1.0, 4.0),
                seed=9, probe_freqs=(12e9, 20e9), name="adjoint-small")
    base.update(kw)
    return scenes.adjoint(**base)


TERMS = [(0, 0, 1.0), (0, 1, 0.5)]


def objective(sim):
    X = sim.dft()
    return sum(w * abs(X[p, f]) ** 2 for p, f, w in TERMS)


def wide_sources(sc):
    """The first source widened to 3 x 2 x 2 Ez edges (the transposed source
    coupling then meets source edges on the cell itself and at i+1, j+1 of
    its neighbours), plus a single edge inside the x-CPML slab (no kappa on
    its coefficients)."""
    s0 = sc.sources[0]
    i, j, k = s0.ijk[:3]
    sc.sources[0] = scenes.Source((i, j, k, i + 2, j + 1, k + 2), 50.0, s0.wtype, s0.amp, s0.t0, s0.tau, s0.freq)
    sc.sources.append(scenes.z_edge(2, sc.ny // 2, sc.nz // 2, sc.dt))
    return sc


@pytest.mark.parametrize("sources", ["single", "wide"])
def test_adjoint_matches_fp64_oracle(engine, sources):
    from oracle.oracle import Oracle
    sc = small_design()
    if sources == "wide":
        sc = wide_sources(sc)
    print(f"[adjoint] {sc.nx}x{sc.ny}x{sc.nz} steps={sc.steps} sources={sources}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    J, g = sim.adjoint_gradient(TERMS, sc.steps, 0)
    Jo, go = Oracle(sc, 64).adjoint(TERMS)
    assert abs(J - Jo) / Jo < 1e-4
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"[adjoint] J gpu {J:.6e} fp64 {Jo:.6e}; gradient rel L2 {rel:.2e}", flush=True)
    assert rel < 1e-3


def test_checkpointed_equals_stored_history(engine):
    """Segments are recomputed on one stream while the previous segment's
    reverse sweep runs on another, through two history buffers: bitwise the
    stored-history result for uneven, two-segment, one-step-tail and
    one-step segmentations.  checkpoint_every < 0 picks the stored history
    on a scene this small (one segment, same bits)."""
    sc = small_design(steps=150)
    a = engine.Simulation.from_scene(sc, "f32")
    J0, g0 = a.adjoint_gradient(TERMS, sc.steps, 0)
    assert a.info()["adjoint_segments"] == 1
    Ja, ga = a.adjoint_gradient(TERMS, sc.steps, -1)
    assert a.info()["adjoint_segments"] == 1
    assert Ja == J0
    np.testing.assert_array_equal(ga, g0)
    for K, nseg in ((32, 5), (75, 2), (149, 2), (1, 150)):
        print(f"[adjoint] checkpoint every {K} steps", flush=True)
        J1, g1 = a.adjoint_gradient(TERMS, sc.steps, K)
        assert a.info()["adjoint_segments"] == nseg
        assert J0 == J1
        np.testing.assert_array_equal(g0, g1)


def test_adjoint_vs_central_differences(engine):
    sc = small_design()
    sim = engine.Simulation.from_scene(sc, "f32")
    _, g = sim.adjoint_gradient(TERMS, sc.steps, 0)
    eps = sc.design_eps.astype(np.float32)
    rng = np.random.default_rng(4)
    cells = rng.choice(eps.size, 4, replace=False)
    for q in cells:
        vals, ev = [], []
        for sgn in (1, -1):
            e2 = eps.copy()
            e2[q] = np.float32(eps[q] + sgn * 5e-2)
            ev.append(float(e2[q]))
            sim.set_design_eps(e2)
            sim.run(sc.steps)
            vals.append(objective(sim))
        fd = (vals[0] - vals[1]) / (ev[0] - ev[1])
        print(f"[adjoint] cell {q}: adjoint {g[q]:.6e} fd {fd:.6e}", flush=True)
        assert abs(g[q] - fd) <= 2e-3 * np.abs(g).max()


def test_torch_autograd_binding(engine):
    """The torch.autograd.Function returns the run's J and, on backward, the
    engine's adjoint gradient."""
    import torch
    from fdtd_amd.autograd import design_objective
    sc = small_design(steps=120)
    sim = engine.Simulation.from_scene(sc, "f32")
    eps = torch.tensor(sc.design_eps, dtype=torch.float64, requires_grad=True)
    J = design_objective(eps, sim, TERMS, sc.steps, 32)
    J.backward()
    sim2 = engine.Simulation.from_scene(sc, "f32")
    J2, g2 = sim2.adjoint_gradient(TERMS, sc.steps, 32)
    assert abs(float(J.detach()) - J2) <= 1e-6 * J2
    np.testing.assert_array_equal(eps.grad.numpy(), g2)


@pytest.mark.parametrize("nx", [200, 530])
def test_adjoint_interior_split(engine, monkeypatch, nx):
    """Grids wide enough for interior adjoint tiles (nx >= 512 also takes the
    LDS-sharing path): the split sweeps equal the general variant alone bit for
    bit, and the gradient matches the FP64 CPU adjoint."""
    from oracle.oracle import Oracle
    sc = small_design(nx=nx, ny=26, nz=24, steps=100, design=(8, 6, 5))
    print(f"[adjoint split] {sc.nx}x{sc.ny}x{sc.nz} steps={sc.steps}", flush=True)
    J0, g0 = engine.Simulation.from_scene(sc, "f32").adjoint_gradient(TERMS, sc.steps, 0)
    monkeypatch.setenv("FDTD_INTERIOR", "0")
    J1, g1 = engine.Simulation.from_scene(sc, "f32").adjoint_gradient(TERMS, sc.steps, 0)
    monkeypatch.delenv("FDTD_INTERIOR")
    assert J0 == J1
    np.testing.assert_array_equal(g0, g1)
    Jo, go = Oracle(sc, 64).adjoint(TERMS)
    rel = np.linalg.norm(g0 - go) / np.linalg.norm(go)
    print(f"[adjoint split] gradient rel L2 vs FP64 {rel:.2e}", flush=True)
    assert rel < 1e-3


@pytest.mark.parametrize("K,chunks", [(0, None), (32, "5"), (0, "6")])
def test_adjoint_fused_interior_bitwise(engine, monkeypatch, K, chunks):
    """The fused interior reverse step (E and H sweeps in one march over the
    tiles with no source, probe or design cell, reading one buffer of a
    pair and writing the other) gives the separate sweeps' gradient bit for
    bit, stored history and checkpointed, with z-blocks of 1, 8 and 7
    planes; the widened source puts source edges at i+1 and j+1 of some
    tiles' cells."""
    sc = wide_sources(small_design(nx=530, ny=40, nz=40, steps=90, design=(8, 6, 5)))
    if chunks:
        monkeypatch.setenv("FDTD_ADJ_CHUNKS", chunks)
    out = {}
    for fuse in ("1", "0"):
        monkeypatch.setenv("FDTD_ADJ_FUSE", fuse)
        print(f"[adjoint fused] {sc.nx}x{sc.ny}x{sc.nz} K={K} chunks={chunks} fuse={fuse}", flush=True)
        sim = engine.Simulation.from_scene(sc, "f32")
        out[fuse] = sim.adjoint_gradient(TERMS, sc.steps, K)
        sim.close()
    assert out["1"][0] == out["0"][0]
    np.testing.assert_array_equal(out["1"][1], out["0"][1])


@pytest.mark.parametrize("nx", [200, 530])
def test_adjoint_gather_paths_bitwise(engine, monkeypatch, nx):
    """The two forms of the adjoint sweeps -- every cell's transposed
    quantities stored by the sweep that updates it (narrow grids) and
    computed where they are gathered, shared through LDS (wide grids) -- give
    the same gradient bit for bit, checkpointed and not."""
    sc = small_design(nx=nx, ny=26, nz=24, steps=100, design=(8, 6, 5))
    out = {}
    for path in ("stored", "lds"):
        monkeypatch.setenv("FDTD_ADJ_GATHER", path)
        print(f"[adjoint gather] {sc.nx}x{sc.ny}x{sc.nz} path={path}", flush=True)
        sim = engine.Simulation.from_scene(sc, "f32")
        out[path] = [sim.adjoint_gradient(TERMS, sc.steps, K) for K in (0, 32)]
        sim.close()
    for (J0, g0), (J1, g1) in zip(out["stored"], out["lds"]):
        assert J0 == J1
        np.testing.assert_array_equal(g0, g1)


@pytest.mark.parametrize("storage,gen16", [("f32", None), ("f16", None), ("f16", "pair")])
def test_light_cone_recompute_is_bitwise(engine, monkeypatch, storage, gen16):
    """Checkpointed segments recomputed over the snapshot box dilated by the
    segment length plus two at a segment's first step and a cell less at
    each later one (a small design region in a large grid: 4 % of the cells)
    give the stored-history gradient, the fixed box's and the whole-grid
    recompute's bit for bit (FP32 one-cell sweeps; FP16 pair sweeps with the
    one-cell or their own general variant over the light-cone tiles)."""
    if gen16:
        monkeypatch.setenv("FDTD_GEN16", gen16)
    sc = scenes.adjoint(nx=140, ny=120, nz=100, steps=150, L=8, d=1e-3, design=(6, 5, 4), eps_range=(1.0, 4.0),
                        seed=31, probe_freqs=(12e9,), name="adjoint-lightcone")
    terms = [(0, 0, 1.0)]
    print(f"[adjoint] light cone {sc.nx}x{sc.ny}x{sc.nz} storage={storage} gen16={gen16}", flush=True)
    sim = engine.Simulation.from_scene(sc, storage)
    J0, g0 = sim.adjoint_gradient(terms, sc.steps, 0)
    J1, g1 = sim.adjoint_gradient(terms, sc.steps, 16)
    assert sim.info()["adjoint_segments"] == 10
    monkeypatch.setenv("FDTD_LIGHTCONE", "0")
    J2, g2 = sim.adjoint_gradient(terms, sc.steps, 16)
    monkeypatch.setenv("FDTD_LIGHTCONE", "fixed")
    J3, g3 = sim.adjoint_gradient(terms, sc.steps, 16)
    sim.close()
    assert np.abs(g0).max() > 0
    assert J0 == J1 == J2 == J3
    np.testing.assert_array_equal(g1, g0)
    np.testing.assert_array_equal(g2, g0)
    np.testing.assert_array_equal(g3, g0)
