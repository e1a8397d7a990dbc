"""GPU parity: the HIP engine against the CPU oracle on the same seeded scenes.

FP32 storage must reproduce the oracle's FP32 restatement of the reference
bit for bit (same operations, same order, no contraction) and the FP64
restatement within 1e-5 relative L2 on every probe series.
"""
import numpy as np
import pytest

from fdtd_amd import scenes
from fdtd_amd.scenes import EX, EY, EZ, HX, HY, HZ

pytestmark = pytest.mark.gpu

FIELDS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)


def scene_vacuum():
    sc = scenes.yee64(n=24, steps=120, L=5, n_probes=4, seed=11)
    # probes of every component, one inside the CPML and one on a face
    sc.probes += [(EX, 12, 12, 12), (EY, 12, 12, 12), (HX, 12, 12, 12), (HY, 13, 11, 12),
                  (HZ, 11, 13, 12), (EZ, 2, 12, 12), (EZ, 0, 5, 5)]
    return sc


def scene_dielectric():
    sc = scenes.dielectric(nx=40, ny=36, nz=32, steps=150, L=6, n_boxes=6, n_src=2, n_probes=8,
                           n_freqs=4, seed=21, name="dielectric-small")
    sc.cpml = scenes.graded_cpml((5, 6, 4, 7, 3, 5), sc.dx)
    return sc


def scene_magnetic():
    """Anisotropic and magnetic lossy boxes (classed H sweep), three waveform
    types, a two-edge source, the reference's default CPML grading."""
    d = 1e-3
    dt = scenes.courant_dt(d, d, d, 0.9)
    sc = scenes.Scene(name="magnetic", nx=30, ny=28, nz=26, dx=d, dy=d, dz=d, dt=dt, steps=140,
                      cpml=scenes.Cpml(thick=(5, 5, 5, 5, 5, 5)))
    sc.boxes = [scenes.Box((8, 7, 6, 20, 18, 16), (2.0, 3.0, 4.0), (0.01, 0.0, 0.02), (1.5, 2.0, 1.0),
                           (100.0, 0.0, 50.0)),
                scenes.Box((14, 12, 10, 24, 22, 20), (5.0, 5.0, 5.0), (0.0, 0.0, 0.0), (3.0, 3.0, 3.0),
                           (0.0, 0.0, 0.0))]
    sc.sources = [scenes.Source((10, 10, 9, 10, 10, 11), 50.0, scenes.WAVE_GAUSSIAN, 1.0, 30 * dt, 8 * dt),
                  scenes.Source((18, 16, 12, 18, 16, 13), 75.0, scenes.WAVE_SINE, 0.5, 0.0, 1.0, 20e9),
                  scenes.z_edge(12, 20, 14, dt)]
    sc.probes = [(EZ, 15, 14, 13), (HX, 15, 14, 13), (HY, 16, 12, 11), (HZ, 9, 9, 9), (EX, 20, 20, 20),
                 (EY, 7, 15, 18)]
    sc.freqs = [5e9, 12e9]
    return sc


def scene_design_u32():
    """A design region of per-cell eps_r large enough for 32-bit classes."""
    return scenes.adjoint(nx=56, ny=56, nz=44, steps=60, L=4, d=1e-3, design=(48, 48, 36),
                          eps_range=(1.0, 4.0), seed=31, name="design-u32")


def scene_wide():
    """Wide enough in x for an interior tile box in both the 64-column and
    the 128-column (16-bit pair) sweeps, so both kernel variants run; a
    source and probes inside the box, probes in the CPML too."""
    sc = scenes.dielectric(nx=300, ny=40, nz=36, steps=100, L=5, n_boxes=6, n_src=2, n_probes=8, n_freqs=2,
                           seed=41, name="wide")
    sc.sources.append(scenes.z_edge(150, 20, 18, sc.dt))
    sc.probes += [(EZ, 140, 20, 18), (EX, 100, 12, 10), (HY, 200, 30, 20), (EZ, 2, 20, 18), (HX, 298, 20, 18)]
    return sc


def scene_pec_faces():
    """CPML on three faces only (x-, y+ and z+ are bare outer faces)."""
    sc = scenes.dielectric(nx=30, ny=22, nz=18, steps=60, L=4, n_boxes=4, n_src=1, n_probes=6, n_freqs=2, seed=77,
                           name="pec-faces")
    sc.cpml = scenes.graded_cpml((0, 5, 4, 0, 3, 0), sc.dx)
    sc.probes += [(EX, 0, 10, 9), (EY, 15, 21, 9), (EZ, 15, 10, 17)]
    return sc


def scene_tiny():
    """9 x 7 x 6 cells, one-cell CPML: a single partial tile in x and y and a
    single z-block; probes on the outer faces and corners.  24 steps: the
    z-face CPML cells sit two cells from the source, and their H psi in
    scaled FP16 passes 65504 at step 29 (in the oracle's FP16 storage model
    as in the engine)."""
    sc = scenes.dielectric(nx=24, ny=24, nz=24, steps=24, L=1, n_boxes=1, n_src=1, n_probes=2, n_freqs=2, seed=78,
                           name="tiny")
    sc.nx, sc.ny, sc.nz = 9, 7, 6
    sc.cpml = scenes.graded_cpml(1, sc.dx)
    sc.boxes = [scenes.Box((2, 1, 1, 6, 5, 4), (4.0,) * 3, (0.002,) * 3)]
    sc.sources = [scenes.z_edge(4, 3, 2, sc.dt)]
    sc.probes = [(EZ, 4, 3, 2), (EX, 5, 3, 3), (HY, 3, 3, 2), (HZ, 0, 0, 0), (EY, 8, 6, 5)]
    return sc


def scene_thin_z():
    """Four planes along z (the smallest grid the engine takes) with no
    z-CPML (bare z faces), CPML of four different thicknesses in x and y."""
    sc = scenes.dielectric(nx=70, ny=66, nz=24, steps=60, L=4, n_boxes=3, n_src=1, n_probes=2, n_freqs=2, seed=79,
                           name="thin-z")
    sc.nz = 4
    sc.boxes = [scenes.Box((b.ijk[0], b.ijk[1], 0, b.ijk[3], b.ijk[4], 3), b.eps_r, b.sigma_e) for b in sc.boxes]
    sc.cpml = scenes.graded_cpml((4, 5, 6, 4, 0, 0), sc.dx)
    sc.sources = [scenes.z_edge(35, 33, 1, sc.dt)]
    sc.probes = [(EZ, 35, 33, 1), (EZ, 20, 20, 1), (HX, 40, 30, 1), (EY, 35, 62, 2), (HY, 36, 33, 2)]
    return sc


SCENES = {"vacuum": scene_vacuum, "dielectric": scene_dielectric, "magnetic": scene_magnetic,
          "design_u32": scene_design_u32, "wide": scene_wide, "pec_faces": scene_pec_faces, "tiny": scene_tiny,
          "thin_z": scene_thin_z}


@pytest.fixture(scope="module")
def oracle_mod():
    from oracle import oracle
    return oracle


@pytest.mark.parametrize("name", list(SCENES))
def test_fp32_bitwise_and_fp64_parity(name, engine, oracle_mod):
    sc = SCENES[name]()
    print(f"[parity] scene={name} grid={sc.nx}x{sc.ny}x{sc.nz} steps={sc.steps}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    info = sim.info()
    sim.run(sc.steps)
    gp = sim.probes()
    o32 = oracle_mod.Oracle(sc, 32)
    o32.steps(sc.steps)
    o64 = oracle_mod.Oracle(sc, 64)
    o64.steps(sc.steps)
    p32, p64 = o32.probes(), o64.probes()
    assert gp.shape == p32.shape
    # bit-exact against the FP32 restatement, probes and whole fields
    mism = np.flatnonzero(gp.astype(np.float64).ravel() != p32.ravel())
    assert mism.size == 0, f"{mism.size} probe samples differ, first at {mism[:5]}"
    for c, nm in enumerate(FIELDS):
        g = sim.field(c)
        o = o32.field(nm).astype(np.float32)
        bad = np.flatnonzero(g.ravel() != o.ravel())
        assert bad.size == 0, f"{nm}: {bad.size} cells differ, max |d| = {np.abs(g - o).max()}"
    # FP64 restatement: 1e-5 relative L2 on every probe series that carries
    # signal; a series that is zero by symmetry (e.g. Hz beside a centred
    # z-source) is held to 1e-5 of the largest series instead.
    n64 = np.linalg.norm(p64, axis=1)
    sig = n64 >= 1e-3 * n64.max()
    r = rel_l2(gp[sig], p64[sig])
    assert np.all(r <= 1e-5), f"relative L2 vs FP64 {r.max():.3e}"
    dn = np.linalg.norm(gp.astype(np.float64) - p64, axis=1)
    assert np.all(dn[~sig] <= 1e-5 * n64.max())
    if sc.freqs:
        gd, od = sim.dft(), o32.dft()
        np.testing.assert_array_equal(gd, od)
    if name == "design_u32":
        assert info["n_classes"] > 65536
    sim.close()


@pytest.mark.parametrize("chunks", [None, "1", "7"])
def test_sweep_chunking_bitwise(chunks, engine, oracle_mod, monkeypatch):
    """Any number of z-blocks per sweep gives the FP32 oracle's results bit
    for bit."""
    if chunks:
        monkeypatch.setenv("FDTD_CHUNKS", chunks)
    sc = scene_dielectric()
    print(f"[parity] chunks={chunks}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    o32 = oracle_mod.Oracle(sc, 32)
    o32.steps(sc.steps)
    np.testing.assert_array_equal(sim.probes().astype(np.float64), o32.probes())
    for c, nm in enumerate(FIELDS):
        np.testing.assert_array_equal(sim.field(c), o32.field(nm).astype(np.float32))


@pytest.mark.parametrize("interior,chunks", [(None, None), ("0", None), (None, "5"), (None, "36")])
def test_interior_split_and_z_blocks_bitwise(interior, chunks, engine, oracle_mod, monkeypatch):
    """The interior/general split of each sweep and any z-blocking give the
    FP32 oracle's results bit for bit (wide scene: the box is non-empty)."""
    if interior:
        monkeypatch.setenv("FDTD_INTERIOR", interior)
    if chunks:
        monkeypatch.setenv("FDTD_CHUNKS", chunks)
    sc = scene_wide()
    print(f"[parity] wide interior={interior} chunks={chunks}", flush=True)
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    o32 = oracle_mod.Oracle(sc, 32)
    o32.steps(sc.steps)
    np.testing.assert_array_equal(sim.probes().astype(np.float64), o32.probes())
    for c, nm in enumerate(FIELDS):
        np.testing.assert_array_equal(sim.field(c), o32.field(nm).astype(np.float32))


@pytest.mark.parametrize("storage", ["f32", "bq16"])
def test_resume_from_checkpoint_is_bitwise(engine, storage):
    """A run resumed from a checkpoint equals the uninterrupted run bit for
    bit (block-scaled storage: its exponent bytes travel in the blob)."""
    sc = scene_dielectric()
    a = engine.Simulation.from_scene(sc, storage)
    a.run(sc.steps)
    ref = a.probes()
    ref_ez = a.field(EZ)
    b = engine.Simulation.from_scene(sc, storage)
    b.run(60)
    blob = b.checkpoint_save()
    b.run(30)            # diverge, then rewind
    b.checkpoint_restore(blob)
    assert b.info()["steps_done"] == 60
    b.run(sc.steps - 60)
    np.testing.assert_array_equal(b.probes(), ref)
    np.testing.assert_array_equal(b.field(EZ), ref_ez)
    # a flipped byte is caught by the checksum
    bad = bytearray(blob)
    bad[len(bad) // 2] ^= 0x40
    with pytest.raises(engine.FdtdError) as ei:
        b.checkpoint_restore(bytes(bad))
    assert ei.value.code == -8


def test_reset_and_profile_reproduce_run(engine):
    sc = scene_vacuum()
    sim = engine.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    ref = sim.probes()
    sim.reset()
    sim.profile(sc.steps)
    np.testing.assert_array_equal(sim.probes(), ref)
    inf = sim.info()
    assert inf["step_kernel_ms"] > 0
    # the same run one fdtd_step at a time
    sim.reset()
    for _ in range(sc.steps):
        sim.step()
    assert sim.info()["steps_done"] == sc.steps
    np.testing.assert_array_equal(sim.probes(), ref)


@pytest.mark.parametrize("storage,nz", [("f32", 1022), ("f16", 2046)])
def test_grid_past_the_32bit_offsets_is_refused(storage, nz, engine):
    """A field array of 4 GiB or more (1024 x 1024 cells a plane, guard
    planes included) is refused at finalize, before anything is built; one
    plane less is accepted by the same check (not run: 24 GB of fields)."""
    sc = scenes.Scene(name="huge", nx=1024, ny=1024, nz=nz, dx=1e-3, dy=1e-3, dz=1e-3,
                      dt=scenes.courant_dt(1e-3, 1e-3, 1e-3, 0.9), steps=1, cpml=scenes.graded_cpml(8, 1e-3),
                      sources=[], probes=[], seed=0)
    print(f"[limits] {storage} 1024x1024x{nz}", flush=True)
    sim = engine.Simulation.from_scene(sc, storage, finalize=False)
    with pytest.raises(engine.FdtdError) as ei:
        sim.finalize(1)
    assert ei.value.code == -6 and "4 GiB" in str(ei.value)
    sim.close()


@pytest.mark.parametrize("fail_at", [1, 4, 9, 16])
def test_allocation_failure_is_reported_and_recoverable(engine, monkeypatch, fail_at):
    """A device allocation that fails during fdtd_finalize (forced from the
    host: FDTD_FAIL_ALLOC=n fails the context's n-th allocation) returns
    FDTD_E_NOMEM with the reason, releases what was allocated, and the same
    context finalizes and runs normally afterwards, bit for bit as a fresh one."""
    sc = scene_dielectric()
    ref = engine.Simulation.from_scene(sc, "f32")
    ref.run(sc.steps)
    want = ref.probes()
    ref.close()
    sim = engine.Simulation.from_scene(sc, "f32", finalize=False)
    monkeypatch.setenv("FDTD_FAIL_ALLOC", str(fail_at))
    with pytest.raises(engine.FdtdError) as ei:
        sim.finalize(sc.steps)
    assert ei.value.code == engine.FDTD_E_NOMEM and "injected" in str(ei.value)
    assert sim.info()["device_bytes"] == 0
    monkeypatch.delenv("FDTD_FAIL_ALLOC")
    sim.finalize(sc.steps)
    sim.run(sc.steps)
    np.testing.assert_array_equal(sim.probes(), want)
    sim.close()
