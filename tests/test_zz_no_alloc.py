"""fdtd_adjoint_gradient_noalloc through the C ABI: the stored-history
adjoint on the library's static device scratch must equal
fdtd_adjoint_gradient(..., checkpoint_every = 0, ...) bit for bit, on two
calls in a row (the second on a smaller scene, after the first left its
contents in the scratch), on the full adjoint128 workload (1000 steps, the
largest run the scratch holds), and refuse a run that needs more scratch
than there is."""
import ctypes as C

import numpy as np
import pytest

from fdtd_amd import scenes

pytestmark = pytest.mark.gpu

TERMS = [(0, 0, 1.0), (0, 1, 0.5)]


def design_scene(**kw):
    base = dict(nx=28, ny=24, nz=22, steps=160, L=4, d=1e-3, design=(6, 5, 4), eps_range=(1.0, 4.0),
                seed=9, probe_freqs=(12e9, 20e9), name="adjoint-small")
    base.update(kw)
    return scenes.adjoint(**base)


def test_noalloc_equals_allocating_adjoint_twice(engine):
    big = design_scene()
    small = design_scene(nx=22, ny=20, nz=18, steps=90, L=3, design=(5, 4, 3), seed=11)
    refs, sims = [], []
    for sc in (big, small):
        sim = engine.Simulation.from_scene(sc, "f32")
        refs.append(sim.adjoint_gradient(TERMS, sc.steps, 0))
        sims.append(sim)
    # the two no-allocation calls back to back: the second, smaller one runs
    # on a scratch that still holds the first call's buffers
    for (sc, sim, (J0, g0)) in zip((big, small), sims, refs):
        print(f"[noalloc] {sc.nx}x{sc.ny}x{sc.nz}, {sc.steps} steps", flush=True)
        J1, g1 = sim.adjoint_gradient_noalloc(TERMS, sc.steps)
        assert J1 == J0
        np.testing.assert_array_equal(g1, g0)
        assert sim.info()["adjoint_segments"] == 1
    for sim in sims:
        sim.close()


def test_noalloc_adjoint128_and_size_limit(engine):
    sc = scenes.adjoint128()
    terms = [(p, f, 1.0) for p in range(len(sc.probes)) for f in range(len(sc.freqs))]
    sim = engine.Simulation.from_scene(sc, "f32", n_steps=1500)
    print("[noalloc] adjoint128, 1000 steps", flush=True)
    J0, g0 = sim.adjoint_gradient(terms, sc.steps, 0)
    J1, g1 = sim.adjoint_gradient_noalloc(terms, sc.steps)
    assert J1 == J0
    np.testing.assert_array_equal(g1, g0)
    # 1500 steps of design history do not fit: a clear error, nothing run
    print("[noalloc] adjoint128, 1500 steps (over the scratch)", flush=True)
    with pytest.raises(RuntimeError, match="static scratch holds"):
        sim.adjoint_gradient_noalloc(terms, 1500)
    # the context is still usable, and the scratch gives the same answer again
    J2, g2 = sim.adjoint_gradient_noalloc(terms, sc.steps)
    assert J2 == J0
    np.testing.assert_array_equal(g2, g0)
    sim.close()


def test_noalloc_through_the_c_abi_directly(engine):
    """The symbol as a plain C call (ctypes on the library), NULL arguments
    refused with FDTD_E_INVALID."""
    lib = engine.load_library()
    f = lib.fdtd_adjoint_gradient_noalloc
    assert f(None, None, 10, None, None, 0) == engine.FDTD_E_INVALID
    sc = design_scene(steps=60)
    sim = engine.Simulation.from_scene(sc, "f32")
    n = len(TERMS)
    obj = engine.FdtdObjective(n, (C.c_int * n)(*[t[0] for t in TERMS]), (C.c_int * n)(*[t[1] for t in TERMS]),
                               (C.c_double * n)(*[t[2] for t in TERMS]))
    J = C.c_double()
    g = np.zeros(sim.design_cells, dtype=np.float64)
    rc = f(sim.h, C.byref(obj), sc.steps, C.byref(J), g.ctypes.data_as(C.POINTER(C.c_double)), g.size)
    assert rc == engine.FDTD_OK
    J0, g0 = sim.adjoint_gradient(TERMS, sc.steps, 0)
    assert J.value == J0
    np.testing.assert_array_equal(g, g0)
    sim.close()
