"""Port voltage / current / S11 (the reference's results1003/1004.dat path).

CPU part: the oracle's restatement of voltMgpu / currMgpu / correcI /
c_zinS11 (oracle/port.py) pinned by what the discrete equations fix:
  * the current loop integral equals the displacement current through the
    loop (discrete Ampere law of the Yee update, in vacuum, FP64), for all
    three loop directions -- this fixes the loop's edges, plane and signs;
  * at a resistive lumped source the port voltage and current obey the
    source's own update (I = (Vs - V)/R minus the gap's displacement
    current) -- this fixes the voltage's sign and scale;
  * the voltage over a 2x2 cross-section is the mean of its four columns;
  * the spectrum: a resistive load Z gives Zin = Z and S11 = (Z-z0)/(Z+z0),
    and at f = k/(n dt) the DFT is numpy's rfft bin k.
GPU part: the engine's port series against the FP32 oracle's, bit for bit,
and its spectrum and result files.
"""
import ctypes as C
import os

import numpy as np
import pytest

from fdtd_amd import scenes
from oracle import port as oport
from oracle.oracle import Oracle

EPS0 = scenes.EPS0


def _vacuum(n=20, steps=90, L=4, src=(10, 10, 10), boxes=()):
    d = 1e-3
    dt = scenes.courant_dt(d, d, d, 0.99)
    return scenes.Scene(name="port_vacuum", nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=dt, steps=steps,
                        cpml=scenes.graded_cpml(L, d), sources=[scenes.z_edge(*src, dt)], probes=[],
                        boxes=list(boxes), seed=7)


# A dielectric block off to one side: around z-directed sources alone the
# field is TM to z (Hz = 0), which would leave the z-lines of the x- and
# y-directed loops unchecked.  The loops' enclosed edges stay in vacuum.
_BLOCK = scenes.Box(ijk=(11, 3, 5, 15, 8, 12), eps_r=(4.0, 4.0, 4.0), sigma_e=(0.0, 0.0, 0.0))


def _run_oracle(sc, cells, precision=64):
    sc.probes = list(cells)
    o = Oracle(sc, precision=precision)
    o.steps(sc.steps)
    hist = o.probes()
    lookup = {c: hist[q] for q, c in enumerate(cells)}
    return o, (lambda comp, i, j, k: lookup[(comp, i, j, k)])


def _ampere_edges(ibox, direction):
    """E edges the loop of currMgpu encloses, with the two cell sizes of
    the loop's plane (E component, edges, (da, db))."""
    is_, js, ks, ie, je, ke = ibox
    if direction == 3:
        return 2, [(i, j, ke - 1) for j in range(js, je + 1) for i in range(is_, ie + 1)], ("dx", "dy")
    if direction == 1:
        return 0, [(ie - 1, j, k) for k in range(ks, ke + 1) for j in range(js, je + 1)], ("dy", "dz")
    return 1, [(i, je - 1, k) for k in range(ks, ke + 1) for i in range(is_, ie + 1)], ("dz", "dx")


@pytest.mark.parametrize("direction,ibox", [(3, (7, 7, 12, 9, 9, 14)), (1, (7, 7, 7, 9, 9, 9)),
                                            (2, (7, 7, 7, 9, 9, 9)), (3, (6, 8, 13, 7, 10, 13))])
def test_port_current_is_the_discrete_ampere_law(direction, ibox):
    sc = _vacuum(boxes=[_BLOCK])
    comp, edges, (da, db) = _ampere_edges(ibox, direction)
    vbox = (10, 10, 10, 10, 10, 11)
    cells = oport.cells(vbox, ibox, direction, sc.nz) + [(comp, *e) for e in edges]
    cells = list(dict.fromkeys(cells))
    _, get = _run_oracle(sc, cells)
    raw = oport.current_raw(get, ibox, direction, sc.dx, sc.dy, sc.dz, sc.nz, dtype=np.float64)
    # H^{m+1/2} drives E^{m+1} -> E^{m+2}; probe m holds E^{m+1}
    dE = sum(np.diff(get(comp, *e)) for e in edges)
    disp = EPS0 * getattr(sc, da) * getattr(sc, db) / sc.dt * dE
    scale = np.max(np.abs(raw))
    assert scale > 0
    assert np.max(np.abs(raw[:-1] - disp)) < 1e-9 * scale
    # every H sample of the loop carries field (none is checked by a zero)
    h_lines = [c for c in cells if c[0] >= 3]
    assert min(np.max(np.abs(get(*c))) for c in h_lines) > 1e-6 * max(np.max(np.abs(get(*c))) for c in h_lines)


def test_port_voltage_and_current_at_a_resistive_source():
    sc = _vacuum()
    src = sc.sources[0]
    vbox = ibox = src.ijk            # the source edge, the loop around it
    cells = oport.cells(vbox, ibox, 3, sc.nz)
    o, get = _run_oracle(sc, cells)
    v = oport.voltage(get, vbox, sc.dz, sc.nz, dtype=np.float64)
    raw = oport.current_raw(get, ibox, 3, sc.dx, sc.dy, sc.dz, sc.nz, dtype=np.float64)
    vs = o.array("vpn")[:sc.steps]   # one edge: the whole source voltage
    R = src.R
    c_gap = EPS0 * sc.dx * sc.dy / sc.dz
    # source update (fdtdrav.h:1402-1421) as a current balance over the loop
    pred = -c_gap / sc.dt * np.diff(v) - (v[1:] + v[:-1]) / (2 * R) + vs[1:] / R
    scale = np.max(np.abs(raw))
    assert np.max(np.abs(raw[:-1] - pred)) < 1e-9 * scale
    # the pulse drives the gap: V follows Vs in sign
    m = int(np.argmax(np.abs(vs)))
    assert np.sign(v[m]) == np.sign(vs[m])


def test_port_voltage_averages_the_cross_section():
    sc = _vacuum()
    box = (9, 9, 8, 10, 10, 12)
    cols = [(i, j, 8, i, j, 12) for j in (9, 10) for i in (9, 10)]
    cells = []
    for b in [box] + cols:
        cells += oport.cells(b, (9, 9, 8, 10, 10, 12), 3, sc.nz)
    cells = list(dict.fromkeys(cells))
    _, get = _run_oracle(sc, cells)
    v = oport.voltage(get, box, sc.dz, sc.nz, dtype=np.float64)
    vc = np.mean([oport.voltage(get, b, sc.dz, sc.nz, dtype=np.float64) for b in cols], axis=0)
    assert np.max(np.abs(v - vc)) < 1e-12 * np.max(np.abs(v))
    assert np.max(np.abs(v)) > 0


def test_correcI_averages_neighbours():
    raw = np.array([1.0, 3.0, 7.0, 15.0], dtype=np.float32)
    assert oport.correct(raw).tolist() == [2.0, 5.0, 11.0, 11.0]
    assert oport.correct(raw[:1]).tolist() == [1.0]


def test_port_spectrum_closed_forms():
    rng = np.random.default_rng(3)
    n, dt, z0, Z = 64, 1e-12, 50.0, 75.0
    i = rng.standard_normal(n)
    v = Z * i
    k = np.arange(1, 8)
    sp = oport.spectrum(v, i, dt, z0, k / (n * dt))
    assert np.allclose(sp["Zin"], Z, rtol=1e-12)
    assert np.allclose(sp["S11"], (Z - z0) / (Z + z0), rtol=1e-12)
    assert np.allclose(sp["V"], np.fft.rfft(v)[k], rtol=1e-10, atol=1e-10 * np.abs(np.fft.rfft(v)).max())


# ---- GPU -----------------------------------------------------------------------

def _engine_port(sc, vbox, ibox, idir, storage="f32"):
    from fdtd_amd import api
    sim = api.Simulation.from_scene(sc, storage, finalize=False)
    sim.set_port(vbox, ibox, vdir=3, idir=idir)
    sim.finalize(sc.steps)
    sim.run(sc.steps)
    return sim


@pytest.mark.gpu
@pytest.mark.parametrize("idir,ibox", [(3, None), (1, (9, 8, 9, 11, 11, 12)), (2, (8, 9, 9, 11, 11, 12))])
def test_port_series_bitwise_vs_oracle(idir, ibox):
    sc = scenes.dielectric(nx=24, ny=22, nz=26, steps=120, L=4, n_boxes=6, n_src=1, n_probes=2,
                           n_freqs=2, seed=41)
    src = sc.sources[0]
    vbox = src.ijk
    ibox = ibox or src.ijk
    print(f"port case idir={idir} ibox={ibox} vbox={vbox}", flush=True)
    sim = _engine_port(sc, vbox, ibox, idir)
    v, i = sim.port()
    sim.close()
    cells = oport.cells(vbox, ibox, idir, sc.nz)
    _, get = _run_oracle(sc, cells, precision=32)
    vo = oport.voltage(get, vbox, sc.dz, sc.nz)
    io = oport.correct(oport.current_raw(get, ibox, idir, sc.dx, sc.dy, sc.dz, sc.nz))
    assert np.max(np.abs(vo)) > 0 and np.max(np.abs(io)) > 0
    assert np.array_equal(v, vo)
    assert np.array_equal(i, io)


@pytest.mark.gpu
def test_port_spectrum_and_result_files(tmp_path):
    sc = _vacuum(n=24, steps=200)
    src = sc.sources[0]
    sim = _engine_port(sc, src.ijk, src.ijk, 3)
    v, i = sim.port()
    freqs = np.array([2e9, 5e9, 10e9, 20e9])
    sp = sim.port_spectrum(50.0, freqs)
    ref = oport.spectrum(v, i, sc.dt, 50.0, freqs)
    for key in ("V", "I", "S11", "Zin"):
        assert np.allclose(sp[key], ref[key], rtol=1e-9, atol=1e-12 * np.abs(ref[key]).max()), key
    vp, ip = tmp_path / "results1003.dat", tmp_path / "results1004.dat"
    sim.save_port(str(vp), str(ip))
    lines = vp.read_text().splitlines()
    assert len(lines) == sc.steps and all(ln.endswith(" ") for ln in lines)
    assert lines[5] == "%e " % float(v[5])
    assert np.allclose(np.loadtxt(vp), v, rtol=1e-6, atol=0)
    assert np.allclose(np.loadtxt(ip), i, rtol=1e-6, atol=0)
    sim.close()


@pytest.mark.gpu
def test_port_errors():
    from fdtd_amd import api
    sc = _vacuum(n=16, steps=10)
    sim = api.Simulation.from_scene(sc, "f32", finalize=False)
    lib, h = sim.lib, sim.h
    assert lib.fdtd_set_port(h, 1, 5, 5, 5, 6, 6, 6, 3, 5, 5, 5, 6, 6, 6) == -6     # vdir 1: unsupported
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 6, 3, 0, 5, 5, 6, 6, 6) == -3     # H ring at i = -1
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 17, 3, 5, 5, 5, 6, 6, 6) == -3    # V edges past nz-1
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 16, 6, 3, 5, 5, 5, 6, 6, 6) == -3    # V box past ny-1
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 6, 3, 5, 5, 5, 6, 6, 16) == -3    # H ring past nz-1
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 6, 4, 5, 5, 5, 6, 6, 6) == -1     # idir 4
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 6, 3, 5, 5, 5, 6, 6, 6) == 0
    assert lib.fdtd_set_port(h, 3, 5, 5, 5, 6, 6, 6, 3, 5, 5, 5, 6, 6, 6) == -2     # one port
    v = np.zeros(4, np.float32)
    ptr = v.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.fdtd_get_port(h, ptr, ptr, 4) == -2                                 # not finalized
    sim.finalize(sc.steps)
    sim.run(4)
    assert lib.fdtd_get_port(h, ptr, ptr, 3) == -1                                 # too small
    assert lib.fdtd_get_port(h, ptr, ptr, 4) == 0
    out = np.zeros(8, np.float64)
    f = np.array([1e9])
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.fdtd_port_spectrum(h, -50.0, 1, dp(f), dp(out), 8) == -1
    assert lib.fdtd_port_spectrum(h, 50.0, 1, dp(f), dp(out), 7) == -1
    assert lib.fdtd_port_spectrum(h, 50.0, 1, dp(f), dp(out), 8) == 0
    assert lib.fdtd_save_port(h, os.fsencode("/nonexistent-dir/v.dat"), b"/tmp/i.dat") == -1
    sim.close()
