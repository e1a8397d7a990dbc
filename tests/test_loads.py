"""Conducting plates and sectors, lumped resistors and x-directed lumped
sources (definePlateX/Y/Z, defineSec, defineResistor, defineSourceX) and the
reference's own patch-antenna scene.

CPU part (oracle, FP64), pinned by geometry and physics rather than by the
reference's formulas:
  * a plate changes exactly the E edges that lie in its rectangle (an Ex
    edge spans [i, i+1] x {j} x {k}, and so on) -- for each plate axis;
  * a plate of sigma 1e10 is a conductor: the tangential E on it stays
    below 1e-6 of the field around it;
  * a sector changes exactly the Ex / Ey edges (within its index ranges) of
    the nodes inside the wedge, found by angle and radius from the apex, and
    a plate applied after it wins where they overlap;
  * a resistor edge with the port around it obeys I = -V/R minus the gap's
    displacement current, step by step (its update as a current balance);
  * an x-directed source edge obeys the same balance along x, with the
    source's R split over (1+je-js)(1+ke-ks) parallel columns of (ie-is)
    series edges; its loaded edges are is<=i<ie, js<=j<=ie, ks<=k<=ke (the
    reference bounds j by ie), and the only drive it adds is the Ez
    injection through setZero's Cezj = 1.2e-38 on the box's Ez edges.
GPU part: the engine bit for bit against the FP32 oracle on a scene with
plates, a resistor and a port, and on the patch-antenna scene.
"""
import math

import numpy as np
import pytest

from fdtd_amd import scenes
from oracle import port as oport
from oracle.oracle import Oracle

EPS0 = scenes.EPS0


def _scene(n=18, steps=60, plates=(), resistors=(), src=(9, 9, 5), L=4, probes=(), boxes=(),
           d=(1e-3, 1.2e-3, 0.8e-3)):
    dt = scenes.courant_dt(*d, 0.99)
    return scenes.Scene(name="loads", nx=n, ny=n, nz=n, dx=d[0], dy=d[1], dz=d[2], dt=dt, steps=steps,
                        cpml=scenes.graded_cpml(L, d[0]), sources=[scenes.z_edge(*src, dt)],
                        plates=list(plates), resistors=list(resistors), probes=list(probes),
                        boxes=list(boxes), seed=5)


def _on_plate(comp, i, j, k, axis, b):
    """Does the E edge (comp, i, j, k) lie in the plate's rectangle?"""
    is_, js, ks, ie, je, ke = b
    lo = [i, j, k]
    hi = [i, j, k]
    hi[comp] += 1                      # the edge spans one cell along its own axis
    a = axis - 1                       # plate normal
    if comp == a:
        return False                   # normal edges are not tangential
    if not (lo[a] == b[a] == hi[a]):
        return False
    rng = [(is_, ie), (js, je), (ks, ke)]
    return all(rng[q][0] <= lo[q] and hi[q] <= rng[q][1] for q in range(3) if q != a)


@pytest.mark.parametrize("axis,box", [(3, (5, 6, 8, 11, 10, 8)), (1, (7, 4, 5, 7, 12, 11)),
                                      (2, (4, 9, 6, 12, 9, 10))])
def test_plate_changes_exactly_the_edges_on_it(axis, box):
    sc = _scene(plates=[scenes.Plate(axis=axis, ijk=box, sigma=(5.0, 7.0, 9.0))], steps=1)
    o = Oracle(sc, precision=64)
    vac = _scene(steps=1)
    ov = Oracle(vac, precision=64)
    n = sc.nx
    for comp, name in enumerate(("Cexe", "Ceye", "Ceze")):
        a = o.array(name).reshape(n, n, n)
        v = ov.array(name).reshape(n, n, n)
        changed = {(i, j, k) for k, j, i in zip(*np.nonzero(a != v))}
        # the source edge changes Ceze on its own; leave it out
        changed.discard((9, 9, 5))
        expect = {(i, j, k) for k in range(1, n - 1) for j in range(1, n - 1) for i in range(1, n - 1)
                  if _on_plate(comp, i, j, k, axis, box)}
        assert changed == expect, (name, sorted(changed ^ expect)[:6])
        sig = (5.0, 7.0, 9.0)[comp]
        for (i, j, k) in list(expect)[:3]:   # the closed form with the plate's sigma (vacuum eps)
            ce = (2 * EPS0 - sc.dt * sig) / (2 * EPS0 + sc.dt * sig)
            assert abs(a[k, j, i] - ce) < 1e-12


def test_pec_plate_holds_tangential_e_at_zero():
    box = (4, 4, 10, 14, 14, 10)
    probes = [(c, i, j, 10) for c in (0, 1) for i in (7, 9, 11) for j in (7, 9, 11)]
    near = [(c, i, j, 8) for c in (0, 1) for i in (7, 9, 11) for j in (7, 9, 11)]
    sc = _scene(plates=[scenes.Plate(axis=3, ijk=box, sigma=(1e10, 1e10, 1e10))], steps=120,
                probes=probes + near)
    o = Oracle(sc, precision=64)
    o.steps(sc.steps)
    h = o.probes()
    on, off = np.abs(h[:len(probes)]).max(), np.abs(h[len(probes):]).max()
    assert off > 0
    assert on < 1e-6 * off


def test_resistor_balance_with_the_port():
    R = 75.0
    res = scenes.Resistor(ijk=(11, 9, 5, 11, 9, 6), R=R)
    sc = _scene(resistors=[res], steps=100)
    cells = oport.cells(res.ijk, res.ijk, 3, sc.nz)
    sc.probes = cells
    o = Oracle(sc, precision=64)
    o.steps(sc.steps)
    hist = o.probes()
    look = {c: hist[q] for q, c in enumerate(cells)}
    get = lambda comp, i, j, k: look[(comp, i, j, k)]
    v = oport.voltage(get, res.ijk, sc.dz, sc.nz, dtype=np.float64)
    raw = oport.current_raw(get, res.ijk, 3, sc.dx, sc.dy, sc.dz, sc.nz, dtype=np.float64)
    c_gap = EPS0 * sc.dx * sc.dy / sc.dz
    pred = -c_gap / sc.dt * np.diff(v) - (v[1:] + v[:-1]) / (2 * R)
    scale = np.max(np.abs(raw))
    assert scale > 0
    assert np.max(np.abs(raw[:-1] - pred)) < 1e-9 * scale


def _xsource(ijk, R=60.0, amp=1.0, dt=None):
    return scenes.Source(ijk=ijk, R=R, dir=1, **scenes.gauss_deriv(dt, amp))


def test_xsource_loads_exactly_its_edges():
    box = (6, 7, 8, 10, 9, 9)                      # je = 9 < ie = 10
    sc = _scene(steps=1)
    sc.sources.append(_xsource(box, dt=sc.dt))
    o = Oracle(sc, precision=64)
    ov = Oracle(_scene(steps=1), precision=64)
    n = sc.nx
    changed = {(i, j, k) for k, j, i in zip(*np.nonzero(o.array("Cexe").reshape(n, n, n)
                                                        != ov.array("Cexe").reshape(n, n, n)))}
    expect = {(i, j, k) for i in range(6, 10) for j in range(7, 11) for k in (8, 9)}
    assert changed == expect, sorted(changed ^ expect)[:6]
    for name in ("Ceze", "Cezhy", "Cezhx", "Ceye"):
        assert np.array_equal(o.array(name), ov.array(name)), name
    cezj = o.array("Cezj").reshape(n, n, n) - ov.array("Cezj").reshape(n, n, n)
    inj = {(i, j, k) for k, j, i in zip(*np.nonzero(cezj))}
    assert inj == {(i, 7 + dj, 8) for i in range(6, 11) for dj in range(3)}


def test_xsource_edge_balance():
    R = 60.0
    box = (8, 9, 9, 10, 10, 10)                    # 2 x 2 parallel columns of 2 series edges
    sc = _scene(steps=100)
    sc.sources.append(_xsource(box, R=R, dt=sc.dt))
    i, j, k = 8, 9, 9
    sc.probes = [(0, i, j, k), (5, i, j, k), (5, i, j - 1, k), (4, i, j, k), (4, i, j, k - 1)]
    o = Oracle(sc, precision=64)
    o.steps(sc.steps)
    ex, hz, hzm, hy, hym = o.probes()
    r_comp = R * (1 + box[4] - box[1]) * (1 + box[5] - box[2]) / (box[3] - box[0])
    circ = (hz - hzm)[:-1] * sc.dz - (hy - hym)[:-1] * sc.dy
    pred = EPS0 * sc.dy * sc.dz / sc.dt * np.diff(ex) + sc.dx * (ex[1:] + ex[:-1]) / (2 * r_comp)
    scale = np.max(np.abs(circ))
    assert scale > 0
    assert np.max(np.abs(circ - pred)) < 1e-9 * scale


def test_xsource_alone_drives_only_through_tiny_cezj():
    box = (7, 7, 8, 10, 8, 10)
    sc = _scene(steps=1)
    sc.sources = [_xsource(box, dt=sc.dt, amp=1e3)]
    o = Oracle(sc, precision=64)
    o.steps(1)
    n = sc.nx
    ez = o.array("Ez").reshape(n, n, n)
    nz = {(i, j, k) for k, j, i in zip(*np.nonzero(ez))}
    assert nz == {(i, j, k) for i in range(7, 11) for j in (7, 8) for k in (8, 9)}
    vals = ez[ez != 0]
    assert np.all(vals == vals[0]) and 0 < abs(vals[0]) < 1e-30
    assert not np.any(o.array("Ex")) and not np.any(o.array("Ey"))


SECTOR = dict(points=(31, 21, 20, 18, 23, 31), kz=9, radius=10.3e-3, max_x=28, max_y=27)


def _sector_nodes(sc, points, radius, max_i, max_j):
    """Nodes (i, j) of the wedge by angle: apex (i1, j1), between the rays
    towards (i0, j0) and (i2, j2) (the arc under pi), within radius metres.
    Independent of the side tests; asserts no node lies near an edge."""
    i0, j0, i1, j1, i2, j2 = points
    a0 = np.arctan2((j0 - j1) * sc.dy, (i0 - i1) * sc.dx)
    a2 = np.arctan2((j2 - j1) * sc.dy, (i2 - i1) * sc.dx)
    span = (a2 - a0) % (2 * np.pi)
    assert 0 < span < np.pi
    out = set()
    for j in range(max_j + 1):
        for i in range(max_i + 1):
            wx, wy = (i - i1) * sc.dx, (j - j1) * sc.dy
            r = np.hypot(wx, wy)
            if r == 0:
                out.add((i, j))
                continue
            t = (np.arctan2(wy, wx) - a0) % (2 * np.pi)
            assert abs(r - radius) > 1e-6 * radius
            assert min(abs(t), abs(t - span), abs(t - 2 * np.pi)) > 1e-6 or r > radius
            if t <= span and r <= radius:
                out.add((i, j))
    return out


def test_sector_loads_exactly_the_wedge_edges():
    sc = _scene(n=40, steps=1, src=(8, 8, 5))
    sc.sectors = [scenes.Sector(sigma=(5.0, 7.0), **SECTOR)]
    # a plate over part of the sector, applied after it (later wins)
    sc.plates = [scenes.Plate(axis=3, ijk=(18, 16, 9, 22, 20, 9), sigma=(50.0, 70.0, 0.0))]
    o = Oracle(sc, precision=64)
    vac = _scene(n=40, steps=1, src=(8, 8, 5))
    ov = Oracle(vac, precision=64)
    n = sc.nx
    inside = _sector_nodes(sc, SECTOR["points"], SECTOR["radius"], SECTOR["max_x"], SECTOR["max_y"])
    assert len(inside) > 40
    kz = SECTOR["kz"]
    for comp, name, sig, rng in ((0, "Cexe", 5.0, lambda i, j: i < SECTOR["max_x"]),
                                 (1, "Ceye", 7.0, lambda i, j: j < SECTOR["max_y"])):
        a = o.array(name).reshape(n, n, n)
        v = ov.array(name).reshape(n, n, n)
        changed = {(i, j, k) for k, j, i in zip(*np.nonzero(a != v))}
        on_plate = {(i, j, kz) for j in range(16, 21) for i in range(18, 23)
                    if _on_plate(comp, i, j, kz, 3, (18, 16, 9, 22, 20, 9))}
        expect = {(i, j, kz) for (i, j) in inside if rng(i, j)} | on_plate
        assert changed == expect, (name, sorted(changed ^ expect)[:6])
        for (i, j, k) in changed:
            s_ = (50.0, 70.0)[comp] if (i, j, k) in on_plate else sig
            ce = (2 * EPS0 - sc.dt * s_) / (2 * EPS0 + sc.dt * s_)
            assert abs(a[k, j, i] - ce) < 1e-12, (name, i, j, k)


def test_calc_index_and_sector_points():
    # calc_index rounds halves away from zero, as C's round() does
    assert [scenes.calc_index(v, 0.5) for v in (1.25, -1.25, 0.75, 0.2)] == [3, -3, 2, 0]
    z, d = (0, 0), (1e-3, 1e-3)
    # a 60-degree sector of radius 10 cells opening towards -x from the apex
    assert scenes.sector_points(10e-3, math.pi / 3, 0.0, (0.0, 0.0), 0.0, 0.0, 0.0, 0, z, z, d) == \
        (-9, -5, 0, 0, -9, 5)
    # rotate_point turns clockwise: (x, y) -> (y, -x) at 90 degrees
    assert scenes.sector_points(10e-3, math.pi / 3, math.pi / 2, (0.0, 0.0), 0.0, 0.0, 0.0, 0, z, z, d) == \
        (-5, 9, 0, 0, 5, 9)
    # gap 2 cells, apex 1 cell out: every point slides by the 1-cell shortfall,
    # away from the origin for sectors 0-2 and towards it for 3-5; then the
    # translation and the air + CPML offsets
    assert scenes.sector_points(10e-3, math.pi / 3, 0.0, (-1e-3, 0.0), 2e-3, 0.0, 0.0, 1, z, z, d) == \
        (-12, -5, -2, 0, -12, 5)
    assert scenes.sector_points(10e-3, math.pi / 3, 0.0, (-1e-3, 0.0), 2e-3, 0.0, 0.0, 4, z, z, d) == \
        (-10, -5, 0, 0, -10, 5)
    assert scenes.sector_points(10e-3, math.pi / 3, 0.0, (0.0, 0.0), 0.0, 50e-3, 30e-3, 0, (10, 10), (15, 15),
                                d) == (-9 + 75, -5 + 55, 75, 55, -9 + 75, 5 + 55)


def test_patch_antenna_scene_matches_source_cu():
    sc = scenes.patch_antenna(steps=10)
    # 60 x 100 x 3 cells of object, 15 air + 10 CPML on every side
    assert (sc.nx, sc.ny, sc.nz) == (110, 150, 53)
    assert [p.ijk for p in sc.plates] == [(25, 25, 25, 85, 125, 25), (44, 25, 28, 50, 75, 28),
                                          (39, 75, 28, 71, 115, 28)]
    assert sc.sources[0].ijk == (44, 25, 25, 50, 25, 28)
    assert sc.port.vbox == (44, 35, 25, 50, 35, 28) and sc.port.ibox == (44, 35, 28, 50, 35, 28)
    assert sc.boxes[0].ijk == (25, 25, 25, 85, 125, 28)


# ---- GPU -------------------------------------------------------------------------

def _engine_vs_oracle(sc):
    from fdtd_amd import api
    sim = api.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    v, i = sim.port()
    pr = sim.probes()
    sim.close()
    p = sc.port
    cells = oport.cells(p.vbox, p.ibox, p.idir, sc.nz)
    sc2 = scenes.Scene(**{**sc.__dict__, "probes": list(sc.probes) + cells, "port": None})
    o = Oracle(sc2, precision=32)
    o.steps(sc.steps)
    h = o.probes()
    look = {c: h[len(sc.probes) + q] for q, c in enumerate(cells)}
    get = lambda comp, i_, j, k: look[(comp, i_, j, k)]
    vo = oport.voltage(get, p.vbox, sc.dz, sc.nz)
    io = oport.correct(oport.current_raw(get, p.ibox, p.idir, sc.dx, sc.dy, sc.dz, sc.nz))
    return v, i, pr[:len(sc.probes)], vo, io, h[:len(sc.probes)].astype(np.float32)


@pytest.mark.gpu
def test_plates_resistor_port_bitwise_vs_oracle():
    sc = _scene(n=22, steps=150, src=(10, 11, 8),
                plates=[scenes.Plate(axis=3, ijk=(5, 5, 12, 16, 16, 12), sigma=(1e10, 1e10, 1e10)),
                        scenes.Plate(axis=1, ijk=(15, 6, 6, 15, 14, 11), sigma=(3.0, 2e3, 1e10)),
                        scenes.Plate(axis=2, ijk=(6, 15, 5, 13, 15, 10), sigma=(50.0, 1.0, 0.5))],
                resistors=[scenes.Resistor(ijk=(13, 11, 7, 13, 12, 9), R=100.0)],
                boxes=[scenes.Box(ijk=(6, 6, 4, 15, 16, 8), eps_r=(3.0, 3.0, 3.0), sigma_e=(0.01, 0.01, 0.01))],
                probes=[(2, 12, 12, 10), (0, 8, 9, 13), (1, 14, 10, 9)])
    sc.port = scenes.Port(vbox=(10, 11, 8, 10, 11, 9), ibox=(10, 11, 8, 10, 11, 9), idir=3)
    print("loads case: 22^3, 3 plates, 1 resistor, port", flush=True)
    v, i, pr, vo, io, po = _engine_vs_oracle(sc)
    assert np.max(np.abs(vo)) > 0
    assert np.array_equal(pr, po)
    assert np.array_equal(v, vo)
    assert np.array_equal(i, io)


@pytest.mark.gpu
def test_patch_antenna_bitwise_vs_oracle():
    sc = scenes.patch_antenna(steps=300)
    print("patch antenna: 110x150x53, 300 steps", flush=True)
    v, i, _, vo, io, _ = _engine_vs_oracle(sc)
    assert np.max(np.abs(vo)) > 0 and np.max(np.abs(io)) > 0
    assert np.array_equal(v, vo)
    assert np.array_equal(i, io)


@pytest.mark.gpu
def test_xsource_bitwise_vs_oracle():
    """An x-directed source beside a z-source (its loaded edges then carry
    the field), and the x-source alone (fields of order 1e-38, denormal in
    FP32)."""
    from fdtd_amd import api
    base = _scene(n=22, steps=120, src=(8, 11, 8),
                  probes=[(0, 12, 10, 11), (0, 13, 12, 12), (2, 13, 11, 11), (4, 14, 12, 11)])
    base.sources.append(_xsource((12, 10, 11, 15, 11, 12), R=40.0, dt=base.dt))
    alone = _scene(n=22, steps=60, probes=base.probes)
    alone.sources = [_xsource((12, 10, 11, 15, 11, 12), R=40.0, dt=alone.dt, amp=1e3)]
    for name, sc in (("with z-source", base), ("alone", alone)):
        print("x-source case:", name, flush=True)
        sim = api.Simulation.from_scene(sc, "f32")
        sim.run(sc.steps)
        pr = sim.probes()
        sim.close()
        o = Oracle(sc, precision=32)
        o.steps(sc.steps)
        po = o.probes().astype(np.float32)
        assert np.max(np.abs(po)) > 0, name
        assert np.array_equal(pr, po), (name, np.max(np.abs(pr - po)))


@pytest.mark.gpu
def test_xsource_errors():
    from fdtd_amd import api
    sc = _scene(n=18, steps=10)
    sim = api.Simulation.from_scene(sc, "f32", finalize=False)
    w = scenes.gauss_deriv(sc.dt)
    add = lambda ijk, d=1: sim.add_lumped_source(ijk, 50.0, w["wtype"], w["amp"], w["t0"], w["tau"], direction=d)
    for ijk, d, code in (((6, 6, 6, 9, 6, 6), 2, -6),      # y-directed: not in the reference
                         ((6, 6, 6, 6, 6, 6), 1, -1),      # ie == is: no x edge
                         ((6, 9, 6, 8, 9, 6), 1, -1),      # js > ie: no row
                         ((0, 3, 6, 4, 6, 6), 1, -3),      # i = 0 Ez edge is not updated
                         ((6, 6, 6, 9, 6, 17), 1, -3)):    # past the updated Ex planes
        with pytest.raises(api.FdtdError) as ei:
            add(ijk, d)
        assert ei.value.code == code, (ijk, d)
    add((1, 4, 6, 5, 6, 6))                                # accepted, but inside the x CPML slab
    with pytest.raises(api.FdtdError) as ei:
        sim.finalize(sc.steps)
    assert "CPML" in str(ei.value)
    sim.close()


@pytest.mark.gpu
def test_sectors_bitwise_vs_oracle():
    """Two overlapping sectors (the later wins) and a plate over part of
    them, against the FP32 oracle bit for bit."""
    from fdtd_amd import api
    sc = _scene(n=40, steps=150, src=(14, 12, 9),
                boxes=[scenes.Box(ijk=(10, 10, 3, 30, 30, 9), eps_r=(2.2, 2.2, 2.2), sigma_e=(4e-4, 4e-4, 4e-4))],
                probes=[(0, 22, 22, 9), (1, 24, 21, 9), (2, 22, 22, 8), (0, 26, 25, 10), (4, 22, 22, 9),
                        (1, 19, 19, 9)])
    sc.sectors = [scenes.Sector(sigma=(1e10, 1e10), **SECTOR),
                  scenes.Sector(points=(12, 30, 22, 22, 30, 30), kz=9, radius=7.5e-3, max_x=30, max_y=30,
                                sigma=(3.0, 40.0))]
    sc.plates = [scenes.Plate(axis=3, ijk=(18, 16, 9, 22, 20, 9), sigma=(50.0, 70.0, 0.0))]
    print("sector case: 40^3, 2 sectors, 1 plate", flush=True)
    sim = api.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    pr = sim.probes()
    sim.close()
    o = Oracle(sc, precision=32)
    o.steps(sc.steps)
    po = o.probes().astype(np.float32)
    assert np.max(np.abs(po)) > 0
    assert np.array_equal(pr, po), np.max(np.abs(pr - po))


@pytest.mark.gpu
def test_sector_errors():
    from fdtd_amd import api
    sc = _scene(n=18, steps=10)
    sim = api.Simulation.from_scene(sc, "f32", finalize=False)
    pts = (12, 9, 9, 9, 9, 12)
    for args, code in (((pts, 0, 5e-3, 10, 10), -3),      # plane k = 0 has no updated Ex / Ey
                       ((pts, 17, 5e-3, 10, 10), -3),     # plane k = nz-1 neither
                       ((pts, 8, 5e-3, 18, 10), -3),      # max_x past nx-1
                       ((pts, 8, 0.0, 10, 10), -1),       # radius must be > 0
                       ((pts, 8, float("nan"), 10, 10), -1)):
        with pytest.raises(api.FdtdError) as ei:
            sim.add_sector(*args, 1e10, 1e10)
        assert ei.value.code == code, args
    with pytest.raises(api.FdtdError) as ei:
        sim.add_sector(pts, 8, 5e-3, 10, 10, -1.0, 1.0)
    assert ei.value.code == -1
    sim.add_sector(pts, 8, 5e-3, 10, 10, 1e10, 1e10)
    sim.finalize(sc.steps)
    sim.close()
