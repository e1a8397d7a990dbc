"""Pins for the CPU oracle (the checker of the GPU engine).

The reference ships no tests, golden vectors or fixtures for its update path,
and its CUDA sources cannot be built here, so each part of the oracle is
pinned against what the mathematics fixes, chosen so that a plausible mistake
(a dropped term, a wrong sign, index or cell size, a transposed operand)
fails at least one check:

  * update equations: exact discrete eigenmodes of a PEC cavity (three
    polarisations, non-cubic cells) must obey the Yee leapfrog's three-term
    recurrence with the discrete Laplacian eigenvalue;
  * coefficients: Maxwell's curl signs and the semi-implicit lossy update
    (Taflove & Hagness eq. 3.32-3.33) in vacuum and inside a lossy box;
  * material averaging: Yee-edge / Yee-face adjacency of a one-cell box;
  * lumped source: the resistive voltage-source update (Elsherbeni & Demir,
    eq. 4.6-4.9);
  * CPML: reduction to the plain exponential integrator (kappa = 1,
    alpha = 0); kappa alone equals a grid stretched by kappa along the
    slab's axis; the per-cell recursion converges (first order in dt) to
    1/s(w) of the complex-frequency-shifted stretch s = kappa +
    sigma/(alpha + j w eps0) (Roden & Gedney 2000); sigma, kappa and alpha
    recovered from the arrays grade in the CFS-PML's directions; grading
    monotonicity; a reflection bound against a domain large enough that
    nothing comes back;
  * pulse: the Gaussian derivative's closed-form peak and odd symmetry;
  * DFT: the accumulated transform against a direct FP64 sum with
    t_m = m dt;
  * linearity, FP32-vs-FP64 agreement;
  * adjoint: the dot-product (transpose) identity and central finite
    differences of the FP64 forward model.
"""
import math

import numpy as np
import pytest

from fdtd_amd import scenes
from fdtd_amd.scenes import Box, Cpml, Scene, Source
from oracle.oracle import Oracle

NO_CPML = Cpml(thick=(0, 0, 0, 0, 0, 0))


def small_scene(nx, ny, nz, dx=1e-3, dy=1e-3, dz=1e-3, steps=10, cpml=NO_CPML, **kw):
    dt = scenes.courant_dt(dx, dy, dz, 0.99)
    return Scene(name="t", nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, dt=dt, steps=steps, cpml=cpml, **kw)


def idx(sc, i, j, k):
    return (k * sc.ny + j) * sc.nx + i


# ---- coefficients -------------------------------------------------------------

def test_vacuum_coefficients_follow_maxwell_curl_signs():
    sc = small_scene(12, 10, 8, dx=1e-3, dy=1.3e-3, dz=0.8e-3)
    o = Oracle(sc, 64)
    x = idx(sc, 5, 4, 3)
    e0, u0, dt = sc.eps0, sc.mu0, sc.dt
    # dEx/dt = (dHz/dy - dHy/dz)/eps, dEy/dt = (dHx/dz - dHz/dx)/eps, dEz/dt = (dHy/dx - dHx/dy)/eps
    want = {"Cexe": 1.0, "Cexhz": dt / (e0 * sc.dy), "Cexhy": -dt / (e0 * sc.dz),
            "Ceye": 1.0, "Ceyhx": dt / (e0 * sc.dz), "Ceyhz": -dt / (e0 * sc.dx),
            "Ceze": 1.0, "Cezhy": dt / (e0 * sc.dx), "Cezhx": -dt / (e0 * sc.dy),
            # dHx/dt = (dEy/dz - dEz/dy)/mu, dHy/dt = (dEz/dx - dEx/dz)/mu, dHz/dt = (dEx/dy - dEy/dx)/mu
            "Chxh": 1.0, "Chxey": dt / (u0 * sc.dz), "Chxez": -dt / (u0 * sc.dy),
            "Chyh": 1.0, "Chyez": dt / (u0 * sc.dx), "Chyex": -dt / (u0 * sc.dz),
            "Chzh": 1.0, "Chzex": dt / (u0 * sc.dy), "Chzey": -dt / (u0 * sc.dx)}
    for name, v in want.items():
        assert o.array(name)[x] == pytest.approx(v, rel=1e-13), name


def test_lossy_dielectric_coefficients():
    eps_r, sig = 4.0, 0.01
    sc = small_scene(12, 12, 12, boxes=[Box((3, 3, 3, 9, 9, 9), (eps_r,) * 3, (sig,) * 3)])
    o = Oracle(sc, 64)
    x = idx(sc, 6, 6, 6)
    eps = eps_r * sc.eps0
    den = 2 * eps + sig * sc.dt
    # float32 material inputs: eps_r and sigma enter as stored float values
    eps = float(np.float32(eps_r)) * sc.eps0
    sig = float(np.float32(sig))
    den = 2 * eps + sig * sc.dt
    assert o.array("Cexe")[x] == pytest.approx((2 * eps - sig * sc.dt) / den, rel=1e-12)
    assert o.array("Cexhz")[x] == pytest.approx(2 * sc.dt / (den * sc.dy), rel=1e-12)
    assert o.array("Ceze")[x] < 1.0


def test_edge_average_of_one_cell_box():
    """A one-cell box (cell (6,5,4) spans [6,7]x[5,6]x[4,5]) touches 4 edges
    of each E component; each of those sees the arithmetic mean of the 4
    cells around it, (3 + eps_r)/4, every other edge sees vacuum."""
    eps_r = 5.0
    sc = small_scene(14, 12, 10, boxes=[Box((6, 5, 4, 7, 6, 5), (eps_r,) * 3, (0.0,) * 3)])
    o = Oracle(sc, 64)
    touch = {"x": {(6, j, k) for j in (5, 6) for k in (4, 5)},
             "y": {(i, 5, k) for i in (6, 7) for k in (4, 5)},
             "z": {(i, j, 4) for i in (6, 7) for j in (5, 6)}}
    curl = {"x": ("Cexhz", "dy"), "y": ("Ceyhx", "dz"), "z": ("Cezhy", "dx")}
    for comp, (name, d) in curl.items():
        cb = o.array(name).reshape(sc.nz, sc.ny, sc.nx)
        for k in range(2, 8):
            for j in range(2, 10):
                for i in range(2, 12):
                    eps_eff = sc.dt / (cb[k, j, i] * getattr(sc, d)) / sc.eps0
                    want = (3.0 + eps_r) / 4.0 if (i, j, k) in touch[comp] else 1.0
                    assert eps_eff == pytest.approx(want, rel=1e-12), (comp, i, j, k)


def test_face_harmonic_mean_of_one_cell_box():
    """H components sit on the faces of a cell: the face between the box cell
    and a vacuum cell sees 2 mu_r / (mu_r + 1)."""
    mu_r = 3.0
    sc = small_scene(14, 12, 10, boxes=[Box((6, 5, 4, 7, 6, 5), (1.0,) * 3, (0.0,) * 3, (mu_r,) * 3)])
    o = Oracle(sc, 64)
    faces = {"x": {(6, 5, 4), (7, 5, 4)}, "y": {(6, 5, 4), (6, 6, 4)}, "z": {(6, 5, 4), (6, 5, 5)}}
    curl = {"x": ("Chxey", "dz"), "y": ("Chyez", "dx"), "z": ("Chzex", "dy")}
    for comp, (name, d) in curl.items():
        cb = o.array(name).reshape(sc.nz, sc.ny, sc.nx)
        for k in range(2, 8):
            for j in range(2, 10):
                for i in range(2, 12):
                    mu_eff = sc.dt / (cb[k, j, i] * getattr(sc, d)) / sc.mu0
                    want = 2 * mu_r / (mu_r + 1) if (i, j, k) in faces[comp] else 1.0
                    assert mu_eff == pytest.approx(want, rel=1e-12), (comp, i, j, k)


def test_lumped_resistive_source_coefficients():
    R = 50.0
    sc = small_scene(12, 12, 12, sources=[Source((6, 6, 6, 6, 6, 7), R, scenes.WAVE_GAUSS_DERIV, 1.0,
                                                 40 * 1e-12, 12 * 1e-12)])
    o = Oracle(sc, 64)
    x = idx(sc, 6, 6, 6)
    e, dt = sc.eps0, sc.dt
    rf = dt * sc.dz / (R * sc.dx * sc.dy)          # dt dz / (R dx dy)
    D = 2 * e + rf
    assert o.array("Ceze")[x] == pytest.approx((2 * e - rf) / D, rel=1e-12)
    assert o.array("Cezhy")[x] == pytest.approx(2 * dt / sc.dx / D, rel=1e-12)
    assert o.array("Cezhx")[x] == pytest.approx(-2 * dt / sc.dy / D, rel=1e-12)
    assert o.array("Cezj")[x] == pytest.approx(-2 * dt / (R * sc.dx * sc.dy) / D, rel=1e-12)
    # the neighbouring edge is untouched
    assert o.array("Ceze")[idx(sc, 6, 6, 7)] == pytest.approx(1.0, rel=1e-13)


# ---- CPML ---------------------------------------------------------------------

def test_cpml_reduces_to_exponential_integrator():
    """kappa = 1, alpha = 0: b = exp(-sigma dt / eps0) and a d = b - 1."""
    L = 8
    sc = small_scene(24, 24, 24, cpml=Cpml(thick=(L,) * 6, sigma_max=(40.0,) * 3, kappa_max=1.0,
                                           alpha_max=0.0, alpha_min=0.0))
    o = Oracle(sc, 64)
    b, a = o.array("b_ex"), o.array("a_ex")
    inside = [i for i in range(sc.nx) if b[i] != 0.0]
    assert len(inside) == 2 * L
    for i in inside:
        assert 0.0 < b[i] < 1.0
        assert a[i] * sc.dx == pytest.approx(b[i] - 1.0, rel=1e-12)
    # exponent -sigma dt/eps0 grows towards the outer face on both sides
    s = -np.log(b[inside]) * sc.eps0 / sc.dt
    assert np.all(np.diff(s[:L]) < 0) and np.all(np.diff(s[L:]) > 0)
    assert np.all(b[L:sc.nx - L] == 0.0)


def test_cpml_profile_grading():
    L = 10
    sc = small_scene(30, 30, 30, cpml=scenes.graded_cpml(L, 1e-3))
    o = Oracle(sc, 64)
    for ax in "xyz":
        be, bm, ae = o.array("b_e" + ax), o.array("b_m" + ax), o.array("a_e" + ax)
        n = len(be)
        lo, hi = list(range(L)), list(range(n - L, n))
        assert np.all((be[lo + hi] > 0) & (be[lo + hi] < 1)) and np.all(ae[lo + hi] <= 0)
        # loss grows towards each outer face, for E and for H nodes
        assert np.all(np.diff(be[lo]) > 0) and np.all(np.diff(be[hi]) < 0)
        assert np.all(np.diff(bm[lo]) > 0) and np.all(np.diff(bm[hi][:-1]) < 0)


def _init_fields(o, sc, seed):
    """Seeded random values in all six fields (an initial state instead of a
    source, so no lumped-source loading enters the comparison)."""
    rng = np.random.default_rng(seed)
    for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        scale = 1.0 if name[0] == "E" else 1.0 / scenes.ETA0
        o.set_array(name, scale * rng.standard_normal(sc.cells))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_cpml_kappa_alone_is_a_coordinate_stretch(axis):
    """With sigma = 0 and alpha = 0 the CPML's complex stretch
    s = kappa + sigma/(alpha + j w eps0) reduces to the real constant kappa
    (Roden & Gedney 2000): a slab spanning the whole extent of one axis, with
    a flat profile (grading order 0, so kappa = kappa_max on every cell), is
    the same grid with that axis' cell size multiplied by kappa.  Both runs
    start from the same random fields in a lossy, anisotropic, magnetic box;
    every field must agree to FP64 round-off after 40 steps.  This pins the
    kappa division of all four curl coefficients of the axis, E and H side."""
    kappa = 3.5
    nx, ny, nz = 15, 13, 11
    d = [1.0e-3, 1.2e-3, 0.9e-3]
    dt = scenes.courant_dt(*d, 0.99)
    box = Box((4, 3, 3, 10, 9, 8), (3.0, 2.0, 5.0), (0.02, 0.01, 0.03), (2.0, 1.5, 3.0), (10.0, 0.0, 5.0))
    thick = [0] * 6
    thick[2 * axis] = (nx, ny, nz)[axis]
    # sigma_max <= 0 selects the reference's default, so "no loss" is 1e-300
    stretched = Cpml(thick=tuple(thick), order=0.0, sigma_max=(1e-300,) * 3, kappa_max=kappa,
                     alpha_max=0.0, alpha_min=0.0)
    dd = list(d)
    dd[axis] *= kappa
    runs = []
    for cells, cp in ((d, stretched), (dd, NO_CPML)):
        sc = Scene(name="t", nx=nx, ny=ny, nz=nz, dx=cells[0], dy=cells[1], dz=cells[2], dt=dt, steps=40,
                   cpml=cp, boxes=[box])
        o = Oracle(sc, 64)
        _init_fields(o, sc, 7)
        o.steps(40)
        runs.append({n: o.array(n) for n in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")})
        o.close()
    for n, want in runs[1].items():
        got = runs[0][n]
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max(), (axis, n)


# (curl coefficient, its cpsi, the axis' 1-D profile a/b names, sign of the
# vacuum coefficient, axis) for the 12 stretched derivatives of setABCPML
_STRETCHED = [
    ("Ceyhz", "cpsi_eyx", "ex", -1, 0), ("Cezhy", "cpsi_ezx", "ex", +1, 0),
    ("Cexhz", "cpsi_exy", "ey", +1, 1), ("Cezhx", "cpsi_ezy", "ey", -1, 1),
    ("Cexhy", "cpsi_exz", "ez", -1, 2), ("Ceyhx", "cpsi_eyz", "ez", +1, 2),
    ("Chyez", "cpsi_hyx", "mx", +1, 0), ("Chzey", "cpsi_hzx", "mx", -1, 0),
    ("Chxez", "cpsi_hxy", "my", -1, 1), ("Chzex", "cpsi_hzy", "my", +1, 1),
    ("Chxey", "cpsi_hxz", "mz", +1, 2), ("Chyex", "cpsi_hyz", "mz", -1, 2),
]


def _stretch_error(sigma, kappa, alpha, dt_factor):
    """Largest |G(w) s(w) / C0 - 1| over the 12 stretched derivatives and
    four frequencies, where G(w) = C + cpsi a / (1 - b e^{-j w dt}) is the
    per-cell transfer function of the recursion psi^n = b psi^{n-1} + a D^n
    applied as C D + cpsi psi, C0 the unstretched (vacuum) coefficient, and
    s(w) = kappa + sigma / (alpha + j w eps0) the CFS stretch (Roden &
    Gedney 2000).  The profile is flat (order 0, alpha_max = alpha_min), so
    every slab cell carries exactly (sigma, kappa, alpha)."""
    n, L = 20, 6
    d = (1.0e-3, 1.3e-3, 0.8e-3)
    dt = scenes.courant_dt(*d, 0.99) * dt_factor
    cp = Cpml(thick=(L,) * 6, order=0.0, sigma_max=(sigma,) * 3, kappa_max=kappa, alpha_max=alpha,
              alpha_min=alpha)
    sc = Scene(name="t", nx=n, ny=n, nz=n, dx=d[0], dy=d[1], dz=d[2], dt=dt, steps=1, cpml=cp)
    o = Oracle(sc, 64)
    worst = 0.0
    for cname, pname, prof, sign, axis in _STRETCHED:
        cell = [n // 2] * 3
        cell[axis] = 2                          # inside this axis' slab only
        x = idx(sc, *cell)
        mat = sc.eps0 if cname[1] == "e" else sc.mu0
        C0 = sign * dt / (mat * d[axis])
        C, cpsi = o.array(cname)[x], o.array(pname)[x]
        a, b = o.array("a_" + prof)[cell[axis]], o.array("b_" + prof)[cell[axis]]
        for f in (0.3e9, 1e9, 3e9, 10e9):
            w = 2 * math.pi * f
            G = C + cpsi * a / (1.0 - b * np.exp(-1j * w * dt))
            s = kappa + sigma / (alpha + 1j * w * sc.eps0)
            worst = max(worst, abs(G * s / C0 - 1.0))
    o.close()
    return worst


@pytest.mark.parametrize("sigma,kappa,alpha", [(8.5, 5.0, 0.05), (3.0, 2.0, 0.5), (20.0, 8.0, 0.0)])
def test_cpml_recursion_converges_to_cfs_stretch(sigma, kappa, alpha):
    """The discrete CPML (b, a, cpsi, C/kappa) realises 1/s(w) of the
    complex-frequency-shifted stretch to first order in dt: the error is
    small and halves when dt halves.  A kappa or alpha misplaced in a or b,
    or a missing kappa division, leaves an O(1) error that does not
    converge."""
    e1 = _stretch_error(sigma, kappa, alpha, 0.1)
    e2 = _stretch_error(sigma, kappa, alpha, 0.05)
    assert e1 < 3e-2, e1
    assert 0.4 < e2 / e1 < 0.6, (e1, e2)


def test_cpml_grading_directions():
    """The recovered per-cell (sigma, kappa, alpha) of the graded CPML follow
    the CFS-PML design: sigma and kappa grow from the interface towards the
    outer face (kappa from about 1 to about kappa_max), alpha falls from about
    alpha_max at the interface to about alpha_min at the outer face, on both
    sides of every axis, for E and H nodes.  Values are recovered from the
    oracle's own arrays: kappa = cpsi / (C d), sigma/kappa + alpha =
    -ln(b) eps0/dt, and sigma = a d kappa^2 (sigma/kappa + alpha) / (b - 1)
    (the relations the convergence pin above fixes)."""
    L, n = 10, 30
    d = (1.0e-3, 1.3e-3, 0.8e-3)
    amax = 0.05
    cp = scenes.graded_cpml(L, d[0])
    sc = Scene(name="t", nx=n, ny=n, nz=n, dx=d[0], dy=d[1], dz=d[2], dt=scenes.courant_dt(*d, 0.99), steps=1,
               cpml=cp)
    o = Oracle(sc, 64)
    for cname, pname, prof, sign, axis in _STRETCHED:
        mat = sc.eps0 if cname[1] == "e" else sc.mu0
        C, cpsi = o.array(cname), o.array(pname)
        a, b = o.array("a_" + prof), o.array("b_" + prof)
        for side in ("low", "high"):
            # slab cells ordered from the interface to the outer face
            order = list(range(L - 1, -1, -1)) if side == "low" else list(range(n - L, n))
            # nodes on the outer walls take no update (tangential E at c = 0 and
            # n-1, H at n-1: the ranges of H:249-573 / H:577-915), so no coefficient
            order = [c for c in order if c < n - 1 and (c > 0 or cname[1] == "h")]
            sig, kap, alp = [], [], []
            for c in order:
                cell = [n // 2] * 3
                cell[axis] = c
                x = idx(sc, *cell)
                k = cpsi[x] / (C[x] * d[axis])
                g = -math.log(b[c]) * mat / sc.dt
                s = a[c] * d[axis] * k * k * g / (b[c] - 1.0)
                scale = sc.eps0 / mat                 # H-node sigma_m, alpha_m carry mu0/eps0
                sig.append(s * scale)
                kap.append(k)
                alp.append((g - s / k) * scale)
            sig, kap, alp = map(np.array, (sig, kap, alp))
            where = (cname, side)
            assert np.all(np.diff(sig) > 0) and np.all(np.diff(kap) > 0), where
            assert np.all(np.diff(alp) < 0), where
            assert 1.0 <= kap[0] < 1.01 and kap[-1] > 0.5 * cp.kappa_max, where
            assert alp[0] > 0.9 * amax and 0.0 <= alp[-1] < 0.25 * amax, where
            assert sig[0] < 0.01 * cp.sigma_max[axis] and sig[-1] > 0.4 * cp.sigma_max[axis], where
    o.close()


def _reflection(L, cpml=True, T=110):
    """Max error (relative to the peak) of two probes 2 cells in front of
    the x- and z- CPML faces against the same geometry embedded in a domain
    too large for anything to come back within T steps (PEC outer walls)."""
    d = 1e-3
    dt = scenes.courant_dt(d, d, d, 0.99)
    n = 10 + 2 * L
    c = n // 2
    M = int(math.ceil(0.99 / math.sqrt(3) * T)) + 2   # c dt / d cells per step
    pr = [(scenes.EZ, L + 2, c, c), (scenes.EZ, c, c, L + 2)]

    def run(n_dom, sh, cp):
        src = Source((c + sh, c + sh, c + sh, c + sh, c + sh, c + sh + 1), 50.0, scenes.WAVE_GAUSS_DERIV, 1.0,
                     26 * dt, 8 * dt)
        sc = small_scene(n_dom, n_dom, n_dom, steps=T, cpml=cp, sources=[src],
                         probes=[(p[0], p[1] + sh, p[2] + sh, p[3] + sh) for p in pr])
        o = Oracle(sc, 64)
        o.steps(T)
        return o.probes()

    ref = run(n + 2 * M, M, NO_CPML)
    got = run(n, 0, scenes.graded_cpml(L, d) if cpml else NO_CPML)
    return max(float(np.abs(got[p] - ref[p]).max() / np.abs(ref[p]).max()) for p in range(len(pr)))


def test_cpml_reflection_is_small_and_falls_with_thickness():
    e5, e10 = _reflection(5), _reflection(10)
    assert e10 < 1.5e-2, e10
    assert e10 < 0.5 * e5, (e5, e10)
    # PEC walls at the same place reflect at order one
    assert _reflection(10, cpml=False) > 10 * e10


# ---- update equations: PEC cavity eigenmodes ------------------------------------

CAV = dict(nx=13, ny=11, nz=9, dx=1.0e-3, dy=1.3e-3, dz=0.8e-3)


def _cavity_mode(comp):
    """Exact discrete eigenmode of the PEC box (tangential E = 0 on the walls
    i = 0, nx-1 / j = 0, ny-1 / k = 0, nz-1) with one E component, uniform
    along that component's own axis."""
    nx, ny, nz = CAV["nx"], CAV["ny"], CAV["nz"]
    k_, j_, i_ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if comp == "Ez":
        f = np.sin(np.pi * i_ / (nx - 1)) * np.sin(np.pi * j_ / (ny - 1))
        f[k_ == nz - 1] = 0.0
        lam = 4 / CAV["dx"] ** 2 * math.sin(math.pi / (2 * (nx - 1))) ** 2 + \
            4 / CAV["dy"] ** 2 * math.sin(math.pi / (2 * (ny - 1))) ** 2
        probes = [(scenes.EZ, 4, 5, 3), (scenes.EZ, 7, 3, 0), (scenes.EZ, 9, 8, 7)]
    elif comp == "Ex":
        f = np.sin(np.pi * j_ / (ny - 1)) * np.sin(np.pi * k_ / (nz - 1))
        f[:, :, nx - 1] = 0.0
        lam = 4 / CAV["dy"] ** 2 * math.sin(math.pi / (2 * (ny - 1))) ** 2 + \
            4 / CAV["dz"] ** 2 * math.sin(math.pi / (2 * (nz - 1))) ** 2
        probes = [(scenes.EX, 3, 4, 5), (scenes.EX, 0, 6, 2), (scenes.EX, 11, 2, 6)]
    else:
        f = np.sin(np.pi * i_ / (nx - 1)) * np.sin(np.pi * k_ / (nz - 1))
        f[:, ny - 1, :] = 0.0
        lam = 4 / CAV["dx"] ** 2 * math.sin(math.pi / (2 * (nx - 1))) ** 2 + \
            4 / CAV["dz"] ** 2 * math.sin(math.pi / (2 * (nz - 1))) ** 2
        probes = [(scenes.EY, 3, 4, 5), (scenes.EY, 6, 0, 2), (scenes.EY, 10, 9, 6)]
    return f, lam, probes


@pytest.mark.parametrize("comp", ["Ez", "Ex", "Ey"])
def test_pec_cavity_mode_obeys_yee_recurrence(comp):
    f, lam, probes = _cavity_mode(comp)
    T = 80
    sc = small_scene(CAV["nx"], CAV["ny"], CAV["nz"], CAV["dx"], CAV["dy"], CAV["dz"], steps=T, probes=probes)
    o = Oracle(sc, 64)
    o.set_array(comp, f.ravel())
    o.steps(T)
    e = o.probes()
    c2dt2 = sc.dt ** 2 / (sc.eps0 * sc.mu0)
    r = e[:, 2:] + e[:, :-2] - (2.0 - c2dt2 * lam) * e[:, 1:-1]
    assert np.abs(e).max() > 0.1
    assert np.abs(r).max() < 1e-11 * np.abs(e).max()
    # the mode really oscillates: it changes sign within the window
    assert e[0].min() < 0 < e[0].max()


def test_pec_cavity_is_lossless_and_bounded():
    f, _, probes = _cavity_mode("Ez")
    sc = small_scene(CAV["nx"], CAV["ny"], CAV["nz"], CAV["dx"], CAV["dy"], CAV["dz"], steps=400, probes=probes)
    o = Oracle(sc, 64)
    o.set_array("Ez", f.ravel())
    o.steps(400)
    e = o.probes()[0]
    # a single-frequency oscillation of constant amplitude
    assert np.abs(e[200:]).max() == pytest.approx(np.abs(e[:200]).max(), rel=2e-2)


# ---- source pulse, DFT, linearity ---------------------------------------------

def test_gauss_derivative_pulse_closed_form():
    T = 120
    sc = small_scene(12, 12, 12, steps=T)
    dt = sc.dt
    src = Source((6, 6, 4, 6, 6, 7), 50.0, scenes.WAVE_GAUSS_DERIV, 2.5, 40 * dt, 12 * dt)
    sc.sources = [src]
    o = Oracle(sc, 64)
    v = o.array("vpn") * 3   # per-node share of a 3-edge source
    # V(t) = amp sqrt(2e) u exp(-u^2), u = (t - t0)/tau: peak amp at u = 1/sqrt(2)
    assert v.max() <= 2.5 * (1 + 1e-12)
    m_pk = 40 + 12 / math.sqrt(2)
    assert abs(int(np.argmax(v)) - m_pk) <= 1
    # samples lie within dt/2 of the peak: f(1/sqrt(2) + du) >= 1 - 2 du^2, du = 1/24
    assert v.max() == pytest.approx(2.5, rel=2 * (1 / 24) ** 2 + 1e-6)
    assert v[40] == 0.0
    assert np.allclose(v[40 + np.arange(1, 40)], -v[40 - np.arange(1, 40)], rtol=0, atol=1e-14)


def test_dft_is_direct_fp64_sum_with_tm_equal_m_dt():
    sc = scenes.dielectric(nx=28, ny=26, nz=24, steps=90, L=5, n_boxes=4, n_src=2, n_probes=5, n_freqs=3, seed=11)
    o = Oracle(sc, 64)
    o.steps(sc.steps)
    x = o.probes()
    X = o.dft()
    m = np.arange(sc.steps)
    for q, f in enumerate(sc.freqs):
        ph = np.exp(-2j * np.pi * f * m * sc.dt) * sc.dt
        direct = x @ ph
        assert np.allclose(X[:, q], direct, rtol=1e-12, atol=1e-12 * np.abs(direct).max())


def test_linearity_in_source_amplitude():
    base = scenes.dielectric(nx=24, ny=24, nz=20, steps=60, L=4, n_boxes=3, n_src=1, n_probes=4, n_freqs=0, seed=5)
    o1 = Oracle(base, 64)
    o1.steps(60)
    base.sources = [Source(s.ijk, s.R, s.wtype, 2.0 * s.amp, s.t0, s.tau, s.freq) for s in base.sources]
    o2 = Oracle(base, 64)
    o2.steps(60)
    a, b = o1.probes(), o2.probes()
    assert np.abs(a).max() > 0
    assert np.allclose(b, 2.0 * a, rtol=1e-12, atol=1e-15 * np.abs(a).max())


def test_fp32_oracle_tracks_fp64_within_1e5():
    sc = scenes.yee64(n=40, steps=200, L=8)
    o64, o32 = Oracle(sc, 64), Oracle(sc, 32)
    o64.steps(sc.steps)
    o32.steps(sc.steps)
    a, b = o64.probes(), o32.probes()
    scale = np.abs(a).max()
    for p in range(a.shape[0]):
        if np.linalg.norm(a[p]) < 1e-3 * scale * math.sqrt(a.shape[1]):
            continue
        assert np.linalg.norm(b[p] - a[p]) / np.linalg.norm(a[p]) < 1e-5, p


# ---- adjoint ------------------------------------------------------------------

F = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]
P = ["eyx", "ezx", "hyx", "hzx", "exy", "ezy", "hxy", "hzy", "exz", "eyz", "hxz", "hyz"]
ADJ = {"Ex": "eb_x", "Ey": "eb_y", "Ez": "eb_z", "Hx": "hb_x", "Hy": "hb_y", "Hz": "hb_z"}
ADJ.update({"Psi_" + p: "pb_" + p for p in P})


def test_adjoint_step_is_the_transpose():
    """<A x, y> = <x, A^T y> for one step of the linear update with CPML,
    anisotropic materials and loss (no sources)."""
    sc = scenes.dielectric(nx=20, ny=18, nz=16, steps=4, L=4, n_boxes=3, n_src=0, n_probes=0, n_freqs=0, seed=3)
    sc.cpml = scenes.graded_cpml((3, 4, 2, 5, 3, 3), sc.dx)
    sc.boxes.append(Box((5, 6, 4, 12, 11, 9), (2.0, 3.0, 4.0), (0.01, 0.0, 0.02), (1.5, 1.0, 2.0), (0.0, 0.0, 0.0)))
    o = Oracle(sc, 64)
    o.adjoint_alloc()
    rng = np.random.default_rng(0)
    N = sc.cells
    keys = F + ["Psi_" + p for p in P]
    x = {k: rng.standard_normal(N) for k in keys}
    for k, v in x.items():
        o.set_array(k, v)
    o.forward_step(0)
    Ax = {k: o.array(k) for k in keys}
    y = {k: rng.standard_normal(N) for k in keys}
    for k, v in y.items():
        o.set_array(ADJ[k], v)
    o.reverse_step(0)
    ATy = {k: o.array(ADJ[k]) for k in keys}
    lhs = sum(float(np.dot(Ax[k], y[k])) for k in keys)
    rhs = sum(float(np.dot(x[k], ATy[k])) for k in keys)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_adjoint_gradient_matches_finite_differences():
    kw = dict(nx=28, ny=24, nz=22, steps=160, L=4, d=1e-3, design=(6, 5, 4), eps_range=(1.0, 4.0), seed=9,
              probe_freqs=(12e9, 20e9))
    sc = scenes.adjoint(**kw)
    terms = [(0, 0, 1.0), (0, 1, 0.5)]
    J, g = Oracle(sc, 64).adjoint(terms)
    assert J > 0 and np.abs(g).max() > 0

    def objective(eps):
        s2 = scenes.adjoint(**kw)
        s2.design_eps = eps
        o = Oracle(s2, 64)
        o.steps(s2.steps)
        X = o.dft()
        return sum(w * abs(X[p, f]) ** 2 for p, f, w in terms)

    h = 1e-3
    for q in (0, 17, 61, g.size - 1):
        ep, em = sc.design_eps.copy(), sc.design_eps.copy()
        ep[q] = np.float32(ep[q] + h)
        em[q] = np.float32(em[q] - h)
        fd = (objective(ep) - objective(em)) / (float(ep[q]) - float(em[q]))
        assert g[q] == pytest.approx(fd, rel=1e-4, abs=1e-6 * np.abs(g).max()), q


# ---- storage-rounding model (the checker of the 16-bit storage forms) ---------

def test_block_scaled_model_stores_int16_blocks():
    """The oracle's block-scaled int16 model (mode 40 + 8: fields in blocks of
    8 x-cells, psi FP16): after every step each 8-cell run of every field is
    exactly int16 mantissas times one power of two, |mantissa| <= 32767, and
    the largest mantissa of a non-zero block uses the top binade (the scale
    is not wasted); its largest probe error against FP64 is at most a fifth
    of FP16's."""
    sc = scenes.dielectric(nx=36, ny=30, nz=28, steps=150, L=5, n_boxes=4, n_src=2, n_probes=6, n_freqs=2,
                           seed=31, name="model-bq")
    ref = Oracle(sc, 64)
    ref.steps(sc.steps)
    r = ref.probes()
    errs = {}
    for mode in (16, 48):
        o = Oracle(sc, 32)
        o.set_quant(mode, 1.0, scenes.ETA0)
        o.steps(sc.steps)
        p = o.probes()
        errs[mode] = np.max(np.linalg.norm(p - r, axis=1) / np.linalg.norm(r, axis=1))
        if mode == 48:
            for name in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                f = o.field(name).reshape(-1, sc.nx)[:, :sc.nx // 8 * 8].reshape(-1, 8)
                mx = np.abs(f).max(axis=1)
                live = mx > 0
                e = np.ceil(np.log2(mx[live] / 32767.0))
                q = np.exp2(e)[:, None]
                m = f[live] / q
                assert np.all(m == np.round(m)) and np.all(np.abs(m) <= 32767), name
                assert np.all(np.abs(m).max(axis=1) >= 16384), name
        o.close()
    assert errs[48] < 0.2 * errs[16], errs
