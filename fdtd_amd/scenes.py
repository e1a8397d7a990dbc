"""Seeded synthetic scenes: grid, CPML, dielectric boxes, lumped sources,
probes and DFT frequencies.

One portable generator (splitmix64) feeds the engine, the CPU oracle and the
benchmark, so every consumer sees the same inputs for a given seed.  The five
named workloads are:

  yee64          64^3, 1 mm, CPML 10, vacuum, 1 source at the centre, 8 Ez
                 probes, 500 steps
  dielectric256  256^3, 0.5 mm, CPML 12, 64 random boxes, 4 sources, 64
                 probes, DFT at 16 frequencies 1-30 GHz, 2000 steps
  slab512        512x512x256, 0.5 mm, CPML 16, 256 boxes, 8 sources, 128
                 probes, the same DFT set, 3000 steps (one GPU's share of
                 a z-slab run)
  adjoint128     128x128x64, 1 mm, CPML 10, design region 32x32x16 with
                 eps_r ~ U[1,4), 1 source, 1 probe 20 cells past the region,
                 DFT at 10 GHz, 1000 steps
  adjoint768     768x768x384, 0.5 mm, CPML 16, design 128x128x64 with eps_r ~
                 U[1,6), 4 sources, 8 probes x {5,10,15,20} GHz, 4000 steps

Every builder also takes smaller sizes so tests can run the same recipe on
grids the CPU oracle finishes in seconds.

Conventions (the reference's, fdtdrav.h): cell indices include the CPML;
boxes are half-open [is,ie) x [js,je) x [ks,ke) and later boxes win
(setBrick, fdtdrav.h:1524-1571); a z-source box covers the Ez edges
is<=i<=ie, js<=j<=je, ks<=k<ke (defineSourceZ, fdtdrav.h:1402).
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

EPS0 = 8.854187817e-12
MU0 = 4.0 * math.pi * 1e-7
C0 = 1.0 / math.sqrt(MU0 * EPS0)
ETA0 = math.sqrt(MU0 / EPS0)

WAVE_GAUSSIAN, WAVE_GAUSS_DERIV, WAVE_SINE = 1, 2, 3
EX, EY, EZ, HX, HY, HZ = range(6)

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): portable seeded 64-bit stream."""

    def __init__(self, seed: int):
        self.s = seed & _MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        """Uniform in [lo, hi) with 53 random bits."""
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        return lo + self.next() % (hi - lo + 1)


@dataclass
class Cpml:
    thick: Tuple[int, int, int, int, int, int]
    order: float = 3.0
    sigma_max: Tuple[float, float, float] = (0.0, 0.0, 0.0)   # <= 0: reference default
    sigma_factor: float = 1.5
    kappa_max: float = 5.0
    alpha_max: float = 0.05
    alpha_min: float = 0.0


@dataclass
class Box:
    ijk: Tuple[int, int, int, int, int, int]
    eps_r: Tuple[float, float, float]
    sigma_e: Tuple[float, float, float]
    mu_r: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    sigma_m: Tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class Source:
    ijk: Tuple[int, int, int, int, int, int]
    R: float
    wtype: int
    amp: float
    t0: float
    tau: float
    freq: float = 0.0
    dir: int = 3      # 3: z-directed (defineSourceZ); 1: x-directed (defineSourceX: loaded Ex edges)


def _c_round(v: float) -> int:
    """C's round(): halves away from zero (Python's round() goes to even)."""
    a = abs(v)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1
    return int(r if v >= 0 else -r)


def calc_index(coord, d) -> int:
    """calc_index (fdtdrav.h:198-206): round(coord / d) of float32 values."""
    return _c_round(float(np.float32(np.float32(coord) / np.float32(d))))


_LIBM = None


def _libm_f(name):
    """glibc's float function `name` (cosf, sinf, tanf): the host build of
    the reference calls these."""
    global _LIBM
    import ctypes
    import ctypes.util
    if _LIBM is None:
        _LIBM = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
    f = getattr(_LIBM, name)
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float]
    return lambda x: np.float32(f(float(x)))


def sector_points(raio, sec_angle, rot_angle, p1, gap, trans_x, trans_y, i, pml_n, air_n, d):
    """sectorPoints (fdtdrav.h:2550-2618), float32 throughout: the ray ends
    at angle +-sec_angle/2 and distance raio from the apex, offset by gap,
    rotated by rot_angle (rotate_point, fdtdrav.h:218-232: x' = x cos + y
    sin, y' = -x sin + y cos), slid by desloc = gap - |p1'.x| along x (to the
    left for sectors i < 3) with the matching y shift, translated by
    (trans_x, trans_y) and indexed with calc_index plus the air and CPML
    offsets.  p1 is the apex (ponto1); pml_n, air_n and d are (x, y) pairs.
    Returns (i0, j0, i1, j1, i2, j2) for Sector.points."""
    f32 = np.float32
    cosf, sinf, tanf = _libm_f("cosf"), _libm_f("sinf"), _libm_f("tanf")
    raio, gap, rot = f32(raio), f32(gap), f32(rot_angle)
    half = f32(f32(sec_angle) / f32(2))

    def rotate(x, y):
        c, s_ = cosf(rot), sinf(rot)
        return f32(f32(x * c) + f32(y * s_)), f32(f32(f32(-x) * s_) + f32(y * c))

    p2 = (f32(-f32(f32(raio * cosf(half)) + gap)), f32(raio * sinf(half)))
    p0 = (f32(-f32(f32(raio * cosf(half)) + gap)), f32(-f32(raio * sinf(half))))
    p0 = rotate(*p0)
    q1 = rotate(f32(p1[0]), f32(p1[1]))
    p2 = rotate(*p2)
    desloc = f32(gap - abs(q1[0]))
    sh = f32(desloc * tanf(rot))
    sgn = 1 if i < 3 else -1
    pts = []
    for (x, y) in (p0, q1, p2):
        x = f32(x - desloc) if sgn > 0 else f32(x + desloc)
        y = f32(y + sh) if sgn > 0 else f32(y - sh)
        x = f32(x - f32(-f32(trans_x)))        # trans_point(x, -trans_x)
        y = f32(y - f32(-f32(trans_y)))
        pts += [air_n[0] + pml_n[0] + calc_index(x, d[0]), air_n[1] + pml_n[1] + calc_index(y, d[1])]
    return tuple(pts)


@dataclass
class Sector:
    """Zero-thickness conducting sector in the plane k = kz (defineSec,
    fdtdrav.h:1681-1744): apex (i1, j1), rays towards (i0, j0) and (i2, j2),
    radius in metres; Ex edges with i < max_x, j <= max_y and Ey edges with
    i <= max_x, j < max_y of the nodes inside get sigma = (sigma_x, sigma_y)."""
    points: Tuple[int, int, int, int, int, int]
    kz: int
    radius: float
    max_x: int
    max_y: int
    sigma: Tuple[float, float]


@dataclass
class Plate:
    """Zero-thickness conducting plate (definePlateX/Y/Z, fdtdrav.h:1574-1678):
    axis 1/2/3 and the inclusive index box of its calc_index'ed extents."""
    axis: int
    ijk: Tuple[int, int, int, int, int, int]
    sigma: Tuple[float, float, float]


@dataclass
class Resistor:
    """Lumped resistor on the Ez edges is<=i<=ie, js<=j<=je, ks<=k<ke
    (defineResistor, fdtdrav.h:1429-1484)."""
    ijk: Tuple[int, int, int, int, int, int]
    R: float


@dataclass
class Port:
    """Sampled-voltage box (z) and sampled-current box of direction idir,
    inclusive indices (fdtdrav.h:3614-3650)."""
    vbox: Tuple[int, int, int, int, int, int]
    ibox: Tuple[int, int, int, int, int, int]
    idir: int = 3


@dataclass
class Scene:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    steps: int
    cpml: Cpml
    boxes: List[Box] = field(default_factory=list)
    sources: List[Source] = field(default_factory=list)
    probes: List[Tuple[int, int, int, int]] = field(default_factory=list)   # (comp, i, j, k)
    freqs: List[float] = field(default_factory=list)
    sectors: List[Sector] = field(default_factory=list)
    plates: List[Plate] = field(default_factory=list)
    resistors: List[Resistor] = field(default_factory=list)
    port: Optional[Port] = None
    design: Optional[Tuple[int, int, int, int, int, int]] = None
    design_eps: Optional[np.ndarray] = None      # float32, x-fastest over the design region
    eps0: float = EPS0
    mu0: float = MU0
    seed: int = 0

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def describe(self) -> dict:
        return {"workload": self.name, "grid": [self.nx, self.ny, self.nz],
                "cell_m": [self.dx, self.dy, self.dz], "dt_s": self.dt, "steps": self.steps,
                "cpml": list(self.cpml.thick), "boxes": len(self.boxes),
                "sources": len(self.sources), "probes": len(self.probes),
                "plates": len(self.plates), "resistors": len(self.resistors),
                "dft_freqs": len(self.freqs), "seed": self.seed}


def courant_dt(dx, dy, dz, factor, eps0=EPS0, mu0=MU0) -> float:
    """dt = factor / (c sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)) (source.cu:45-46)."""
    c = 1.0 / math.sqrt(mu0 * eps0)
    return factor / (c * math.sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)))


def graded_cpml(L, d, order=3.0) -> Cpml:
    """CPML of the workloads: sigma_max = 0.8 (m+1) / (eta0 d), kappa_max 5,
    alpha 0.05 -> 0, polynomial grading m = 3."""
    smax = 0.8 * (order + 1.0) / (ETA0 * d)
    thick = L if isinstance(L, tuple) else (L,) * 6
    return Cpml(thick=thick, order=order, sigma_max=(smax, smax, smax), kappa_max=5.0,
                alpha_max=0.05, alpha_min=0.0)


def gauss_deriv(dt, amp=1.0) -> dict:
    """The workloads' pulse: Gaussian derivative, t0 = 40 dt, tau = 12 dt."""
    return dict(wtype=WAVE_GAUSS_DERIV, amp=amp, t0=40.0 * dt, tau=12.0 * dt)


def z_edge(i, j, k, dt, R=50.0, amp=1.0) -> Source:
    return Source(ijk=(i, j, k, i, j, k + 1), R=R, **gauss_deriv(dt, amp))


def _interior(rng, n, L):
    return rng.randint(L + 2, n - L - 3)


def _random_boxes(rng, n_boxes, nx, ny, nz, L, side=(8, 48), eps=(1.0, 12.0), sig=(0.0, 0.01)):
    boxes = []
    dims = (nx, ny, nz)
    for _ in range(n_boxes):
        lo, hi = [], []
        for n in dims:
            smax = min(side[1], n - 2 * L)
            smin = min(side[0], smax)
            s = rng.randint(smin, smax)
            a = rng.randint(L, n - L - s)
            lo.append(a)
            hi.append(a + s)
        e = rng.uniform(*eps)
        sg = rng.uniform(*sig)
        boxes.append(Box(ijk=(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]), eps_r=(e, e, e),
                         sigma_e=(sg, sg, sg)))
    return boxes


def _random_sources_probes(rng, n_src, n_probe, nx, ny, nz, L, dt, comp=EZ):
    srcs, edges = [], set()
    while len(srcs) < n_src:
        i, j, k = _interior(rng, nx, L), _interior(rng, ny, L), _interior(rng, nz, L)
        if (i, j, k) in edges:
            continue
        edges.add((i, j, k))
        srcs.append(z_edge(i, j, k, dt))
    probes, seen = [], set()
    while len(probes) < n_probe:
        i, j, k = _interior(rng, nx, L), _interior(rng, ny, L), _interior(rng, nz, L)
        if (i, j, k) in seen or (i, j, k) in edges:
            continue
        seen.add((i, j, k))
        probes.append((comp, i, j, k))
    return srcs, probes


def dft_set_1_30ghz(n=16):
    return [1e9 + 29e9 * q / (n - 1) for q in range(n)]


def yee64(n=64, steps=500, L=10, d=1e-3, n_probes=8, seed=1001) -> Scene:
    """Vacuum cube, one centred z-source, seeded Ez probes."""
    dt = courant_dt(d, d, d, 0.99)
    rng = SplitMix64(seed)
    c = n // 2
    src = [z_edge(c, c, c, dt)]
    _, probes = _random_sources_probes(rng, 0, n_probes, n, n, n, L, dt)
    probes = [p for p in probes]
    return Scene(name=f"yee{n}", nx=n, ny=n, nz=n, dx=d, dy=d, dz=d, dt=dt, steps=steps,
                 cpml=graded_cpml(L, d), sources=src, probes=probes, seed=seed)


def dielectric(nx=256, ny=256, nz=256, steps=2000, L=12, d=0.5e-3, n_boxes=64, n_src=4,
               n_probes=64, n_freqs=16, seed=1002, name=None) -> Scene:
    """Random dielectric boxes (last written wins), z-sources, Ez probes, DFT."""
    dt = courant_dt(d, d, d, 0.99)
    rng = SplitMix64(seed)
    boxes = _random_boxes(rng, n_boxes, nx, ny, nz, L)
    srcs, probes = _random_sources_probes(rng, n_src, n_probes, nx, ny, nz, L, dt)
    return Scene(name=name or f"dielectric{nx}x{ny}x{nz}", nx=nx, ny=ny, nz=nz, dx=d, dy=d, dz=d,
                 dt=dt, steps=steps, cpml=graded_cpml(L, d), boxes=boxes, sources=srcs,
                 probes=probes, freqs=dft_set_1_30ghz(n_freqs), seed=seed)


def dielectric256(**kw) -> Scene:
    return dielectric(**{**dict(nx=256, ny=256, nz=256, steps=2000, L=12, n_boxes=64, n_src=4,
                                n_probes=64, seed=1002, name="dielectric256"), **kw})


def slab512(**kw) -> Scene:
    return dielectric(**{**dict(nx=512, ny=512, nz=256, steps=3000, L=16, n_boxes=256, n_src=8,
                                n_probes=128, seed=1003, name="slab512"), **kw})


def adjoint(nx=128, ny=128, nz=64, steps=1000, L=10, d=1e-3, design=(32, 32, 16),
            eps_range=(1.0, 4.0), n_src=1, probe_freqs=(10e9,), n_probes=1, seed=1004,
            name=None) -> Scene:
    """Centred design region with seeded per-cell eps_r; sources before and
    probes 20 cells past the region along x."""
    dt = courant_dt(d, d, d, 0.99)
    rng = SplitMix64(seed)
    dnx, dny, dnz = design
    i0, j0, k0 = (nx - dnx) // 2, (ny - dny) // 2, (nz - dnz) // 2
    reg = (i0, j0, k0, i0 + dnx, j0 + dny, k0 + dnz)
    eps = np.array([rng.uniform(*eps_range) for _ in range(dnx * dny * dnz)], dtype=np.float32)
    cj, ck = ny // 2, nz // 2
    srcs = []
    for q in range(n_src):
        off = (q - (n_src - 1) / 2.0) * 4
        srcs.append(z_edge(max(L + 2, i0 - 20), int(cj + off), ck, dt))
    probes = []
    for q in range(n_probes):
        off = (q - (n_probes - 1) / 2.0) * 3
        probes.append((EZ, min(nx - L - 3, i0 + dnx - 1 + 20), int(cj + off), ck))
    return Scene(name=name or f"adjoint{nx}x{ny}x{nz}", nx=nx, ny=ny, nz=nz, dx=d, dy=d, dz=d,
                 dt=dt, steps=steps, cpml=graded_cpml(L, d), sources=srcs, probes=probes,
                 freqs=list(probe_freqs), design=reg, design_eps=eps, seed=seed)


def adjoint128(**kw) -> Scene:
    return adjoint(**{**dict(name="adjoint128"), **kw})


def adjoint768(**kw) -> Scene:
    return adjoint(**{**dict(nx=768, ny=768, nz=384, steps=4000, L=16, d=0.5e-3,
                             design=(128, 128, 64), eps_range=(1.0, 6.0), n_src=4,
                             probe_freqs=(5e9, 10e9, 15e9, 20e9), n_probes=8, seed=1005,
                             name="adjoint768"), **kw})



def reference_sigma_max(pi, d, order=3.0, sigma_factor=1.5):
    """setABCPML's sigma_max for one axis in float arithmetic, as the
    reference evaluates it with its own pi (fdtdrav.h:1793-1808):
    sigma_factor * (npml + 1) / (150 * pi * 1 * d)."""
    f32 = np.float32
    sigma_opt = f32(f32(order) + f32(1)) / f32(f32(f32(f32(150) * f32(pi)) * f32(1)) * f32(d))
    return float(f32(f32(sigma_factor) * f32(sigma_opt)))


def patch_antenna(steps=3000) -> Scene:
    """The reference's own scene (source.cu:17-226): a microstrip-fed patch on
    a 60 x 100 x 3 cell substrate (eps_r 2.2, sigma 4e-4), ground plane, feed
    line and patch as z-plates with sigma 1e10, a 50-ohm z-source of 7 x 3
    edges driven by a Gaussian (nc = 20), the port 10 cells down the feed
    line.  CPML 10 and 15 cells of air on every side.  Constants as the
    reference's: pi = 3.1415 (also in the CPML's sigma_max, evaluated in
    float), Courant factor 1, FP32 cell sizes.  Expressions
    follow C's arithmetic: float products, promoted to double by a double
    literal (mu0, the third plate's y maximum) or by sqrt/pow of a float
    (c, dt, as a C build evaluates them), rounded to float on assignment.

    The bounding box of source.cu:256-273 compares every maximum with `<`,
    which would keep the first object's maximum; here the maximum over the
    objects is taken (x 60 dx, y 100 dy, z 3 dz), as the loop intends.
    """
    f32 = np.float32
    eps0 = f32(8.854187817e-12)
    pi = f32(3.1415)
    mu0 = f32(float(f32(4 * pi)) * 1e-7)          # (4*pi)*1e-7: float times a double literal
    c = f32(1.0 / math.sqrt(float(f32(mu0 * eps0))))
    dx, dy, dz = f32(0.389e-3), f32(0.4e-3), f32(0.265e-3)
    dt = f32(1.0 / (float(c) * math.sqrt(1.0 / float(dx) ** 2 + 1.0 / float(dy) ** 2 + 1.0 / float(dz) ** 2)))
    pml, air = 10, 15
    o = pml + air

    def ci(coord, d):   # calc_index, fdtdrav.h:198-206
        return calc_index(coord, d)

    X, Y, Z = ci(60 * dx, dx), ci(100 * dy, dy), ci(3 * dz, dz)
    nx, ny, nz = X + 2 * o, Y + 2 * o, Z + 2 * o
    brick = Box(ijk=(o, o, o, o + ci(60 * dx, dx), o + ci(100 * dy, dy), o + ci(3 * dz, dz)),
                eps_r=(2.2, 2.2, 2.2), sigma_e=(0.0004, 0.0004, 0.0004),
                mu_r=(1.0, 1.0, 1.0), sigma_m=(1.2e-38, 1.2e-38, 1.2e-38))
    pec = [((0.0, 0.0, 0.0), (60 * dx, 100 * dy, 0.0)),
           ((7.535e-3, 0.0, 3 * dz), (9.869e-3, 50 * dy, 3 * dz)),
           ((5.445e-3, 50 * dy, 3 * dz), (17.895e-3, f32(float(50 * dy) + 16e-3), 3 * dz))]
    plates = [Plate(axis=3, ijk=(o + ci(a[0], dx), o + ci(a[1], dy), o + ci(a[2], dz),
                                 o + ci(b[0], dx), o + ci(b[1], dy), o + ci(b[2], dz)),
                    sigma=(1e10, 1e10, 1e10)) for a, b in pec]
    tau = f32(f32(20 * dy) / f32(2 * c))
    src = Source(ijk=(o + ci(7.535e-3, dx), o + ci(0.0, dy), o + ci(0.0, dz),
                      o + ci(9.869e-3, dx), o + ci(0.0, dy), o + ci(3 * dz, dz)),
                 R=50.0, wtype=WAVE_GAUSSIAN, amp=1.0, t0=float(f32(3 * tau)), tau=float(tau), freq=1e8)
    vbox = (o + ci(7.535e-3, dx), o + ci(10 * dy, dy), o + ci(0.0, dz),
            o + ci(9.869e-3, dx), o + ci(10 * dy, dy), o + ci(3 * dz, dz))
    ibox = (o + ci(7.535e-3, dx), o + ci(10 * dy, dy), o + ci(3 * dz, dz),
            o + ci(9.869e-3, dx), o + ci(10 * dy, dy), o + ci(3 * dz, dz))
    cp = Cpml(thick=(pml,) * 6, order=3.0, sigma_max=tuple(reference_sigma_max(pi, d) for d in (dx, dy, dz)),
              kappa_max=5.0, alpha_max=0.05, alpha_min=0.0, sigma_factor=1.5)
    return Scene(name="patch_antenna", nx=nx, ny=ny, nz=nz, dx=float(dx), dy=float(dy), dz=float(dz),
                 dt=float(dt), steps=steps, cpml=cp, boxes=[brick], sources=[src], plates=plates,
                 port=Port(vbox=vbox, ibox=ibox, idir=2), eps0=float(eps0), mu0=float(mu0), seed=0)

WORKLOADS = {"yee64": yee64, "dielectric256": dielectric256, "slab512": slab512,
             "adjoint128": adjoint128, "adjoint768": adjoint768}
