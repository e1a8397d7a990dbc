"""ctypes binding of the engine's C ABI (include/fdtd.h).

`Simulation` mirrors the C calls one to one; `Simulation.from_scene` builds and
finalizes a context from a `scenes.Scene`.  The library is the in-tree
fdtd_amd/libfdtd_amd.so; there is no fallback: if it is missing, importing it
raises, and every compute call goes to the HIP kernels.
"""
import ctypes as C
import os

import numpy as np

from . import scenes

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_amd.so")

FDTD_OK = 0
FDTD_E_INVALID, FDTD_E_STATE, FDTD_E_RANGE, FDTD_E_NOMEM = -1, -2, -3, -4
FDTD_E_DEVICE, FDTD_E_UNSUPPORTED, FDTD_E_NONFINITE, FDTD_E_CHECKSUM = -5, -6, -7, -8
ERRORS = {-1: "FDTD_E_INVALID", -2: "FDTD_E_STATE", -3: "FDTD_E_RANGE", -4: "FDTD_E_NOMEM",
          -5: "FDTD_E_DEVICE", -6: "FDTD_E_UNSUPPORTED", -7: "FDTD_E_NONFINITE",
          -8: "FDTD_E_CHECKSUM"}
STORAGE = {"f32": 0, "f16": 1, "bf16": 2, "bq16": 3}   # bq16: block-scaled int16 fields

# Every function the header declares, with its ctypes signature.
_P = C.c_void_p
EXPORTS = {
    "fdtd_create": (C.c_int, [_P, C.c_int, C.POINTER(C.c_void_p)]),
    "fdtd_destroy": (C.c_int, [_P]),
    "fdtd_last_error": (C.c_char_p, [_P]),
    "fdtd_set_cpml": (C.c_int, [_P, _P]),
    "fdtd_add_box": (C.c_int, [_P] + [C.c_int] * 6 + [C.POINTER(C.c_float)] * 4),
    "fdtd_add_lumped_source": (C.c_int, [_P, C.c_int] + [C.c_int] * 6 + [C.c_double, _P]),
    "fdtd_add_plate": (C.c_int, [_P, C.c_int] + [C.c_int] * 6 + [C.POINTER(C.c_float)]),
    "fdtd_add_sector": (C.c_int, [_P] + [C.c_int] * 7 + [C.c_float, C.c_int, C.c_int, C.c_float, C.c_float]),
    "fdtd_add_resistor": (C.c_int, [_P, C.c_int] + [C.c_int] * 6 + [C.c_double]),
    "fdtd_add_probe": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
    "fdtd_set_dft": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double)]),
    "fdtd_set_design_region": (C.c_int, [_P] + [C.c_int] * 6),
    "fdtd_set_design_eps": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "fdtd_finalize": (C.c_int, [_P, C.c_int]),
    "fdtd_run": (C.c_int, [_P, C.c_int]),
    "fdtd_step": (C.c_int, [_P]),
    "fdtd_reset": (C.c_int, [_P]),
    "fdtd_profile": (C.c_int, [_P, C.c_int]),
    "fdtd_get_info": (C.c_int, [_P, _P]),
    "fdtd_get_probes": (C.c_int, [_P, C.POINTER(C.c_float), C.c_size_t]),
    "fdtd_get_dft": (C.c_int, [_P, C.POINTER(C.c_double), C.c_size_t]),
    "fdtd_get_field": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float), C.c_size_t]),
    "fdtd_set_port": (C.c_int, [_P] + [C.c_int] * 14),
    "fdtd_get_port": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t]),
    "fdtd_port_spectrum": (C.c_int, [_P, C.c_double, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_size_t]),
    "fdtd_save_port": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "fdtd_checkpoint_size": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "fdtd_checkpoint_save": (C.c_int, [_P, _P, C.c_size_t]),
    "fdtd_checkpoint_restore": (C.c_int, [_P, _P, C.c_size_t]),
    "fdtd_adjoint_gradient": (C.c_int, [_P, _P, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.c_size_t]),
    "fdtd_adjoint_gradient_noalloc": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.c_size_t]),
}


class FdtdGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("courant", C.c_double),
                ("eps0", C.c_double), ("mu0", C.c_double)]


class FdtdCpml(C.Structure):
    _fields_ = [("thick", C.c_int * 6), ("order", C.c_double), ("sigma_max", C.c_double * 3),
                ("sigma_factor", C.c_double), ("kappa_max", C.c_double),
                ("alpha_max", C.c_double), ("alpha_min", C.c_double)]


class FdtdWaveform(C.Structure):
    _fields_ = [("type", C.c_int), ("amp", C.c_double), ("t0", C.c_double),
                ("tau", C.c_double), ("freq", C.c_double)]


class FdtdObjective(C.Structure):
    _fields_ = [("n_terms", C.c_int), ("probe", C.POINTER(C.c_int)),
                ("freq", C.POINTER(C.c_int)), ("weight", C.POINTER(C.c_double))]


class FdtdInfo(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("steps_done", C.c_int), ("steps_capacity", C.c_int),
                ("n_probes", C.c_int), ("n_freqs", C.c_int), ("n_sources", C.c_int),
                ("n_classes", C.c_int), ("storage", C.c_int), ("dt", C.c_double),
                ("last_run_ms", C.c_double), ("e_kernel_ms", C.c_double),
                ("h_kernel_ms", C.c_double), ("step_kernel_ms", C.c_double),
                ("interior_split", C.c_int), ("pad1", C.c_int), ("adjoint_forward_ms", C.c_double),
                ("adjoint_total_ms", C.c_double), ("adjoint_segments", C.c_int), ("pad0", C.c_int),
                ("device_bytes", C.c_uint64),
                ("bytes_per_cell_e", C.c_uint64), ("bytes_per_cell_h", C.c_uint64),
                ("adjoint_device_bytes", C.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class FdtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libfdtd_amd.so and bind every exported symbol (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `python -m fdtd_amd.build` "
                          "(there is no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    """One engine context on the current HIP device."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt=0.0, courant=0.99, eps0=0.0, mu0=0.0,
                 storage="f32"):
        self.lib = load_library()
        g = FdtdGrid(nx, ny, nz, dx, dy, dz, dt, courant, eps0, mu0)
        h = C.c_void_p()
        self.storage = storage
        self._check(self.lib.fdtd_create(C.byref(g), STORAGE[storage], C.byref(h)), ctx=None)
        self.h = h
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n_probes = 0
        self.n_freqs = 0
        self.design_cells = 0

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc, ctx=-1):
        if rc != FDTD_OK:
            hp = self.h if ctx == -1 else None
            msg = self.lib.fdtd_last_error(hp)
            raise FdtdError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fdtd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- scene ----------------------------------------------------------------
    def set_cpml(self, cp: scenes.Cpml):
        s = FdtdCpml((C.c_int * 6)(*cp.thick), cp.order, (C.c_double * 3)(*cp.sigma_max),
                     cp.sigma_factor, cp.kappa_max, cp.alpha_max, cp.alpha_min)
        self._check(self.lib.fdtd_set_cpml(self.h, C.byref(s)))

    def add_box(self, ijk, eps_r, sigma_e, mu_r=(1, 1, 1), sigma_m=(0, 0, 0)):
        arr = [(C.c_float * 3)(*v) for v in (eps_r, sigma_e, mu_r, sigma_m)]
        self._check(self.lib.fdtd_add_box(self.h, *ijk, *arr))

    def add_lumped_source(self, ijk, R, wtype, amp, t0, tau, freq=0.0, direction=3):
        w = FdtdWaveform(wtype, amp, t0, tau, freq)
        self._check(self.lib.fdtd_add_lumped_source(self.h, direction, *ijk, R, C.byref(w)))

    def add_plate(self, axis, ijk, sigma):
        sg = (C.c_float * 3)(*sigma)
        self._check(self.lib.fdtd_add_plate(self.h, axis, *ijk, sg))

    def add_sector(self, points, kz, radius, max_x, max_y, sigma_x, sigma_y):
        """points = (i0, j0, i1, j1, i2, j2), apex (i1, j1); radius in metres."""
        self._check(self.lib.fdtd_add_sector(self.h, *[int(v) for v in points], int(kz), float(radius),
                                             int(max_x), int(max_y), float(sigma_x), float(sigma_y)))

    def add_resistor(self, ijk, R, direction=3):
        self._check(self.lib.fdtd_add_resistor(self.h, direction, *ijk, float(R)))

    def add_probe(self, comp, i, j, k) -> int:
        idx = C.c_int()
        self._check(self.lib.fdtd_add_probe(self.h, comp, i, j, k, C.byref(idx)))
        self.n_probes += 1
        return idx.value

    def set_dft(self, freqs):
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        self._check(self.lib.fdtd_set_dft(self.h, len(f), _dptr(f)))
        self.n_freqs = len(f)

    def set_port(self, vbox, ibox, vdir=3, idir=3):
        """Port voltage box and current box (is, js, ks, ie, je, ke, inclusive
        indices as the reference's sampled boxes).  Adds the probes the two
        reductions read after the caller's own."""
        self._check(self.lib.fdtd_set_port(self.h, vdir, *vbox, idir, *ibox))
        self.n_probes = self.info()["n_probes"]

    def set_design_region(self, ijk):
        self._check(self.lib.fdtd_set_design_region(self.h, *ijk))
        self.design_cells = (ijk[3] - ijk[0]) * (ijk[4] - ijk[1]) * (ijk[5] - ijk[2])

    def set_design_eps(self, eps):
        e = np.ascontiguousarray(eps, dtype=np.float32).ravel()
        self._check(self.lib.fdtd_set_design_eps(self.h, _fptr(e), e.size))

    def finalize(self, n_steps):
        self._check(self.lib.fdtd_finalize(self.h, int(n_steps)))

    # -- time loop ------------------------------------------------------------
    def run(self, nsteps):
        self._check(self.lib.fdtd_run(self.h, int(nsteps)))

    def step(self):
        """One step (E then H), as one call of the reference's solver."""
        self._check(self.lib.fdtd_step(self.h))

    def profile(self, nsteps):
        self._check(self.lib.fdtd_profile(self.h, int(nsteps)))

    def reset(self):
        self._check(self.lib.fdtd_reset(self.h))

    def info(self) -> dict:
        inf = FdtdInfo()
        self._check(self.lib.fdtd_get_info(self.h, C.byref(inf)))
        return inf.as_dict()

    # -- outputs --------------------------------------------------------------
    def probes(self) -> np.ndarray:
        steps = self.info()["steps_done"]
        out = np.zeros((self.n_probes, steps), dtype=np.float32)
        self._check(self.lib.fdtd_get_probes(self.h, _fptr(out), out.size))
        return out

    def dft(self) -> np.ndarray:
        out = np.zeros((self.n_probes, self.n_freqs, 2), dtype=np.float64)
        self._check(self.lib.fdtd_get_dft(self.h, _dptr(out), out.size))
        return out[..., 0] + 1j * out[..., 1]

    def port(self):
        """(V, I): port voltage and averaged current per step (float32)."""
        steps = self.info()["steps_done"]
        v = np.zeros(steps, dtype=np.float32)
        i = np.zeros(steps, dtype=np.float32)
        self._check(self.lib.fdtd_get_port(self.h, _fptr(v), _fptr(i), steps))
        return v, i

    def port_spectrum(self, z0, freqs) -> dict:
        """V(f), I(f), S11 and Zin (complex128 arrays) at the given frequencies."""
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        out = np.zeros((len(f), 8), dtype=np.float64)
        self._check(self.lib.fdtd_port_spectrum(self.h, float(z0), len(f), _dptr(f), _dptr(out),
                                                out.size))
        c = out[:, 0::2] + 1j * out[:, 1::2]
        return {"V": c[:, 0], "I": c[:, 1], "S11": c[:, 2], "Zin": c[:, 3]}

    def save_port(self, vpath, ipath):
        """results1003.dat / results1004.dat style text files of V and I."""
        self._check(self.lib.fdtd_save_port(self.h, os.fsencode(vpath), os.fsencode(ipath)))

    def field(self, comp) -> np.ndarray:
        out = np.zeros((self.nz, self.ny, self.nx), dtype=np.float32)
        self._check(self.lib.fdtd_get_field(self.h, comp, _fptr(out), out.size))
        return out

    def checkpoint_save(self) -> bytes:
        n = C.c_size_t()
        self._check(self.lib.fdtd_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._check(self.lib.fdtd_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def checkpoint_restore(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(self.lib.fdtd_checkpoint_restore(self.h, buf, len(blob)))

    def adjoint_gradient_noalloc(self, terms, n_steps):
        """fdtd_adjoint_gradient_noalloc: the stored-history adjoint on the
        library's static device scratch.  Returns (J, grad)."""
        n = len(terms)
        obj = FdtdObjective(n, (C.c_int * n)(*[t[0] for t in terms]), (C.c_int * n)(*[t[1] for t in terms]),
                            (C.c_double * n)(*[t[2] for t in terms]))
        J = C.c_double()
        g = np.zeros(self.design_cells, dtype=np.float64)
        self._check(self.lib.fdtd_adjoint_gradient_noalloc(self.h, C.byref(obj), int(n_steps), C.byref(J),
                                                           _dptr(g), g.size))
        return J.value, g

    def adjoint_gradient(self, terms, n_steps, checkpoint_every):
        """terms: list of (probe, freq_index, weight).  Returns (J, grad)."""
        n = len(terms)
        pr = (C.c_int * n)(*[t[0] for t in terms])
        fr = (C.c_int * n)(*[t[1] for t in terms])
        wt = (C.c_double * n)(*[t[2] for t in terms])
        obj = FdtdObjective(n, pr, fr, wt)
        J = C.c_double()
        g = np.zeros(self.design_cells, dtype=np.float64)
        self._check(self.lib.fdtd_adjoint_gradient(self.h, C.byref(obj), int(n_steps),
                                                   int(checkpoint_every), C.byref(J), _dptr(g),
                                                   g.size))
        return J.value, g

    # -- scene helper -----------------------------------------------------------
    @classmethod
    def from_scene(cls, sc: scenes.Scene, storage="f32", n_steps=None, finalize=True):
        sim = cls(sc.nx, sc.ny, sc.nz, sc.dx, sc.dy, sc.dz, dt=sc.dt, eps0=sc.eps0, mu0=sc.mu0,
                  storage=storage)
        sim.set_cpml(sc.cpml)
        for b in sc.boxes:
            sim.add_box(b.ijk, b.eps_r, b.sigma_e, b.mu_r, b.sigma_m)
        for s in sc.sources:
            sim.add_lumped_source(s.ijk, s.R, s.wtype, s.amp, s.t0, s.tau, s.freq, direction=s.dir)
        for q in getattr(sc, "sectors", []):
            sim.add_sector(q.points, q.kz, q.radius, q.max_x, q.max_y, q.sigma[0], q.sigma[1])
        for p in getattr(sc, "plates", []):
            sim.add_plate(p.axis, p.ijk, p.sigma)
        for r in getattr(sc, "resistors", []):
            sim.add_resistor(r.ijk, r.R)
        for (comp, i, j, k) in sc.probes:
            sim.add_probe(comp, i, j, k)
        if getattr(sc, "port", None) is not None:
            sim.set_port(sc.port.vbox, sc.port.ibox, vdir=3, idir=sc.port.idir)
        if sc.freqs:
            sim.set_dft(sc.freqs)
        if sc.design is not None:
            sim.set_design_region(sc.design)
            if sc.design_eps is not None:
                sim.set_design_eps(sc.design_eps)
        if finalize:
            sim.finalize(n_steps or sc.steps)
        return sim
