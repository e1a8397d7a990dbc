"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The engine never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB points the binding at another build of the same source (the
# planted-mistake tests load deliberately broken copies this way).
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OracleScene(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("dt", C.c_double), ("eps0", C.c_double), ("mu0", C.c_double),
                ("pml", C.c_int * 6), ("order", C.c_double), ("sigma_max", C.c_double * 3),
                ("kappa_max", C.c_double), ("alpha_max", C.c_double), ("alpha_min", C.c_double),
                ("nbox", C.c_int), ("box", C.POINTER(C.c_int)), ("box_mat", C.POINTER(C.c_float)),
                ("nsrc", C.c_int), ("src", C.POINTER(C.c_int)), ("src_R", C.POINTER(C.c_double)),
                ("src_wtype", C.POINTER(C.c_int)), ("src_wpar", C.POINTER(C.c_double)),
                ("nprobe", C.c_int), ("probe", C.POINTER(C.c_int)),
                ("nfreq", C.c_int), ("freqs", C.POINTER(C.c_double)),
                ("design", C.c_int * 6), ("design_eps", C.POINTER(C.c_float)),
                ("nsteps", C.c_int),
                ("nplate", C.c_int), ("plate", C.POINTER(C.c_int)), ("plate_sigma", C.POINTER(C.c_float)),
                ("nres", C.c_int), ("res", C.POINTER(C.c_int)), ("res_R", C.POINTER(C.c_double)),
                ("src_dir", C.POINTER(C.c_int)),
                ("nsector", C.c_int), ("sector", C.POINTER(C.c_int)), ("sector_f", C.POINTER(C.c_float))]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        lib.oracle_create.restype = C.c_void_p
        lib.oracle_create.argtypes = [C.POINTER(OracleScene), C.c_int]
        lib.oracle_destroy.argtypes = [C.c_void_p]
        lib.oracle_steps.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_step_count.argtypes = [C.c_void_p]
        lib.oracle_set_threads.argtypes = [C.c_int]
        lib.oracle_get_probes.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.oracle_get_dft.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.oracle_get_array.restype = C.c_long
        lib.oracle_get_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        lib.oracle_adjoint.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        lib.oracle_adjoint_alloc.argtypes = [C.c_void_p]
        lib.oracle_forward_step.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_reverse_step.argtypes = [C.c_void_p, C.c_int]
        lib.oracle_set_quant.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int]
        lib.oracle_set_array.restype = C.c_long
        lib.oracle_set_array.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        _lib = lib
    return _lib


def _sigma_max(sc):
    """Resolve the CPML sigma_max the way the engine does (<= 0: reference
    default sigma_factor (m+1) / (150 pi d), fdtdrav.h:1783-1798)."""
    out = []
    for a, d in enumerate((sc.dx, sc.dy, sc.dz)):
        s = sc.cpml.sigma_max[a]
        if not s > 0:
            s = sc.cpml.sigma_factor * (sc.cpml.order + 1.0) / (150.0 * np.pi * d)
        out.append(s)
    return out


class Oracle:
    """CPU restatement of the reference for one Scene, FP32 or FP64."""

    def __init__(self, sc, precision=64, nsteps=None):
        lib = load()
        self.sc = sc
        self.nsteps = int(nsteps if nsteps is not None else sc.steps)
        # keep every array alive for the lifetime of the handle
        self._keep = []

        def arr(a, ctype, np_dtype):
            a = np.ascontiguousarray(a, dtype=np_dtype)
            if a.size == 0:
                a = np.zeros(1, dtype=np_dtype)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        boxes = [list(b.ijk) for b in sc.boxes]
        mats = [list(b.eps_r) + list(b.sigma_e) + list(b.mu_r) + list(b.sigma_m) for b in sc.boxes]
        src = [list(s.ijk) for s in sc.sources]
        wpar = [[s.amp, s.t0, s.tau, s.freq] for s in sc.sources]
        os_ = OracleScene()
        os_.nx, os_.ny, os_.nz = sc.nx, sc.ny, sc.nz
        os_.dx, os_.dy, os_.dz, os_.dt = sc.dx, sc.dy, sc.dz, sc.dt
        os_.eps0, os_.mu0 = sc.eps0, sc.mu0
        os_.pml = (C.c_int * 6)(*sc.cpml.thick)
        os_.order = sc.cpml.order
        os_.sigma_max = (C.c_double * 3)(*_sigma_max(sc))
        os_.kappa_max, os_.alpha_max, os_.alpha_min = sc.cpml.kappa_max, sc.cpml.alpha_max, sc.cpml.alpha_min
        os_.nbox = len(boxes)
        os_.box = arr(np.array(boxes, dtype=np.int32).ravel(), C.c_int, np.int32)
        os_.box_mat = arr(np.array(mats, dtype=np.float32).ravel(), C.c_float, np.float32)
        os_.nsrc = len(src)
        os_.src = arr(np.array(src, dtype=np.int32).ravel(), C.c_int, np.int32)
        os_.src_R = arr([s.R for s in sc.sources], C.c_double, np.float64)
        os_.src_wtype = arr([s.wtype for s in sc.sources], C.c_int, np.int32)
        os_.src_wpar = arr(np.array(wpar, dtype=np.float64).ravel(), C.c_double, np.float64)
        os_.src_dir = arr([getattr(s, "dir", 3) for s in sc.sources], C.c_int, np.int32)
        os_.nprobe = len(sc.probes)
        os_.probe = arr(np.array(sc.probes, dtype=np.int32).ravel(), C.c_int, np.int32)
        os_.nfreq = len(sc.freqs)
        os_.freqs = arr(sc.freqs, C.c_double, np.float64)
        if sc.design is not None and sc.design_eps is not None:
            os_.design = (C.c_int * 6)(*sc.design)
            os_.design_eps = arr(sc.design_eps, C.c_float, np.float32)
        else:
            os_.design = (C.c_int * 6)(0, 0, 0, 0, 0, 0)
            os_.design_eps = C.POINTER(C.c_float)()
        os_.nsteps = self.nsteps
        plates = getattr(sc, "plates", []) or []
        res = getattr(sc, "resistors", []) or []
        os_.nplate = len(plates)
        os_.plate = arr(np.array([[p.axis, *p.ijk] for p in plates], dtype=np.int32).ravel(), C.c_int, np.int32)
        os_.plate_sigma = arr(np.array([p.sigma for p in plates], dtype=np.float32).ravel(), C.c_float, np.float32)
        secs = getattr(sc, "sectors", []) or []
        os_.nsector = len(secs)
        os_.sector = arr(np.array([[*q.points, q.kz, q.max_x, q.max_y] for q in secs], dtype=np.int32).ravel(),
                         C.c_int, np.int32)
        os_.sector_f = arr(np.array([[q.radius, *q.sigma] for q in secs], dtype=np.float32).ravel(),
                           C.c_float, np.float32)
        os_.nres = len(res)
        os_.res = arr(np.array([r.ijk for r in res], dtype=np.int32).ravel(), C.c_int, np.int32)
        os_.res_R = arr([r.R for r in res], C.c_double, np.float64)
        self._scene = os_
        self.lib = lib
        self.h = lib.oracle_create(C.byref(os_), int(precision))
        if not self.h:
            raise ValueError("oracle rejected the scene")

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def steps(self, n):
        rc = self.lib.oracle_steps(self.h, int(n))
        if rc != 0:
            raise RuntimeError(f"oracle_steps failed ({rc})")

    @property
    def steps_done(self):
        return self.lib.oracle_step_count(self.h)

    def probes(self):
        out = np.zeros((len(self.sc.probes), self.nsteps), dtype=np.float64)
        self.lib.oracle_get_probes(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[:, :self.steps_done]

    def dft(self):
        out = np.zeros((len(self.sc.probes), len(self.sc.freqs), 2), dtype=np.float64)
        self.lib.oracle_get_dft(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[..., 0] + 1j * out[..., 1]

    def array(self, name):
        n = self.lib.oracle_get_array(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n, dtype=np.float64)
        self.lib.oracle_get_array(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def field(self, name):
        return self.array(name).reshape(self.sc.nz, self.sc.ny, self.sc.nx)

    def set_quant(self, mode, sE=1.0, sH=1.0, keep_radius=-1):
        """Round every stored E/H/psi value to FP16 (16) or BF16 (17) in
        scaled units, or to block-scaled int16 with one power-of-two scale
        per B x-cells: 18 + B for fields and psi, 40 + B for the fields with
        psi in FP16 (the engine's block-scaled storage) -- FP32 oracle only,
        a model of reduced-precision storage for error studies."""
        if self.lib.oracle_set_quant(self.h, int(mode), float(sE), float(sH), int(keep_radius)) != 0:
            raise ValueError("quantisation needs the FP32 oracle")

    def adjoint(self, terms):
        """FP64 discrete adjoint: terms = [(probe, freq_index, weight)].
        Returns (J, dJ/d eps_r per design cell, x-fastest)."""
        sc = self.sc
        n = len(terms)
        tp = (C.c_int * n)(*[t[0] for t in terms])
        tf = (C.c_int * n)(*[t[1] for t in terms])
        tw = (C.c_double * n)(*[t[2] for t in terms])
        d = sc.design
        g = np.zeros((d[3] - d[0]) * (d[4] - d[1]) * (d[5] - d[2]), dtype=np.float64)
        J = C.c_double()
        rc = self.lib.oracle_adjoint(self.h, n, tp, tf, tw, C.byref(J),
                                     g.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError("oracle adjoint needs the FP64 oracle and a design region")
        return J.value, g

    def adjoint_alloc(self):
        self.lib.oracle_adjoint_alloc(self.h)

    def forward_step(self, m):
        """One E+H step on the current arrays (no bookkeeping)."""
        self.lib.oracle_forward_step(self.h, int(m))

    def reverse_step(self, m):
        """One reverse (transposed) step on the adjoint arrays."""
        self.lib.oracle_reverse_step(self.h, int(m))

    def set_array(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        n = self.lib.oracle_set_array(self.h, name.encode(), v.ctypes.data_as(C.POINTER(C.c_double)))
        if n < 0:
            raise KeyError(name)
