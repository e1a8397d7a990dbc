"""CPU restatement of the reference's port voltage / current reductions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The engine
never uses it.

The reference samples Ez into a per-step voltage box and Hx, Hy, Hz into a
per-step current box (sampleField, fdtdrav.h:234-246, called at
fdtdrav.h:740 and 347-391), then after the run reduces them:
  voltMgpu  fdtdrav.h:2185-2298  (direction 3 only: the reference samples
                                  only Ez, see fdtdrav.h:740)
  currMgpu  fdtdrav.h:2301-2487
  correcI   fdtdrav.h:2490-2507
  c_zinS11  fdtdrav.h:2510-2546  (on the DFT of the series; calc_otm's FFT,
                                  fdtdrav.h:2988-3051, is disabled there)
The functions below follow those loop nests literally -- local indices of
the sampled boxes, the reference's `if` filters, FP32 accumulation in loop
order -- over a sampler `get(comp, i, j, k) -> series` of global cells.
"""
import numpy as np

EZ, HX, HY, HZ = 2, 3, 4, 5


def voltage(get, vbox, dz, nz, dtype=np.float32):
    """voltMgpu case 3 (fdtdrav.h:2252-2281) for every step at once, in
    dtype arithmetic (float32 as the reference; float64 for the pins)."""
    is_, js, ks, ie, je, ke = vbox
    volt_nx = ie - is_ + 1
    f32 = dtype
    _is, _js, _ie, _je = f32(is_), f32(js), f32(ie), f32(je)
    csuf = f32(-f32(dz)) / f32((_ie - _is + f32(1)) * (_je - _js + f32(1)))
    acc = None
    for k in range(nz):                              # local k, gpu_offset = 0
        for j in range(js, je + 1):
            for i in range(is_, ie + 1):
                if ks <= k < ke:
                    s = np.asarray(get(EZ, i, j, k), dtype=dtype)
                    acc = s.copy() if acc is None else (acc + s).astype(dtype)
    assert volt_nx > 0
    return (acc * csuf).astype(dtype)


def current_raw(get, ibox, direction, dx, dy, dz, nz, dtype=np.float32):
    """currMgpu (fdtdrav.h:2301-2487) before correcI.  Local index 0 of the
    sampled H box is (is-1, js-1); the box is current_NX x current_NY."""
    is_, js, ks, ie, je, ke = ibox
    cnx, cny = ie - is_ + 2, je - js + 2

    def H(comp, il, jl, k):   # local (il, jl) of the sampled box -> global cell
        return np.asarray(get(comp, is_ - 1 + il, js - 1 + jl, k), dtype=dtype)

    z = None
    sums = {}

    def acc(name, v):
        sums[name] = v.copy() if name not in sums else (sums[name] + v).astype(dtype)

    f32 = dtype
    if direction == 1:
        for k in range(nz):
            for j in range(1, cny):
                if k == ks - 1:
                    acc("yp", H(HY, cnx - 2, j, k))
                if k == ke:
                    acc("yn", H(HY, cnx - 2, j, k))
            if ks <= k <= ke:
                acc("zp", H(HZ, cnx - 2, cny - 1, k))
                acc("zn", H(HZ, cnx - 2, 0, k))
        z = (((f32(dy) * sums["yp"] + f32(dz) * sums["zp"]) - f32(dy) * sums["yn"]) - f32(dz) * sums["zn"])
    elif direction == 2:
        for k in range(nz):
            if ks <= k <= ke:
                acc("zp", H(HZ, 0, cny - 2, k))
                acc("zn", H(HZ, cnx - 1, cny - 2, k))
            for i in range(1, cnx):
                if k == ke:
                    acc("xp", H(HX, i, cny - 2, k))
                if k == ks - 1:
                    acc("xn", H(HX, i, cny - 2, k))
        z = (((f32(dz) * sums["zp"] + f32(dx) * sums["xp"]) - f32(dz) * sums["zn"]) - f32(dx) * sums["xn"])
    elif direction == 3:
        for k in range(nz):
            for i in range(1, cnx):
                if k == ke - 1:
                    acc("xp", H(HX, i, 0, k))
                if k == ke - 1:
                    acc("xn", H(HX, i, cny - 1, k))
            for j in range(1, cny):
                if k == ke - 1:
                    acc("yp", H(HY, cnx - 1, j, k))
                if k == ke - 1:
                    acc("yn", H(HY, 0, j, k))
        z = (((f32(dy) * sums["yp"] + f32(dx) * sums["xp"]) - f32(dy) * sums["yn"]) - f32(dx) * sums["xn"])
    else:
        raise ValueError("direction must be 1, 2 or 3")
    return np.asarray(z, dtype=dtype)


def correct(raw, dtype=np.float32):
    """correcI (fdtdrav.h:2490-2507), reading the samples as they were
    before the pass."""
    raw = np.asarray(raw, dtype=dtype)
    n = raw.size
    out = np.empty_like(raw)
    if n == 1:
        out[:] = raw
        return out
    out[:-1] = ((raw[:-1] + raw[1:]) / dtype(2)).astype(dtype)
    out[-1] = dtype((raw[-1] + raw[-2]) / dtype(2))
    return out


def cells(vbox, ibox, direction, nz):
    """Every (comp, i, j, k) the reductions read (to place probes)."""
    seen = []

    def rec(comp, i, j, k):
        if (comp, i, j, k) not in seen:
            seen.append((comp, i, j, k))
        return np.zeros(1, dtype=np.float32)

    voltage(rec, vbox, 1.0, nz)
    current_raw(rec, ibox, direction, 1.0, 1.0, 1.0, nz)
    return seen


def spectrum(v, cur, dt, z0, freqs):
    """V(f), I(f) as sum_m x[m] exp(-2 pi i f m dt), then c_zinS11's a, b,
    S11 = b / a and Zin = z0 (1 + S11) / (1 - S11) (fdtdrav.h:2510-2546), in
    FP64."""
    m = np.arange(len(v), dtype=np.float64)
    out = {"V": [], "I": [], "S11": [], "Zin": []}
    for f in freqs:
        e = np.exp(-2j * np.pi * f * m * dt)
        V = np.sum(np.asarray(v, np.float64) * e)
        I = np.sum(np.asarray(cur, np.float64) * e)
        den = 2.0 * np.sqrt(z0)
        a = (V + z0 * I) / den
        b = (V - z0 * I) / den
        s11 = b / a
        out["V"].append(V)
        out["I"].append(I)
        out["S11"].append(s11)
        out["Zin"].append(z0 * (1 + s11) / (1 - s11))
    return {k: np.array(x) for k, x in out.items()}
