// adjoint.hip -- reverse-time (discrete adjoint) kernels for gfx950.
//
// The forward time step is linear in (E, H, psi_e, psi_h), so its adjoint is
// the exact transpose.  Reverse step m, for the state after forward step m:
//   adj_e_kernel  (E cells, backward differences, like the forward E sweep)
//       [finalize of the E half-step of step m+1: psi_e adjoint, dJ/deps on
//        the design edges, dJ/dE <- Ce dJ/dE]  then the transpose of the H
//        half-step of step m gathered into dJ/dE, then dJ/d(probe sample m).
//   adj_h_kernel  (H cells, forward differences, like the forward H sweep)
//       [dJ/dH <- Ch dJ/dH, psi_h adjoint] then the transpose of the E
//        half-step of step m gathered into dJ/dH, with the lumped-source
//        edges' own Ez coefficients where the gather meets one.
// Every transposed term is a gather, so no two threads write one address and
// no atomics are needed.  Forward operators: fdtdrav.h:665-911 (E) and
// fdtdrav.h:328-562 (H); edge averaging fdtdrav.h:1135-1177.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "device_common.h"
#include "internal.h"

namespace fdtd {

namespace {

const float4 NOPROF = {0.f, 0.f, 1.f, 0.f};

__device__ __forceinline__ bool xs_h(const KArgs &a, int i) { return i < a.L[0] || (i >= a.nx - a.L[1] && i < a.nx - 1); }
__device__ __forceinline__ bool ys_h(const KArgs &a, int j) { return j < a.L[2] || (j >= a.ny - a.L[3] && j < a.ny - 1); }
__device__ __forceinline__ bool zs_h(const KArgs &a, int k) { return k < a.L[4] || (k >= a.nz - a.L[5] && k < a.nz - 1); }
__device__ __forceinline__ bool xs_e(const KArgs &a, int i) { return (i > 0 && i < a.L[0]) || (i >= a.nx - a.L[1] && i < a.nx - 1); }
__device__ __forceinline__ bool ys_e(const KArgs &a, int j) { return (j > 0 && j < a.L[2]) || (j >= a.ny - a.L[3] && j < a.ny - 1); }
__device__ __forceinline__ bool zs_e(const KArgs &a, int k) { return (k > 0 && k < a.L[4]) || (k >= a.nz - a.L[5] && k < a.nz - 1); }
__device__ __forceinline__ int slab_x(const KArgs &a, int i) { return i < a.L[0] ? i : i - (a.nx - a.L[1]) + a.L[0]; }
__device__ __forceinline__ int slab_y(const KArgs &a, int j) { return j < a.L[2] ? j : j - (a.ny - a.L[3]) + a.L[2]; }
__device__ __forceinline__ int slab_z(const KArgs &a, int k) { return k < a.L[4] ? k : k - (a.nz - a.L[5]) + a.L[4]; }
// element of cell (i, j, k) in the x / y / z slab arrays of psi
__device__ __forceinline__ i64 psi_sx(const KArgs &a, int i, int j, int k) { return ((i64)k * a.ny + j) * a.Lx + slab_x(a, i); }
__device__ __forceinline__ i64 psi_sy(const KArgs &a, int i, int j, int k) { return ((i64)k * a.Ly + slab_y(a, j)) * a.px + i; }
__device__ __forceinline__ i64 psi_sz(const KArgs &a, int i, int j, int k) { return ((i64)slab_z(a, k) * a.ny + j) * a.px + i; }

// The adjoint psi values of one cell, two per slab axis (E cells: pb[0], pb[1]
// in x, pb[4], pb[5] in y, pb[8], pb[9] in z; H cells: pb[2], pb[3], pb[6],
// pb[7], pb[10], pb[11]).  The transposed quantities take them as arguments
// so that the sweeps can read them with the rest of a plane (a read where
// they are used waits for every load in flight, the next plane's included).
struct Psi6 {
    float x0, x1, y0, y1, z0, z1;
};

// psi of H cell (i, j, k) on the slabs it is in (zero elsewhere)
template <int V>
__device__ __forceinline__ Psi6 psih_at(const KArgs &a, const AArgs &b, int i, int j, int k)
{
    Psi6 p{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if ((V & V_X) && xs_h(a, i)) { const i64 q = psi_sx(a, i, j, k); p.x0 = b.pb[2][q]; p.x1 = b.pb[3][q]; }
    if ((V & V_Y) && ys_h(a, j)) { const i64 q = psi_sy(a, i, j, k); p.y0 = b.pb[6][q]; p.y1 = b.pb[7][q]; }
    if ((V & V_Z) && zs_h(a, k)) { const i64 q = psi_sz(a, i, j, k); p.z0 = b.pb[10][q]; p.z1 = b.pb[11][q]; }
    return p;
}
// psi of E cell (i, j, k) on the slabs it is in (zero elsewhere)
template <int V>
__device__ __forceinline__ Psi6 psie_at(const KArgs &a, const AArgs &b, int i, int j, int k)
{
    Psi6 p{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if ((V & V_X) && xs_e(a, i)) { const i64 q = psi_sx(a, i, j, k); p.x0 = b.pb[0][q]; p.x1 = b.pb[1][q]; }
    if ((V & V_Y) && ys_e(a, j)) { const i64 q = psi_sy(a, i, j, k); p.y0 = b.pb[4][q]; p.y1 = b.pb[5][q]; }
    if ((V & V_Z) && zs_e(a, k)) { const i64 q = psi_sz(a, i, j, k); p.z0 = b.pb[8][q]; p.z1 = b.pb[9][q]; }
    return p;
}
// The same, issued on every path (the two-deep pipelines): a lane that needs
// no value of an axis (need false, or the cell off that slab) reads element
// 0 of the array, so that the compiler's waits count the same loads on every
// path; only values on the cell's slabs are used.
template <int V>
__device__ __forceinline__ Psi6 psih_all(const KArgs &a, const AArgs &b, bool need, int i, int j, int k)
{
    Psi6 p{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (V & V_X) { const i64 q = need && xs_h(a, i) ? psi_sx(a, i, j, k) : 0; p.x0 = b.pb[2][q]; p.x1 = b.pb[3][q]; }
    if (V & V_Y) { const i64 q = need && ys_h(a, j) ? psi_sy(a, i, j, k) : 0; p.y0 = b.pb[6][q]; p.y1 = b.pb[7][q]; }
    if (V & V_Z) { const i64 q = need && zs_h(a, k) ? psi_sz(a, i, j, k) : 0; p.z0 = b.pb[10][q]; p.z1 = b.pb[11][q]; }
    return p;
}
template <int V>
__device__ __forceinline__ Psi6 psie_all(const KArgs &a, const AArgs &b, bool need, int i, int j, int k)
{
    Psi6 p{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (V & V_X) { const i64 q = need && xs_e(a, i) ? psi_sx(a, i, j, k) : 0; p.x0 = b.pb[0][q]; p.x1 = b.pb[1][q]; }
    if (V & V_Y) { const i64 q = need && ys_e(a, j) ? psi_sy(a, i, j, k) : 0; p.y0 = b.pb[4][q]; p.y1 = b.pb[5][q]; }
    if (V & V_Z) { const i64 q = need && zs_e(a, k) ? psi_sz(a, i, j, k) : 0; p.z0 = b.pb[8][q]; p.z1 = b.pb[9][q]; }
    return p;
}
// conditional loads (the one-plane-at-a-time marches): only what is needed
template <int V, bool PIPE>
__device__ __forceinline__ Psi6 psih_get(const KArgs &a, const AArgs &b, bool need, int i, int j, int k)
{
    if (PIPE) return psih_all<V>(a, b, need, i, j, k);
    return need ? psih_at<V>(a, b, i, j, k) : Psi6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
}
template <int V, bool PIPE>
__device__ __forceinline__ Psi6 psie_get(const KArgs &a, const AArgs &b, bool need, int i, int j, int k)
{
    if (PIPE) return psie_all<V>(a, b, need, i, j, k);
    return need ? psie_at<V>(a, b, i, j, k) : Psi6{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
}

template <int CM>
__device__ __forceinline__ void coefs(const KArgs &a, const float4 *tab, bool uni, i64 c, float4 &cx, float4 &cy,
                                      float4 &cz)
{
    const int cl = uni ? 0 : class_of<CM>(a.cls, c);
    cx = tab[cl * 3 + 0];
    cy = tab[cl * 3 + 1];
    cz = tab[cl * 3 + 2];
}

// Transposed H half-step quantities of H cell (i,j,k): the factors that
// multiply each E difference it reads (Hx: dzEy -> sy, dyEz -> sz; Hy: dxEz
// -> tz, dzEx -> tx; Hz: dyEx -> rx, dxEy -> ry), main term plus psi_h.
struct HQ {
    float sy, sz, tz, tx, rx, ry;
};

// From the cell's H coefficients and its dJ/dH (and the psi_h adjoint in
// memory).
// Region variant V: bit V_X / V_Y / V_Z set when the cell can meet an x / y
// / z CPML slab or outer face (the predicates of the other axes compile out).
template <int V>
__device__ __forceinline__ HQ hq_of(const KArgs &a, const float4 &hx, const float4 &hy, const float4 &hz, float hbx,
                                    float hby, float hbz, int i, int j, int k, const Psi6 &ps)
{
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;
    HQ q{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    const bool ihi = !PX || i < nx - 1, jhi = !PY || j < ny - 1, khi = !PZ || k < nz - 1;
    const bool ux = jhi && khi, uy = ihi && khi;
    const bool uz = ihi && jhi;
    const bool xs = PX && xs_h(a, i), ys = PY && ys_h(a, j), zs = PZ && zs_h(a, k);
    const float4 pxm = xs ? a.px_prof[nx + i] : NOPROF;
    const float4 pym = ys ? a.py_prof[ny + j] : NOPROF;
    const float4 pzm = zs ? a.pz_prof[nz + k] : NOPROF;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    // curl coefficients over kappa, divided on the slab cells only (kappa is
    // 1 elsewhere and x / 1 = x exactly)
    float hxy = hx.y, hxz = hx.z, hyy = hy.y, hyz = hy.z, hzy = hz.y, hzz = hz.z;
    if (zs) { hxy = hxy / pzm.z; hyz = hyz / pzm.z; }
    if (ys) { hxz = hxz / pym.z; hzy = hzy / pym.z; }
    if (xs) { hyy = hyy / pxm.z; hzz = hzz / pxm.z; }
    if (ux) {
        q.sy = hxy * hbx;
        q.sz = hxz * hbx;
    }
    if (uy) {
        q.tz = hyy * hby;
        q.tx = hyz * hby;
    }
    if (uz) {
        q.rx = hzy * hbz;
        q.ry = hzz * hbz;
    }
    if (xs) {
        q.tz += pxm.y * (ps.x0 + (uy ? (hy.y * dx) * hby : 0.f));   // psi_hyx
        q.ry += pxm.y * (ps.x1 + (uz ? (hz.z * dx) * hbz : 0.f));   // psi_hzx
    }
    if (ys) {
        q.sz += pym.y * (ps.y0 + (ux ? (hx.z * dy) * hbx : 0.f));   // psi_hxy
        q.rx += pym.y * (ps.y1 + (uz ? (hz.y * dy) * hbz : 0.f));   // psi_hzy
    }
    if (zs) {
        q.sy += pzm.y * (ps.z0 + (ux ? (hx.y * dz) * hbx : 0.f));   // psi_hxz
        q.tx += pzm.y * (ps.z1 + (uy ? (hy.z * dz) * hby : 0.f));   // psi_hyz
    }
    return q;
}

// Transposed E half-step quantities of E cell (i,j,k): Ex: dyHz -> ax, dzHy
// -> bx; Ey: dzHx -> ay, dxHz -> by; Ez: dxHy -> az, dyHx -> bz.
struct EQ {
    float ax, bx, ay, by, az, bz;
};

// From the cell's E coefficients and its dJ/dE (and the psi_e adjoint in
// memory).
template <int V>
__device__ __forceinline__ EQ eq_of(const KArgs &a, const float4 &ex, const float4 &ey, const float4 &ez, float ebx,
                                    float eby, float ebz, int i, int j, int k, const Psi6 &ps)
{
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;
    EQ q{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    if ((PX && i >= nx - 1) || (PY && j >= ny - 1) || (PZ && k >= nz - 1)) return q;   // no E update, no psi_e
    const bool ilo = !PX || i > 0, jlo = !PY || j > 0, klo = !PZ || k > 0;
    const bool ux = jlo && klo, uy = ilo && klo, uz = ilo && jlo;
    const bool xs = PX && xs_e(a, i), ys = PY && ys_e(a, j), zs = PZ && zs_e(a, k);
    const float4 pxe = xs ? a.px_prof[i] : NOPROF;
    const float4 pye = ys ? a.py_prof[j] : NOPROF;
    const float4 pze = zs ? a.pz_prof[k] : NOPROF;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    float exy = ex.y, exz = ex.z, eyy = ey.y, eyz = ey.z, ezy = ez.y, ezz = ez.z;
    if (ys) { exy = exy / pye.z; ezz = ezz / pye.z; }
    if (zs) { exz = exz / pze.z; eyy = eyy / pze.z; }
    if (xs) { eyz = eyz / pxe.z; ezy = ezy / pxe.z; }
    if (ux) {
        q.ax = exy * ebx;
        q.bx = exz * ebx;
    }
    if (uy) {
        q.ay = eyy * eby;
        q.by = eyz * eby;
    }
    if (uz) {
        q.az = ezy * ebz;
        q.bz = ezz * ebz;
    }
    if (xs) {
        q.by += pxe.y * (ps.x0 + (uy ? (ey.z * dx) * eby : 0.f));   // psi_eyx
        q.az += pxe.y * (ps.x1 + (uz ? (ez.y * dx) * ebz : 0.f));   // psi_ezx
    }
    if (ys) {
        q.ax += pye.y * (ps.y0 + (ux ? (ex.y * dy) * ebx : 0.f));   // psi_exy
        q.bz += pye.y * (ps.y1 + (uz ? (ez.z * dy) * ebz : 0.f));   // psi_ezy
    }
    if (zs) {
        q.bx += pze.y * (ps.z0 + (ux ? (ex.z * dz) * ebx : 0.f));   // psi_exz
        q.ay += pze.y * (ps.z1 + (uy ? (ey.y * dz) * eby : 0.f));   // psi_eyz
    }
    return q;
}

// The same quantities with the cell's coefficients and adjoint read from
// memory (the wide-grid sweeps, which compute them where they gather).
template <int CM, int V>
__device__ HQ hquant(const KArgs &a, const AArgs &b, bool huni, int i, int j, int k)
{
    const i64 c = (i64)k * a.plane + (i64)j * a.px + i;
    float4 hx, hy, hz;
    coefs<CM>(a, a.htab, huni, c, hx, hy, hz);
    return hq_of<V>(a, hx, hy, hz, b.hb[0][c], b.hb[1][c], b.hb[2][c], i, j, k, psih_at<V>(a, b, i, j, k));
}

template <int CM, int V>
__device__ EQ equant(const KArgs &a, const AArgs &b, bool euni, int i, int j, int k)
{
    if (((V & V_X) && i >= a.nx - 1) || ((V & V_Y) && j >= a.ny - 1) || ((V & V_Z) && k >= a.nz - 1))
        return EQ{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const i64 c = (i64)k * a.plane + (i64)j * a.px + i;
    float4 ex, ey, ez;
    coefs<CM>(a, a.etab, euni, c, ex, ey, ez);
    // dJ/dE after this step's E sweep (the H sweep's operand)
    return eq_of<V>(a, ex, ey, ez, b.ebn[0][c], b.ebn[1][c], b.ebn[2][c], i, j, k, psie_at<V>(a, b, i, j, k));
}

// First lumped-source event of column `col` at or after event `ev` (probe
// events skipped), or -1.  The event list ends every column with a record of
// another column and ends with a terminator.
__device__ __forceinline__ int src_from(const KArgs &a, int ev, int col)
{
    if (ev < 0) return -1;
    while (a.events[ev].col == col && a.events[ev].kind != EV_SOURCE) ++ev;
    return a.events[ev].col == col ? ev : -1;
}

// equant used the material Ez curl coefficients on every edge; a lumped-
// source edge updates Ez with its own (defineSourceZ, fdtdrav.h:1402-1420,
// written after the CPML's kappa division).  The differences times dJ/dEz of
// source edge `id` at (i, j, k): x = the term of dxHy, y = the term of dyHx.
template <int CM>
__device__ float2 src_terms(const KArgs &a, const AArgs &b, bool euni, int i, int j, int k, int id)
{
    const i64 c = (i64)k * a.plane + (i64)j * a.px + i;
    float4 ex, ey, ez;
    coefs<CM>(a, a.etab, euni, c, ex, ey, ez);
    const float4 pxe = xs_e(a, i) ? a.px_prof[i] : NOPROF;
    const float4 pye = ys_e(a, j) ? a.py_prof[j] : NOPROF;
    const SrcEdge se = a.srcs[id];
    const float ebz = b.ebn[2][c];   // after this step's E sweep
    return make_float2(se.cezhy * ebz - (ez.y / pxe.z) * ebz, se.cezhx * ebz - (ez.z / pye.z) * ebz);
}

}  // namespace

__device__ __forceinline__ HQ ld_hq(const AArgs &b, i64 c)
{
    return HQ{b.hq[0][c], b.hq[1][c], b.hq[2][c], b.hq[3][c], b.hq[4][c], b.hq[5][c]};
}
__device__ __forceinline__ void st_hq(const AArgs &b, i64 c, const HQ &q)
{
    b.hq[0][c] = q.sy; b.hq[1][c] = q.sz; b.hq[2][c] = q.tz;
    b.hq[3][c] = q.tx; b.hq[4][c] = q.rx; b.hq[5][c] = q.ry;
}
__device__ __forceinline__ EQ ld_eq(const AArgs &b, i64 c)
{
    return EQ{b.eq[0][c], b.eq[1][c], b.eq[2][c], b.eq[3][c], b.eq[4][c], b.eq[5][c]};
}
__device__ __forceinline__ void st_eq(const AArgs &b, i64 c, const EQ &q)
{
    b.eq[0][c] = q.ax; b.eq[1][c] = q.bx; b.eq[2][c] = q.ay;
    b.eq[3][c] = q.by; b.eq[4][c] = q.az; b.eq[5][c] = q.bz;
}

// Whether the adjoint sweeps of rows nx cells wide store every cell's
// transposed quantities (narrow grids, where the sweeps are latency-bound:
// plain loads replace the neighbours' coefficient lookups) or compute them
// where they gather and share them through LDS (wide grids, where the stored
// arrays' traffic costs more).  Measured: 128x128x64, 1000 steps, stored
// history 98.1 -> 83.0 ms; 768x768x384 FP16, 4000 steps, checkpointed 45.5 s
// computing against 50.6 s storing.  FDTD_ADJ_GATHER=stored / lds overrides.
bool adj_stores_quantities(int nx)
{
    const char *ov = std::getenv("FDTD_ADJ_GATHER");
    if (ov && std::strcmp(ov, "stored") == 0) return true;
    if (ov && std::strcmp(ov, "lds") == 0) return false;
    return nx < 8 * TX;
}

// The adjoint sweeps, one thread per (i, j) column marching a z-block.  Each
// cell's transposed quantities are computed once, by the thread that owns
// it, at the end of the sweep that updates its adjoint (EQ in the E sweep,
// HQ in the H sweep), and stored: the other sweep gathers them from its own
// cell and its neighbours ((i-1, j, k), (i, j-1, k), (i, j, k-1) for the E
// sweep, (i+1, j, k), (i, j+1, k), (i, j, k+1) for the H sweep) as plain
// loads, with no coefficient lookups of the neighbours.  The stored values
// are those a recomputation would give, bit for bit.

// fstep >= 0: finalize the E half-step of forward step fstep first.
// sten: gather the transposed H half-step of step m and inject dJ/ds[m].
template <int CM, int V>
__global__ void __launch_bounds__(TX * BY)
    adj_e_kernel(const KArgs a, const AArgs b, const int m, const int fstep, const int sten, const int euni)
{
    constexpr bool IN = V == V_IN;   // interior tiles: no slab, face or range predicates
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;   // predicated axes
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    const int i = tx * TX + (int)(threadIdx.x & (TX - 1));
    const int j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    if ((PX && i >= nx) || (PY && j >= ny)) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const bool ilo = !PX || i > 0, ihi = !PX || i < nx - 1, jlo = !PY || j > 0, jhi = !PY || j < ny - 1;
    const int mycol = j * nx + i;
    int ev = first_event(a, a.e_first, mycol, k0);
    int ek = ev >= 0 ? a.events[ev].k : INT_MAX;
    const bool in_box = i >= b.sb[0] && i < b.sb[0] + b.sb[3] && j >= b.sb[1] && j < b.sb[1] + b.sb[4];
    const i64 nb = (i64)b.sb[3] * b.sb[4] * b.sb[5];
    const bool xs = PX && xs_e(a, i), ys = PY && ys_e(a, j);
    const float4 pxe = xs ? a.px_prof[i] : NOPROF;
    const float4 pye = ys ? a.py_prof[j] : NOPROF;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    const HQ zq{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

    HQ qkm = zq;
    if (sten && k0 > 0) qkm = ld_hq(b, (i64)(k0 - 1) * a.plane + (i64)j * px + i);
    for (int k = k0; k < k1; ++k) {
        const i64 c = (i64)k * a.plane + (i64)j * px + i;
        const bool klo = !PZ || k > 0, khi = !PZ || k < nz - 1;
        const bool ux = ihi && jlo && jhi && klo && khi;
        const bool uy = ilo && ihi && jhi && klo && khi;
        const bool uz = ilo && ihi && jlo && jhi && khi;
        float ebx = b.eb[0][c], eby = b.eb[1][c], ebz = b.eb[2][c];
        float4 ex, ey, ez;
        coefs<CM>(a, a.etab, euni, c, ex, ey, ez);
        // source edge and probe events of this cell
        float cez_src = 0.f;
        bool is_src = false;
        int pev = -1;
        if (k == ek) {
            pev = ev;
            while (a.events[ev].k == k && a.events[ev].col == mycol) {
                if (a.events[ev].kind == EV_SOURCE) {
                    is_src = true;
                    cez_src = a.srcs[a.events[ev].id].ceze;
                }
                ++ev;
            }
            ek = (a.events[ev].col == mycol) ? a.events[ev].k : INT_MAX;
        }
        if (fstep >= 0 && (ux || uy || uz)) {
            const bool zs = PZ && zs_e(a, k);
            if (xs && jhi && khi) {
                const i64 s = ((i64)k * ny + j) * a.Lx + slab_x(a, i);
                b.pb[0][s] = pxe.x * (b.pb[0][s] + (uy ? (ey.z * dx) * eby : 0.f));
                b.pb[1][s] = pxe.x * (b.pb[1][s] + (uz ? (ez.y * dx) * ebz : 0.f));
            }
            if (ys && ihi && khi) {
                const i64 s = ((i64)k * a.Ly + slab_y(a, j)) * px + i;
                b.pb[4][s] = pye.x * (b.pb[4][s] + (ux ? (ex.y * dy) * ebx : 0.f));
                b.pb[5][s] = pye.x * (b.pb[5][s] + (uz ? (ez.z * dy) * ebz : 0.f));
            }
            if (zs && ihi && jhi) {
                const float4 pze = a.pz_prof[k];
                const i64 s = ((i64)slab_z(a, k) * ny + j) * px + i;
                b.pb[8][s] = pze.x * (b.pb[8][s] + (ux ? (ex.z * dz) * ebx : 0.f));
                b.pb[9][s] = pze.x * (b.pb[9][s] + (uy ? (ey.y * dz) * eby : 0.f));
            }
            if (in_box && k >= b.sb[2] && k < b.sb[2] + b.sb[5]) {
                const i64 q = (i - b.sb[0]) + (i64)b.sb[3] * ((j - b.sb[1]) + (i64)b.sb[4] * (k - b.sb[2]));
                const float *h0 = b.hist + (i64)(fstep - b.hist_base) * 3 * nb;
                const float *h1 = h0 + 3 * nb;
                const float e_[3] = {ebx, eby, ebz};
                const bool u_[3] = {ux, uy, uz};
                for (int cc = 0; cc < 3; ++cc) {
                    const float4 dc = b.dcoef[cc * nb + q];
                    if (!u_[cc] || (dc.x == 0.f && dc.y == 0.f)) continue;
                    const double em = h0[cc * nb + q], ep = h1[cc * nb + q];
                    const double dE = (double)dc.x * em - (double)dc.y * (ep - (double)dc.z * em);
                    b.gedge[cc * nb + q] += (double)e_[cc] * dE;
                }
            }
            if (ux) ebx = ex.x * ebx;
            if (uy) eby = ey.x * eby;
            if (uz) ebz = (is_src ? cez_src : ez.x) * ebz;
        }
        if (sten) {
            // this cell, (i-1, j, k) and (i, j-1, k)
            const HQ q0 = ld_hq(b, c);
            const HQ qi = ilo && (uy || uz) ? ld_hq(b, c - 1) : zq;
            const HQ qj = jlo && (ux || uz) ? ld_hq(b, c - px) : zq;
            if (ux) {
                ebx += (qkm.tx - q0.tx) + (qj.rx - q0.rx);
                if (uz) ebz += qj.sz - q0.sz;
            }
            if (uy) eby += (qkm.sy - q0.sy) + (qi.ry - q0.ry);
            if (uz) ebz += qi.tz - q0.tz;
            if (uz && !ux) ebz += qj.sz - q0.sz;
            qkm = q0;
            if (pev >= 0) {   // dJ/d(probe sample m) of E probes on this cell
                for (int e = pev; a.events[e].k == k && a.events[e].col == mycol; ++e) {
                    const int kind = a.events[e].kind;
                    if (kind < EV_PROBE || kind > EV_PROBE + 2) continue;
                    const float gv = b.g[(i64)a.events[e].id * b.gsteps + m];
                    if (kind == EV_PROBE) ebx += gv;
                    else if (kind == EV_PROBE + 1) eby += gv;
                    else ebz += gv;
                }
            }
        }
        if (ux) b.eb[0][c] = ebx;
        if (uy) b.eb[1][c] = eby;
        if (uz) b.eb[2][c] = ebz;
        // the transposed E quantities of this cell, for the H sweep's gather
        st_eq(b, c, eq_of<V>(a, ex, ey, ez, ebx, eby, ebz, i, j, k, psie_at<V>(a, b, i, j, k)));
    }
}

template <int CM, int V>
__global__ void __launch_bounds__(TX * BY) adj_h_kernel(const KArgs a, const AArgs b, const int euni, const int huni)
{
    constexpr bool IN = V == V_IN;   // interior tiles: no slab, face or range predicates
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;   // predicated axes
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    const int i = tx * TX + (int)(threadIdx.x & (TX - 1));
    const int j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    if ((PX && i >= nx) || (PY && j >= ny)) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const bool ihi = !PX || i < nx - 1, jhi = !PY || j < ny - 1;
    const bool xs = PX && xs_h(a, i), ys = PY && ys_h(a, j);
    const float4 pxm = xs ? a.px_prof[nx + i] : NOPROF;
    const float4 pym = ys ? a.py_prof[ny + j] : NOPROF;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    const EQ zq{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    EQ q0 = ld_eq(b, (i64)k0 * a.plane + (i64)j * px + i);
    // lumped-source edges among the Ez edges this thread gathers from: its
    // own cell, (i+1, j) and (i, j+1); event cursor and its plane per column.
    // Rows outside the source rows skip the lookups (no load: a kernel
    // argument range).
    const bool srow = j >= b.src_j0 && j <= b.src_j1;
    const int scol[3] = {j * nx + i, j * nx + i + 1, (j + 1) * nx + i};
    const bool sok[3] = {srow, srow && i + 1 < nx, srow && j + 1 < ny};
    int sev[3], sk[3];
    for (int q = 0; q < 3; ++q) {
        sev[q] = sok[q] ? src_from(a, first_event(a, a.e_first, scol[q], k0), scol[q]) : -1;
        sk[q] = sev[q] >= 0 ? a.events[sev[q]].k : INT_MAX;
    }
    for (int k = k0; k < k1; ++k) {
        const i64 c = (i64)k * a.plane + (i64)j * px + i;
        const bool khi = !PZ || k < nz - 1;
        const bool ux = jhi && khi, uy = ihi && khi, uz = ihi && jhi;
        const EQ qk = khi ? ld_eq(b, c + a.plane) : zq;
        float4 hx, hy, hz;
        coefs<CM>(a, a.htab, huni, c, hx, hy, hz);
        float hbx = b.hb[0][c], hby = b.hb[1][c], hbz = b.hb[2][c];
        // psi_h adjoint and dJ/dH <- Ch dJ/dH (the H half-step, pointwise part)
        if (xs) {
            const i64 s = ((i64)k * ny + j) * a.Lx + slab_x(a, i);
            b.pb[2][s] = pxm.x * (b.pb[2][s] + (uy ? (hy.y * dx) * hby : 0.f));
            b.pb[3][s] = pxm.x * (b.pb[3][s] + (uz ? (hz.z * dx) * hbz : 0.f));
        }
        if (ys) {
            const i64 s = ((i64)k * a.Ly + slab_y(a, j)) * px + i;
            b.pb[6][s] = pym.x * (b.pb[6][s] + (ux ? (hx.z * dy) * hbx : 0.f));
            b.pb[7][s] = pym.x * (b.pb[7][s] + (uz ? (hz.y * dy) * hbz : 0.f));
        }
        if (PZ && zs_h(a, k)) {
            const float4 pzm = a.pz_prof[nz + k];
            const i64 s = ((i64)slab_z(a, k) * ny + j) * px + i;
            b.pb[10][s] = pzm.x * (b.pb[10][s] + (ux ? (hx.y * dz) * hbx : 0.f));
            b.pb[11][s] = pzm.x * (b.pb[11][s] + (uy ? (hy.z * dz) * hby : 0.f));
        }
        if (ux) hbx = hx.x * hbx;
        if (uy) hby = hy.x * hby;
        if (uz) hbz = hz.x * hbz;
        // transposed E half-step gathered from the E cells that read this H:
        // this cell, (i+1, j, k) and (i, j+1, k)
        {
            const EQ qi = ihi && (uy || uz) ? ld_eq(b, c + 1) : zq;
            const EQ qj = jhi && (ux || uz) ? ld_eq(b, c + px) : zq;
            if (ux) {
                hbx += (q0.ay - qk.ay) + (q0.bz - qj.bz);
                if (uz) hbz += q0.ax - qj.ax;
            }
            if (uy) hby += (q0.bx - qk.bx) + (q0.az - qi.az);
            if (uz) hbz += q0.by - qi.by;
            if (uz && !ux) hbz += q0.ax - qj.ax;
        }
        if (k == sk[0] || k == sk[1] || k == sk[2]) {
            // source edge on this cell: +terms; at (i+1, j) or (i, j+1): -terms
            for (int q = 0; q < 3; ++q)
                while (k == sk[q]) {
                    const float2 d = src_terms<CM>(a, b, euni, q == 1 ? i + 1 : i, q == 2 ? j + 1 : j, k,
                                                   a.events[sev[q]].id);
                    if (q == 0) {
                        if (uy) hby += d.x;
                        if (ux) hbx += d.y;
                    } else if (q == 1) {
                        if (uy) hby -= d.x;
                    } else {
                        if (ux) hbx -= d.y;
                    }
                    sev[q] = src_from(a, sev[q] + 1, scol[q]);
                    sk[q] = sev[q] >= 0 ? a.events[sev[q]].k : INT_MAX;
                }
        }
        if (ux) b.hb[0][c] = hbx;
        if (uy) b.hb[1][c] = hby;
        if (uz) b.hb[2][c] = hbz;
        // the transposed H quantities of this cell, for the next E sweep's gather
        st_hq(b, c, hq_of<V>(a, hx, hy, hz, hbx, hby, hbz, i, j, k, psih_at<V>(a, b, i, j, k)));
        q0 = qk;
    }
}

// The adjoint sweeps of wide grids: the transposed quantities are computed
// where they are gathered instead of stored.  A thread's gather needs those
// of its cell and of two neighbours in the same plane ((i-1, j) and (i, j-1)
// for the E sweep, (i+1, j) and (i, j+1) for the H sweep); the plane
// neighbour (k-1 / k+1) is carried in registers along the z-march.  Each
// thread computes its own cell's quantities once per plane and shares them
// through LDS.  A wave covers 64 columns starting on a 256-byte boundary, so
// each row of a plane is two whole 128-byte lines; the tile's edge lane (lane
// 0 of the E sweep, lane 63 of the H sweep) computes the quantities of the
// column beyond the tile itself, as the tile's edge row computes its missing
// row.  (Tiles of 63 columns plus a halo lane that only computes put every
// row across three lines: adjoint768 reverse step 6.00 ms, against 5.1 ms
// for aligned 64-column tiles without the edge-lane work.)  The shared
// values are exactly those the neighbour computes for itself, and exactly
// the stored ones of the narrow-grid sweeps above.

// The wide-grid sweeps march like the forward ones: every array of a plane
// is addressed by one 32-bit byte offset from its allocation base (the
// adjoint arrays are under 4 GiB, checked before the run), and the loads of
// plane k+1 -- adjoint fields, the neighbour row's values the tile's edge
// row needs, the class index of the plane after -- are issued before plane k
// is computed, in two register sets that alternate (loop unrolled by two),
// with the coefficient-table entries requested one plane ahead.  The
// arithmetic is that of the per-cell helpers above, in the same order.
typedef std::uint32_t u32;

__device__ __forceinline__ float ldf(const float *base, u32 off) { return *(const float *)((const char *)base + off); }
__device__ __forceinline__ void stf(float *base, u32 off, float v) { *(float *)((char *)base + off) = v; }

struct Tab3 {
    float4 x, y, z;
};

// The uniform (one-class) coefficients, read once per workgroup into scalar
// registers and hidden from the optimizer: told that they equal the table's
// first entry, it re-reads the table through a selected pointer on every
// plane instead of using the registers.
__device__ __forceinline__ float4 sreg4(const float4 *p)
{
    typedef __attribute__((address_space(4))) const float cf;
    const cf *q = (const cf *)p;
    float4 v = make_float4(q[0], q[1], q[2], q[3]);
    v.x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.x)));
    v.y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.y)));
    v.z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.z)));
    v.w = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v.w)));
    asm volatile("" : "+s"(v.x), "+s"(v.y), "+s"(v.z), "+s"(v.w));
    return v;
}
__device__ __forceinline__ Tab3 uniform_tab3(const float4 *tab) { return Tab3{sreg4(tab), sreg4(tab + 1), sreg4(tab + 2)}; }
__device__ __forceinline__ Tab3 tab3(const float4 *tab, int cl)
{
    return Tab3{tab[cl * 3 + 0], tab[cl * 3 + 1], tab[cl * 3 + 2]};
}
template <int CM>
__device__ __forceinline__ int class_u(const void *cls, u32 cell)
{
    if (CM == CLS_U16) return *(const std::uint16_t *)((const char *)cls + cell * 2u);
    if (CM == CLS_U32) return (int)*(const std::uint32_t *)((const char *)cls + cell * 4u);
    return 0;
}

// Reads of one plane k of the adjoint E sweep.
struct AEPlane {
    float e0, e1, e2;   // dJ/dE of the own cell (i, j, k)
    float h0, h1, h2;   // dJ/dH of (i, j, k)
    float g0, g1, g2;   // dJ/dH of (i, j-1, k) (the tile's first row)
    float m0, m1, m2;   // dJ/dH of (i-1, j, k) (the tile's first lane)
    int cln;            // class index of (i, j, k+1) (E coefficients)
    Psi6 ph, phm, phg;  // psi of H cells (i, j, k), (i-1, j, k), (i, j-1, k)
    Psi6 pe;            // psi of E cell (i, j, k), before this sweep updates it
};

// Column constants of the adjoint E sweep.
struct AECol {
    int i, j, row, lane, mycol;
    bool valid, own, ilo, ihi, jlo, jhi, in_box, xs, ys;
    float4 pxe, pye;
    u32 o0, ob, cell0;   // byte offset of (i, j, 0) from the allocation base; bytes per plane; cell index of (i, j, 0)
};

// Interior tile of a workgroup from a list of tiles of the box (each its
// index in the box, x fastest), XCD-aware as remap_tile.
__device__ __forceinline__ void tile_of_list(const KArgs &a, const int *tl, int ntl, int &tx, int &ty, int &tz)
{
    const int e = tl[remap_tile(ntl)];
    const int bx = a.ib[3] - a.ib[0], by = a.ib[4] - a.ib[1];
    tx = a.ib[0] + e % bx;
    ty = a.ib[1] + (e / bx) % by;
    tz = a.ib[2] + e / (bx * by);
}

// PIPE (the two-deep pipeline of the interior variants and of the E sweep's
// x-edge and z-slab ones): every load is issued on every path (lanes and
// planes that need no value read a cell that exists: the first cell of the
// arrays, their own cell, or the last plane), so that the compiler's waits
// for a plane's values count only the younger loads of the next plane.
// Otherwise (variants that march one plane at a time, where the pipeline's
// registers cost more occupancy than it hides) only the needed loads are
// issued.
template <int CM, int V, bool PIPE>
__device__ __forceinline__ void ae_fetch(AEPlane &p, const KArgs &a, const AArgs &b, const AECol &cc,
                                         const float *const (&eb)[3], const float *const (&hb)[3], int euni, int k,
                                         int k1)
{
    constexpr bool IN = V == V_IN;
    const u32 o = cc.o0 + (u32)k * cc.ob;
    // (i, j-1): the tile's first row, and every row of an interior tile
    // (which takes its y-neighbour from memory, not from LDS)
    const bool ne = IN || cc.own, nh = IN || cc.valid, ng = IN || (cc.row == 0 && cc.own && cc.jlo);
    const bool nm = cc.lane == 0 && (IN || (cc.own && cc.ilo));
    if (PIPE) {
        const u32 oe = ne ? o : cc.ob;   // cc.ob: cell (0, 0, 0)
        const u32 oh = nh ? o : cc.ob;
        // lanes and rows that need no neighbour value read one address of
        // the wave (one line) instead of their own cells again
        const u32 ou = (u32)__builtin_amdgcn_readfirstlane((int)oh);
        const u32 og = ng ? o - (u32)a.px * 4u : ou;
        const u32 om = (u32)__builtin_amdgcn_readfirstlane((int)(nm ? o - 4u : oh));
        p.e0 = ldf(eb[0], oe);
        p.e1 = ldf(eb[1], oe);
        p.e2 = ldf(eb[2], oe);
        p.h0 = ldf(hb[0], oh);
        p.h1 = ldf(hb[1], oh);
        p.h2 = ldf(hb[2], oh);
        p.g0 = ldf(hb[0], og);
        p.g1 = IN ? 0.f : ldf(hb[1], og);   // (the y-neighbour quantities read need no Hy)
        p.g2 = ldf(hb[2], og);
        p.m0 = ldf(hb[0], om);
        p.m1 = ldf(hb[1], om);
        p.m2 = ldf(hb[2], om);
        const int kn = min(k + 1, k1 - 1);
        const int cl = class_u<CM>(a.cls, (ne ? cc.cell0 : 0u) + (u32)kn * (u32)a.plane);
        p.cln = euni ? 0 : cl;
    } else {
        p.e0 = p.e1 = p.e2 = p.h0 = p.h1 = p.h2 = p.g0 = p.g1 = p.g2 = p.m0 = p.m1 = p.m2 = 0.f;
        if (ne) {
            p.e0 = ldf(eb[0], o);
            p.e1 = ldf(eb[1], o);
            p.e2 = ldf(eb[2], o);
        }
        if (nh) {
            p.h0 = ldf(hb[0], o);
            p.h1 = ldf(hb[1], o);
            p.h2 = ldf(hb[2], o);
        }
        if (ng) {
            const u32 og = o - (u32)a.px * 4u;
            p.g0 = ldf(hb[0], og);
            p.g1 = ldf(hb[1], og);
            p.g2 = ldf(hb[2], og);
        }
        if (nm) {
            p.m0 = ldf(hb[0], o - 4u);
            p.m1 = ldf(hb[1], o - 4u);
            p.m2 = ldf(hb[2], o - 4u);
        }
        p.cln = (!euni && ne && k + 1 < k1) ? class_u<CM>(a.cls, cc.cell0 + (u32)(k + 1) * (u32)a.plane) : 0;
    }
    // slab psi of the plane, read with its other values
    if (V != V_IN) {
        const int i = cc.i, j = cc.j;
        p.ph = psih_get<V, PIPE>(a, b, cc.valid, i, j, k);
        p.phm = psih_get<V, PIPE>(a, b, nm, i - 1, j, k);
        p.phg = psih_get<V, PIPE>(a, b, ng, i, j - 1, k);
        p.pe = psie_get<V, PIPE>(a, b, cc.own, i, j, k);
    }
}

// H coefficients of cell (i, j, k): the uniform ones loaded before the
// march, else through the class map.
template <int CM>
__device__ __forceinline__ void hcoefs(const KArgs &a, const Tab3 &hu, int huni, int i, int j, int k, float4 &hx,
                                       float4 &hy, float4 &hz)
{
    if (huni) {
        hx = hu.x;
        hy = hu.y;
        hz = hu.z;
    } else {
        coefs<CM>(a, a.htab, false, (i64)k * a.plane + (i64)j * a.px + i, hx, hy, hz);
        // keep the two paths apart (merged, they become one load through a
        // selected pointer into a stack copy of hu)
        asm volatile("" : "+v"(hx.x), "+v"(hx.y), "+v"(hx.z), "+v"(hy.x), "+v"(hy.y), "+v"(hy.z), "+v"(hz.x),
                     "+v"(hz.y), "+v"(hz.z));
    }
}

template <int CM, int V>
__device__ __forceinline__ void ae_update(const KArgs &a, const AArgs &b, const AECol &cc, const AEPlane &p,
                                          const Tab3 &te, const Tab3 &hu, float *const (&eb)[3], HQ (*sq)[TX],
                                          int k, int m, int fstep, int sten, int huni, HQ &qkm, int &ev, int &ek)
{
    constexpr bool IN = V == V_IN;
    constexpr bool PZ = (V & V_Z) != 0;
    const int i = cc.i, j = cc.j, row = cc.row, lane = cc.lane, mycol = cc.mycol;
    const int nz = a.nz, ny = a.ny, px = a.px;
    const u32 o = cc.o0 + (u32)k * cc.ob;
    const bool klo = !PZ || k > 0, khi = !PZ || k < nz - 1;
    const bool own = cc.own, ilo = cc.ilo, ihi = cc.ihi, jlo = cc.jlo, jhi = cc.jhi;
    const bool ux = own && ihi && jlo && jhi && klo && khi;
    const bool uy = own && ilo && ihi && jhi && klo && khi;
    const bool uz = own && ilo && ihi && jlo && jhi && khi;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    const HQ zq{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float ebx = own ? p.e0 : 0.f, eby = own ? p.e1 : 0.f, ebz = own ? p.e2 : 0.f;
    // source edge and probe events of this cell
    float cez_src = 0.f;
    bool is_src = false;
    int pev = -1;
    if (k == ek) {
        pev = ev;
        while (a.events[ev].k == k && a.events[ev].col == mycol) {
            if (a.events[ev].kind == EV_SOURCE) {
                is_src = true;
                cez_src = a.srcs[a.events[ev].id].ceze;
            }
            ++ev;
        }
        ek = (a.events[ev].col == mycol) ? a.events[ev].k : INT_MAX;
    }
    if (fstep >= 0 && (ux || uy || uz)) {
        const float4 ex = te.x, ey = te.y, ez = te.z;
        const bool zs = PZ && zs_e(a, k);
        // psi_e adjoint: the old values read with the plane (p.pe)
        if (cc.xs && jhi && khi) {
            const i64 s = psi_sx(a, i, j, k);
            b.pb[0][s] = cc.pxe.x * (p.pe.x0 + (uy ? (ey.z * dx) * eby : 0.f));
            b.pb[1][s] = cc.pxe.x * (p.pe.x1 + (uz ? (ez.y * dx) * ebz : 0.f));
        }
        if (cc.ys && ihi && khi) {
            const i64 s = psi_sy(a, i, j, k);
            b.pb[4][s] = cc.pye.x * (p.pe.y0 + (ux ? (ex.y * dy) * ebx : 0.f));
            b.pb[5][s] = cc.pye.x * (p.pe.y1 + (uz ? (ez.z * dy) * ebz : 0.f));
        }
        if (zs && ihi && jhi) {
            const float4 pze = a.pz_prof[k];
            const i64 s = psi_sz(a, i, j, k);
            b.pb[8][s] = pze.x * (p.pe.z0 + (ux ? (ex.z * dz) * ebx : 0.f));
            b.pb[9][s] = pze.x * (p.pe.z1 + (uy ? (ey.y * dz) * eby : 0.f));
        }
        if (cc.in_box && k >= b.sb[2] && k < b.sb[2] + b.sb[5]) {
            const i64 nb = (i64)b.sb[3] * b.sb[4] * b.sb[5];
            const i64 q = (i - b.sb[0]) + (i64)b.sb[3] * ((j - b.sb[1]) + (i64)b.sb[4] * (k - b.sb[2]));
            const float *h0 = b.hist + (i64)(fstep - b.hist_base) * 3 * nb;
            const float *h1 = h0 + 3 * nb;
            const float e_[3] = {ebx, eby, ebz};
            const bool u_[3] = {ux, uy, uz};
            for (int c3 = 0; c3 < 3; ++c3) {
                const float4 dc = b.dcoef[c3 * nb + q];
                if (!u_[c3] || (dc.x == 0.f && dc.y == 0.f)) continue;
                const double em = h0[c3 * nb + q], ep = h1[c3 * nb + q];
                const double dE = (double)dc.x * em - (double)dc.y * (ep - (double)dc.z * em);
                b.gedge[c3 * nb + q] += (double)e_[c3] * dE;
            }
        }
        if (ux) ebx = ex.x * ebx;
        if (uy) eby = ey.x * eby;
        if (uz) ebz = (is_src ? cez_src : ez.x) * ebz;
    }
    if (sten) {
        // consecutive planes use the two LDS buffers in turn (sq is one of
        // them): a thread writes plane k's after the barrier of plane k-1,
        // which every thread reaches only once done reading plane k-2's, so
        // one barrier per plane
        HQ q0 = zq;
        if (IN || cc.valid) {
            float4 hx, hy, hz;
            hcoefs<CM>(a, hu, huni, i, j, k, hx, hy, hz);
            q0 = hq_of<V>(a, hx, hy, hz, p.h0, p.h1, p.h2, i, j, k, p.ph);
        }
        // the tile's first lane and first row: the quantities of (i-1, j, k)
        // and (i, j-1, k), beyond the tile, computed beside the cell's own so
        // that their reads overlap
        HQ qim = zq, qjm = zq;
        if (own && lane == 0 && ilo && (uy || uz)) {
            float4 hx, hy, hz;
            hcoefs<CM>(a, hu, huni, i - 1, j, k, hx, hy, hz);
            qim = hq_of<V>(a, hx, hy, hz, p.m0, p.m1, p.m2, i - 1, j, k, p.phm);
        }
        if (IN || (own && row == 0 && jlo && (ux || uz))) {
            float4 hx, hy, hz;
            hcoefs<CM>(a, hu, huni, i, j - 1, k, hx, hy, hz);
            qjm = hq_of<V>(a, hx, hy, hz, p.g0, p.g1, p.g2, i, j - 1, k, p.phg);
        }
        // (i-1, j, k) and (i, j-1, k).  Interior tiles: the x-neighbour's from
        // the adjacent lane (a DPP wave shift), the y-neighbour's computed
        // from its H read above, no LDS and no barrier.  Otherwise both
        // through LDS, the tile's first lane and row taking their own.
        HQ qi = zq, qj = zq;
        if (IN) {
            qi.tz = lane_prev(q0.tz, qim.tz);
            qi.ry = lane_prev(q0.ry, qim.ry);
            qj = qjm;
        } else {
            sq[row][lane] = q0;
            __syncthreads();
            qi = sq[row][lane > 0 ? lane - 1 : 0];
            if (lane == 0) qi = qim;
            qj = sq[row > 0 ? row - 1 : 0][lane];
            if (row == 0) qj = qjm;
        }
        if (own) {
            if (ux) {
                ebx += (qkm.tx - q0.tx) + (qj.rx - q0.rx);
                if (uz) ebz += qj.sz - q0.sz;
            }
            if (uy) eby += (qkm.sy - q0.sy) + (qi.ry - q0.ry);
            if (uz) ebz += qi.tz - q0.tz;
            if (uz && !ux) ebz += qj.sz - q0.sz;
            qkm = q0;
            if (pev >= 0) {   // dJ/d(probe sample m) of E probes on this cell
                for (int e = pev; a.events[e].k == k && a.events[e].col == mycol; ++e) {
                    const int kind = a.events[e].kind;
                    if (kind < EV_PROBE || kind > EV_PROBE + 2) continue;
                    const float gv = b.g[(i64)a.events[e].id * b.gsteps + m];
                    if (kind == EV_PROBE) ebx += gv;
                    else if (kind == EV_PROBE + 1) eby += gv;
                    else ebz += gv;
                }
            }
        }
    }
    if (ux) stf(eb[0], o, ebx);
    if (uy) stf(eb[1], o, eby);
    if (uz) stf(eb[2], o, ebz);
}

// fstep >= 0: finalize the E half-step of forward step fstep first.
// sten: gather the transposed H half-step of step m and inject dJ/ds[m].
template <int CM, int V>
__global__ void __launch_bounds__(TX * BY)
    adj_e_lds_kernel(const KArgs a, const AArgs b, const int m, const int fstep, const int sten, const int euni,
                     const int huni)
{
    constexpr bool IN = V == V_IN;   // interior tiles: no slab, face or range predicates
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;   // predicated axes
    __shared__ HQ sq[2][BY][TX];
    int tx, ty, tz;
    if (IN && b.ntl > 0) tile_of_list(a, b.tl, b.ntl, tx, ty, tz);
    else tile_of_region<V>(a, tx, ty, tz);
    AECol cc;
    cc.lane = (int)(threadIdx.x & (TX - 1));
    cc.row = (int)(threadIdx.x / TX);
    const int i = tx * TX + cc.lane;
    const int j = ty * BY + cc.row;
    cc.i = i;
    cc.j = j;
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    cc.valid = (!PX || i < nx) && (!PY || j < ny);   // a cell this thread updates
    cc.own = cc.valid;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    cc.ilo = !PX || i > 0;
    cc.ihi = !PX || i < nx - 1;
    cc.jlo = !PY || j > 0;
    cc.jhi = !PY || j < ny - 1;
    cc.mycol = j * nx + i;
    cc.in_box = i >= b.sb[0] && i < b.sb[0] + b.sb[3] && j >= b.sb[1] && j < b.sb[1] + b.sb[4];
    cc.xs = PX && cc.own && xs_e(a, i);
    cc.ys = PY && cc.own && ys_e(a, j);
    cc.pxe = cc.xs ? a.px_prof[i] : NOPROF;
    cc.pye = cc.ys ? a.py_prof[j] : NOPROF;
    cc.ob = (u32)a.plane * 4u;
    cc.cell0 = (u32)(j * px + i);
    cc.o0 = cc.cell0 * 4u + cc.ob;
    // allocation bases (guard plane below k = 0): read eb and hb, write ebn
    float *const eb[3] = {b.ebn[0] - a.plane, b.ebn[1] - a.plane, b.ebn[2] - a.plane};
    const float *const ebc[3] = {b.eb[0] - a.plane, b.eb[1] - a.plane, b.eb[2] - a.plane};
    const float *const hb[3] = {b.hb[0] - a.plane, b.hb[1] - a.plane, b.hb[2] - a.plane};
    int ev = cc.own ? first_event(a, a.e_first, cc.mycol, k0) : -1;
    int ek = ev >= 0 ? a.events[ev].k : INT_MAX;

    HQ qkm{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (cc.own && sten && k0 > 0) qkm = hquant<CM, V>(a, b, huni, i, j, k0 - 1);
    const Tab3 hu = uniform_tab3(a.htab);   // the H coefficients when they are one class
    const int cl0 = (!euni && (IN || cc.own)) ? class_u<CM>(a.cls, cc.cell0 + (u32)k0 * (u32)a.plane) : 0;
    if (V == V_GEN) {
        // one plane at a time (the x-edge and z-slab variants pipeline like the
        // interior: adjoint768 x-edge 0.575 -> 0.555 ms, z-slab 0.365 -> 0.357 ms)
        int cl = cl0;
        for (int k = k0; k < k1; ++k) {
            AEPlane P;
            ae_fetch<CM, V, false>(P, a, b, cc, ebc, hb, euni, k, k1);
            const Tab3 t = tab3(a.etab, cl);
            ae_update<CM, V>(a, b, cc, P, t, hu, eb, sq[(k - k0) & 1], k, m, fstep, sten, huni, qkm, ev, ek);
            cl = P.cln;
        }
        return;
    }
    AEPlane A, B;
    Tab3 tA = tab3(a.etab, cl0), tB;
    ae_fetch<CM, V, true>(A, a, b, cc, ebc, hb, euni, k0, k1);
    int k = k0;
    // steady state: the fetches are unconditional, so on every path the next
    // plane's loads are the younger ones when this plane's are awaited
    for (; k + 2 < k1; k += 2) {
        ae_fetch<CM, V, true>(B, a, b, cc, ebc, hb, euni, k + 1, k1);
        tB = tab3(a.etab, A.cln);
        ae_update<CM, V>(a, b, cc, A, tA, hu, eb, sq[0], k, m, fstep, sten, huni, qkm, ev, ek);
        ae_fetch<CM, V, true>(A, a, b, cc, ebc, hb, euni, k + 2, k1);
        tA = tab3(a.etab, B.cln);
        ae_update<CM, V>(a, b, cc, B, tB, hu, eb, sq[1], k + 1, m, fstep, sten, huni, qkm, ev, ek);
    }
    if (k + 1 < k1) {
        ae_fetch<CM, V, true>(B, a, b, cc, ebc, hb, euni, k + 1, k1);
        tB = tab3(a.etab, A.cln);
        ae_update<CM, V>(a, b, cc, A, tA, hu, eb, sq[0], k, m, fstep, sten, huni, qkm, ev, ek);
        ae_update<CM, V>(a, b, cc, B, tB, hu, eb, sq[1], k + 1, m, fstep, sten, huni, qkm, ev, ek);
    } else {
        ae_update<CM, V>(a, b, cc, A, tA, hu, eb, sq[0], k, m, fstep, sten, huni, qkm, ev, ek);
    }
}

// Reads of one plane k of the adjoint H sweep.
struct AHPlane {
    float f0, f1, f2;   // dJ/dE of (i, j, k+1)
    float h0, h1, h2;   // dJ/dH of the own cell (i, j, k)
    float d0, d1, d2;   // dJ/dE of (i, j+1, k) (the tile's last row)
    float x0, x1, x2;   // dJ/dE of (i+1, j, k) (the tile's last lane)
    int cl2;            // class index of (i, j, k+2) (E coefficients of the next plane's z-neighbour)
    int cld;            // class index of (i, j+1, k+1) (the tile's last row, next plane)
    int cli;            // class index of (i+1, j, k+1) (the tile's last lane, next plane)
    Psi6 pf, pfx, pfd;  // psi of E cells (i, j, k+1), (i+1, j, k), (i, j+1, k)
    Psi6 ph;            // psi of H cell (i, j, k), before this sweep updates it
};

// PIPE: as for ae_fetch.
template <int CM, int V, bool PIPE>
__device__ __forceinline__ void ah_fetch(AHPlane &p, const KArgs &a, const AArgs &b, const AECol &cc,
                                         const float *const (&eb)[3], const float *const (&hb)[3], int euni, int k)
{
    constexpr bool IN = V == V_IN;
    const int nz = a.nz;
    const u32 o = cc.o0 + (u32)k * cc.ob;
    // (i, j+1): the tile's last row, and every row of an interior tile
    // (which takes its y-neighbour from memory, not from LDS)
    const bool last = IN || (cc.row == BY - 1 && cc.own && cc.jhi);
    const bool nf = (IN || cc.valid) && (!(V & V_Z) || k + 1 < nz), nh = IN || cc.own;
    const bool xl = cc.lane == TX - 1 && (IN || (cc.own && cc.ihi));   // column i+1 is in the grid
    if (PIPE) {
        const u32 of = nf ? o + cc.ob : cc.ob;   // cc.ob: cell (0, 0, 0)
        const u32 oh = nh ? o : cc.ob;
        // lanes and rows that need no neighbour value read one address of
        // the wave (one line) instead of their own cells again
        const u32 ou = (u32)__builtin_amdgcn_readfirstlane((int)oh);
        const u32 od = last ? o + (u32)a.px * 4u : ou;
        const u32 ox = (u32)__builtin_amdgcn_readlane((int)(xl ? o + 4u : oh), TX - 1);
        p.f0 = ldf(eb[0], of);
        p.f1 = ldf(eb[1], of);
        p.f2 = ldf(eb[2], of);
        p.h0 = ldf(hb[0], oh);
        p.h1 = ldf(hb[1], oh);
        p.h2 = ldf(hb[2], oh);
        p.d0 = ldf(eb[0], od);
        p.d1 = IN ? 0.f : ldf(eb[1], od);   // (the y-neighbour quantities read need no Ey)
        p.d2 = ldf(eb[2], od);
        p.x0 = ldf(eb[0], ox);
        p.x1 = ldf(eb[1], ox);
        p.x2 = ldf(eb[2], ox);
        const u32 c0 = (IN || cc.valid) ? cc.cell0 : 0u;
        const u32 cu = (u32)__builtin_amdgcn_readfirstlane((int)c0);
        const int c2 = class_u<CM>(a.cls, c0 + (u32)min(k + 2, nz - 1) * (u32)a.plane);
        const int cd =
            class_u<CM>(a.cls, (last ? cc.cell0 + (u32)a.px : cu) + (u32)min(k + 1, nz - 1) * (u32)a.plane);
        const u32 cx = (u32)__builtin_amdgcn_readlane((int)(xl ? cc.cell0 + 1u : c0), TX - 1);
        const int ci = class_u<CM>(a.cls, cx + (u32)min(k + 1, nz - 1) * (u32)a.plane);
        p.cl2 = euni ? 0 : c2;
        p.cld = euni ? 0 : cd;
        p.cli = euni ? 0 : ci;
    } else {
        p.f0 = p.f1 = p.f2 = p.h0 = p.h1 = p.h2 = p.d0 = p.d1 = p.d2 = p.x0 = p.x1 = p.x2 = 0.f;
        if (nf) {
            p.f0 = ldf(eb[0], o + cc.ob);
            p.f1 = ldf(eb[1], o + cc.ob);
            p.f2 = ldf(eb[2], o + cc.ob);
        }
        if (nh) {
            p.h0 = ldf(hb[0], o);
            p.h1 = ldf(hb[1], o);
            p.h2 = ldf(hb[2], o);
        }
        if (last) {
            const u32 od = o + (u32)a.px * 4u;
            p.d0 = ldf(eb[0], od);
            p.d1 = ldf(eb[1], od);
            p.d2 = ldf(eb[2], od);
        }
        if (xl) {
            p.x0 = ldf(eb[0], o + 4u);
            p.x1 = ldf(eb[1], o + 4u);
            p.x2 = ldf(eb[2], o + 4u);
        }
        p.cl2 = (!euni && cc.valid && k + 2 < nz) ? class_u<CM>(a.cls, cc.cell0 + (u32)(k + 2) * (u32)a.plane) : 0;
        p.cld = (!euni && last && k + 1 < nz) ? class_u<CM>(a.cls, cc.cell0 + (u32)a.px + (u32)(k + 1) * (u32)a.plane)
                                              : 0;
        p.cli = (!euni && xl && k + 1 < nz) ? class_u<CM>(a.cls, cc.cell0 + 1u + (u32)(k + 1) * (u32)a.plane) : 0;
    }
    // slab psi of the plane, read with its other values
    if (V != V_IN) {
        const int i = cc.i, j = cc.j;
        p.pf = psie_get<V, PIPE>(a, b, nf, i, j, k + 1);
        p.pfx = psie_get<V, PIPE>(a, b, xl, i + 1, j, k);
        p.pfd = psie_get<V, PIPE>(a, b, last, i, j + 1, k);
        p.ph = psih_get<V, PIPE>(a, b, nh, i, j, k);
    }
}

// tk: E coefficients of (i, j, k+1); tj: of (i, j+1, k) (the tile's last row);
// ti: of (i+1, j, k) (the tile's last lane).  q0: the transposed quantities
// of (i, j, k), carried from the previous plane.
template <int CM, int V>
__device__ __forceinline__ void ah_update(const KArgs &a, const AArgs &b, const AECol &cc, const AHPlane &p,
                                          const Tab3 &tk, const Tab3 &tj, const Tab3 &ti, const Tab3 &hu,
                                          float *const (&hb)[3], EQ (*sq)[TX], int k, int euni, int huni, EQ &q0,
                                          int (&sev)[3], int (&sk)[3], const int (&scol)[3])
{
    constexpr bool IN = V == V_IN;
    constexpr bool PZ = (V & V_Z) != 0;
    const int i = cc.i, j = cc.j, row = cc.row, lane = cc.lane;
    const int nz = a.nz, ny = a.ny, px = a.px;
    const u32 o = cc.o0 + (u32)k * cc.ob;
    const bool own = cc.own, ihi = cc.ihi, jhi = cc.jhi;
    const bool khi = !PZ || k < nz - 1;
    const bool ux = own && jhi && khi, uy = own && ihi && khi, uz = own && ihi && jhi;
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    const EQ zq{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const EQ qk = (cc.valid && khi) ? eq_of<V>(a, tk.x, tk.y, tk.z, p.f0, p.f1, p.f2, i, j, k + 1, p.pf) : zq;
    // the tile's last lane and last row: the quantities of (i+1, j, k) and
    // (i, j+1, k), beyond the tile, computed beside the cell's own so that
    // their reads overlap
    EQ qil = zq, qjl = zq;
    if (own && lane == TX - 1 && ihi && (uy || uz))
        qil = eq_of<V>(a, ti.x, ti.y, ti.z, p.x0, p.x1, p.x2, i + 1, j, k, p.pfx);
    if (IN || (own && row == BY - 1 && jhi && (ux || uz)))
        qjl = eq_of<V>(a, tj.x, tj.y, tj.z, p.d0, p.d1, p.d2, i, j + 1, k, p.pfd);
    float hbx = 0.f, hby = 0.f, hbz = 0.f;
    if (own) {
        float4 hx, hy, hz;
        hcoefs<CM>(a, hu, huni, i, j, k, hx, hy, hz);
        hbx = p.h0;
        hby = p.h1;
        hbz = p.h2;
        // psi_h adjoint and dJ/dH <- Ch dJ/dH (the H half-step, pointwise part)
        // psi_h adjoint: the old values read with the plane (p.ph)
        if (cc.xs) {
            const i64 s = psi_sx(a, i, j, k);
            b.pb[2][s] = cc.pxe.x * (p.ph.x0 + (uy ? (hy.y * dx) * hby : 0.f));
            b.pb[3][s] = cc.pxe.x * (p.ph.x1 + (uz ? (hz.z * dx) * hbz : 0.f));
        }
        if (cc.ys) {
            const i64 s = psi_sy(a, i, j, k);
            b.pb[6][s] = cc.pye.x * (p.ph.y0 + (ux ? (hx.z * dy) * hbx : 0.f));
            b.pb[7][s] = cc.pye.x * (p.ph.y1 + (uz ? (hz.y * dy) * hbz : 0.f));
        }
        if (PZ && zs_h(a, k)) {
            const float4 pzm = a.pz_prof[nz + k];
            const i64 s = psi_sz(a, i, j, k);
            b.pb[10][s] = pzm.x * (p.ph.z0 + (ux ? (hx.y * dz) * hbx : 0.f));
            b.pb[11][s] = pzm.x * (p.ph.z1 + (uy ? (hy.z * dz) * hby : 0.f));
        }
        if (ux) hbx = hx.x * hbx;
        if (uy) hby = hy.x * hby;
        if (uz) hbz = hz.x * hbz;
    }
    // transposed E half-step gathered from the E cells that read this H:
    // (i+1, j, k) and (i, j+1, k).  Interior tiles: the x-neighbour's from
    // the adjacent lane (a DPP wave shift), the y-neighbour's computed from
    // its E read with the plane, no LDS and no barrier.  Otherwise from the
    // neighbours in LDS, the tile's last lane and last row computing theirs;
    // consecutive planes use the two LDS buffers in turn, one barrier per
    // plane (as in the E sweep).
    EQ qi = zq, qj = zq;
    if (IN) {
        qi.by = lane_next(q0.by, qil.by);
        qi.az = lane_next(q0.az, qil.az);
        qj = qjl;
    } else {
        sq[row][lane] = q0;
        __syncthreads();
        qi = sq[row][lane < TX - 1 ? lane + 1 : TX - 1];
        if (lane == TX - 1) qi = qil;
        qj = sq[row < BY - 1 ? row + 1 : BY - 1][lane];
        if (row == BY - 1) qj = qjl;
    }
    if (own) {
        if (ux) {
            hbx += (q0.ay - qk.ay) + (q0.bz - qj.bz);
            if (uz) hbz += q0.ax - qj.ax;
        }
        if (uy) hby += (q0.bx - qk.bx) + (q0.az - qi.az);
        if (uz) hbz += q0.by - qi.by;
        if (uz && !ux) hbz += q0.ax - qj.ax;
    }
    if (k == sk[0] || k == sk[1] || k == sk[2]) {
        // source edge on this cell: +terms; at (i+1, j) or (i, j+1): -terms
        for (int q = 0; q < 3; ++q)
            while (k == sk[q]) {
                const float2 d = src_terms<CM>(a, b, euni, q == 1 ? i + 1 : i, q == 2 ? j + 1 : j, k,
                                               a.events[sev[q]].id);
                if (q == 0) {
                    if (uy) hby += d.x;
                    if (ux) hbx += d.y;
                } else if (q == 1) {
                    if (uy) hby -= d.x;
                } else {
                    if (ux) hbx -= d.y;
                }
                sev[q] = src_from(a, sev[q] + 1, scol[q]);
                sk[q] = sev[q] >= 0 ? a.events[sev[q]].k : INT_MAX;
            }
    }
    if (ux) stf(hb[0], o, hbx);
    if (uy) stf(hb[1], o, hby);
    if (uz) stf(hb[2], o, hbz);
    q0 = qk;
}

template <int CM, int V>
__global__ void __launch_bounds__(TX * BY)
    adj_h_lds_kernel(const KArgs a, const AArgs b, const int euni, const int huni)
{
    constexpr bool IN = V == V_IN;   // interior tiles: no slab, face or range predicates
    constexpr bool PX = (V & V_X) != 0, PY = (V & V_Y) != 0, PZ = (V & V_Z) != 0;   // predicated axes
    __shared__ EQ sq[2][BY][TX];
    int tx, ty, tz;
    if (IN && b.ntl > 0) tile_of_list(a, b.tl, b.ntl, tx, ty, tz);
    else tile_of_region<V>(a, tx, ty, tz);
    AECol cc;
    cc.lane = (int)(threadIdx.x & (TX - 1));
    cc.row = (int)(threadIdx.x / TX);
    const int i = tx * TX + cc.lane;
    const int j = ty * BY + cc.row;
    cc.i = i;
    cc.j = j;
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    cc.valid = (!PX || i < nx) && (!PY || j < ny);
    cc.own = cc.valid;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    cc.ilo = true;
    cc.jlo = true;
    cc.ihi = !PX || i < nx - 1;
    cc.jhi = !PY || j < ny - 1;
    cc.mycol = j * nx + i;
    cc.in_box = false;
    cc.xs = PX && cc.own && xs_h(a, i);
    cc.ys = PY && cc.own && ys_h(a, j);
    cc.pxe = cc.xs ? a.px_prof[nx + i] : NOPROF;
    cc.pye = cc.ys ? a.py_prof[ny + j] : NOPROF;
    cc.ob = (u32)a.plane * 4u;
    cc.cell0 = (u32)(j * px + i);
    cc.o0 = cc.cell0 * 4u + cc.ob;
    // read ebn (this step's dJ/dE) and hb, write hbn
    const float *const eb[3] = {b.ebn[0] - a.plane, b.ebn[1] - a.plane, b.ebn[2] - a.plane};
    float *const hb[3] = {b.hbn[0] - a.plane, b.hbn[1] - a.plane, b.hbn[2] - a.plane};
    const float *const hbc[3] = {b.hb[0] - a.plane, b.hb[1] - a.plane, b.hb[2] - a.plane};
    const EQ zq{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    EQ q0 = cc.valid ? equant<CM, V>(a, b, euni, i, j, k0) : zq;
    // lumped-source edges among the Ez edges this thread gathers from: its
    // own cell, (i+1, j) and (i, j+1); event cursor and its plane per column.
    // Rows outside the source rows skip the lookups (no load: a kernel
    // argument range).
    const bool srow = cc.own && j >= b.src_j0 && j <= b.src_j1;
    const int scol[3] = {j * nx + i, j * nx + i + 1, (j + 1) * nx + i};
    const bool sok[3] = {srow, srow && i + 1 < nx, srow && j + 1 < ny};
    int sev[3], sk[3];
    for (int q = 0; q < 3; ++q) {
        sev[q] = sok[q] ? src_from(a, first_event(a, a.e_first, scol[q], k0), scol[q]) : -1;
        sk[q] = sev[q] >= 0 ? a.events[sev[q]].k : INT_MAX;
    }
    const Tab3 hu = uniform_tab3(a.htab);   // the H coefficients when they are one class
    // E coefficients of (i, j, k0+1), of (i, j+1, k0) and of (i+1, j, k0)
    int ck, cj, ci;
    {
        const bool last = IN || (cc.row == BY - 1 && cc.own && cc.jhi);
        const bool xl = cc.lane == TX - 1 && (IN || (cc.own && cc.ihi));
        const u32 c0 = (IN || cc.valid) ? cc.cell0 : 0u;
        ck = euni ? 0 : class_u<CM>(a.cls, c0 + (u32)min(k0 + 1, nz - 1) * (u32)a.plane);
        cj = euni ? 0 : class_u<CM>(a.cls, (last ? cc.cell0 + (u32)px : c0) + (u32)k0 * (u32)a.plane);
        ci = euni ? 0 : class_u<CM>(a.cls, (xl ? cc.cell0 + 1u : c0) + (u32)k0 * (u32)a.plane);
    }
    if (!IN) {
        // one plane at a time (pipelined, the H edge variants measured
        // slower: x-edge 0.57 -> 0.69 ms, z-slab 0.39 -> 0.48 ms on adjoint768)
        for (int k = k0; k < k1; ++k) {
            AHPlane P;
            ah_fetch<CM, V, false>(P, a, b, cc, eb, hbc, euni, k);
            const Tab3 tk = tab3(a.etab, ck), tj = tab3(a.etab, cj), ti = tab3(a.etab, ci);
            ah_update<CM, V>(a, b, cc, P, tk, tj, ti, hu, hb, sq[(k - k0) & 1], k, euni, huni, q0, sev, sk, scol);
            ck = P.cl2;
            cj = P.cld;
            ci = P.cli;
        }
        return;
    }
    AHPlane A, B;
    // E coefficients of (i, j, k+1), of (i, j+1, k) and of (i+1, j, k) for
    // plane k of the set
    Tab3 tkA = tab3(a.etab, ck), tjA = tab3(a.etab, cj), tiA = tab3(a.etab, ci), tkB, tjB, tiB;
    ah_fetch<CM, V, true>(A, a, b, cc, eb, hbc, euni, k0);
    int k = k0;
#define AH_UPD(S, TK, TJ, TI, BUF, K) \
    ah_update<CM, V>(a, b, cc, S, TK, TJ, TI, hu, hb, sq[BUF], K, euni, huni, q0, sev, sk, scol)
    for (; k + 2 < k1; k += 2) {
        ah_fetch<CM, V, true>(B, a, b, cc, eb, hbc, euni, k + 1);
        tkB = tab3(a.etab, A.cl2);
        tjB = tab3(a.etab, A.cld);
        tiB = tab3(a.etab, A.cli);
        AH_UPD(A, tkA, tjA, tiA, 0, k);
        ah_fetch<CM, V, true>(A, a, b, cc, eb, hbc, euni, k + 2);
        tkA = tab3(a.etab, B.cl2);
        tjA = tab3(a.etab, B.cld);
        tiA = tab3(a.etab, B.cli);
        AH_UPD(B, tkB, tjB, tiB, 1, k + 1);
    }
    if (k + 1 < k1) {
        ah_fetch<CM, V, true>(B, a, b, cc, eb, hbc, euni, k + 1);
        tkB = tab3(a.etab, A.cl2);
        tjB = tab3(a.etab, A.cld);
        tiB = tab3(a.etab, A.cli);
        AH_UPD(A, tkA, tjA, tiA, 0, k);
        AH_UPD(B, tkB, tjB, tiB, 1, k + 1);
    } else {
        AH_UPD(A, tkA, tjA, tiA, 0, k);
    }
#undef AH_UPD
}

// E of the design box into snapshot slot (step - hist_base), SI units.
template <class T>
__global__ void snap_kernel(const KArgs a, float *hist, const int slot)
{
    const int bx = a.sb[3], by = a.sb[4], bz = a.sb[5];
    const i64 nb = (i64)bx * by * bz;
    const i64 t = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int i = (int)(t % bx), j = (int)((t / bx) % by), k = (int)(t / ((i64)bx * by));
    const i64 c = (i64)(a.sb[2] + k) * a.plane + (i64)(a.sb[1] + j) * a.px + (a.sb[0] + i);
    float *o = hist + (i64)slot * 3 * nb + t;
    o[0] = ld((const T *)a.ex + c) * a.probe_unscale[0];
    o[nb] = ld((const T *)a.ey + c) * a.probe_unscale[1];
    o[2 * nb] = ld((const T *)a.ez + c) * a.probe_unscale[2];
}

// Fused interior reverse step: the E sweep and the H sweep of one reverse
// step in one march over an interior tile whose cells hold no design edge
// and whose cells and the cells one beyond them at i+1 and j+1 hold no
// source or probe (the host's list).  Plane kk's E update runs one plane
// ahead of plane kk-1's H update, which reads dJ/dE of (i, j, kk).  The E
// values a tile's H cells read beyond the tile -- (i+1, j) for the last
// lane, (i, j+1) for every row (the row above is another wave's, or another
// tile's) -- are evaluated here as well, from the old values: the pass reads
// eb / hb and writes ebn / hbn, the other buffers of a pair, so no workgroup
// reads a value that another has already overwritten.  A cell's arithmetic
// is the E and H sweeps' (the same products and sums in the same order), so
// the result is theirs bit for bit, at about two thirds of their memory
// traffic: each cell's dJ/dE and dJ/dH are read once and written once per
// step.  H coefficients are one class (the host fuses only then).

// Reads of plane kk.  Every load is issued on every lane: a lane that needs
// no neighbour value reads the one address of its wave's lane that does.
struct FPlane {
    float e0, e1, e2;   // old dJ/dE of (i, j, kk)
    float h0, h1, h2;   // old dJ/dH of (i, j, kk)
    float g0, g2;       // old dJ/dH x, z of (i, j-1, kk)
    float m1, m2;       // old dJ/dH y, z of (i-1, j, kk) (the first lane's)
    float y0, y2;       // old dJ/dE x, z of (i, j+1, kk)
    float u0, u1, u2;   // old dJ/dH of (i, j+1, kk)
    float c1;           // old dJ/dH y of (i-1, j+1, kk) (the first lane's)
    float x1, x2;       // old dJ/dE y, z of (i+1, j, kk) (the last lane's)
    float w0, w1, w2;   // old dJ/dH of (i+1, j, kk) (the last lane's)
    float d0;           // old dJ/dH x of (i+1, j-1, kk) (the last lane's)
    int cl, cly, clx;   // E classes of (i, j, kk), (i, j+1, kk), (i+1, j, kk)
};

// Values carried from plane kk-1 to plane kk.
struct FCarry {
    float qtx, qsy;            // H quantities tx, sy of (i, j, kk-1)
    float ytx;                 // tx of (i, j+1, kk-1)
    float xsy;                 // sy of (i+1, j, kk-1)
    EQ q;                      // E quantities of (i, j, kk-1)
    float ybz, yax;            // bz, ax of (i, j+1, kk-1)
    float xaz, xby;            // az, by of (i+1, j, kk-1)
    float hx, hy, hz;          // old dJ/dH of (i, j, kk-1)
};

template <int CM>
__device__ __forceinline__ void fz_fetch(FPlane &p, const KArgs &a, const float *const (&eo)[3],
                                         const float *const (&ho)[3], u32 o0, u32 cell0, u32 ob, int kk, int euni)
{
    const u32 o = o0 + (u32)kk * ob;
    const u32 rw = (u32)a.px * 4u;
    const u32 om = (u32)__builtin_amdgcn_readfirstlane((int)(o - 4u));   // (i-1, j) of the first lane
    const u32 ox = (u32)__builtin_amdgcn_readlane((int)(o + 4u), TX - 1);   // (i+1, j) of the last lane
    p.e0 = ldf(eo[0], o);
    p.e1 = ldf(eo[1], o);
    p.e2 = ldf(eo[2], o);
    p.h0 = ldf(ho[0], o);
    p.h1 = ldf(ho[1], o);
    p.h2 = ldf(ho[2], o);
    p.g0 = ldf(ho[0], o - rw);
    p.g2 = ldf(ho[2], o - rw);
    p.m1 = ldf(ho[1], om);
    p.m2 = ldf(ho[2], om);
    p.y0 = ldf(eo[0], o + rw);
    p.y2 = ldf(eo[2], o + rw);
    p.u0 = ldf(ho[0], o + rw);
    p.u1 = ldf(ho[1], o + rw);
    p.u2 = ldf(ho[2], o + rw);
    p.c1 = ldf(ho[1], om + rw);
    p.x1 = ldf(eo[1], ox);
    p.x2 = ldf(eo[2], ox);
    p.w0 = ldf(ho[0], ox);
    p.w1 = ldf(ho[1], ox);
    p.w2 = ldf(ho[2], ox);
    p.d0 = ldf(ho[0], ox - rw);
    const u32 c = cell0 + (u32)kk * (u32)a.plane;
    const int cl = class_u<CM>(a.cls, c);
    const int cly = class_u<CM>(a.cls, c + (u32)a.px);
    const int clx = class_u<CM>(a.cls, (u32)__builtin_amdgcn_readlane((int)(c + 1u), TX - 1));
    p.cl = euni ? 0 : cl;
    p.cly = euni ? 0 : cly;
    p.clx = euni ? 0 : clx;
}

// The E update of plane kk (the own cell; x and z of (i, j+1); y and z of
// (i+1, j)), the E quantities of the three, and the H update of plane kk-1.
// fin: finalize the forward E half-step (ae_update's fstep >= 0).  e_st:
// store the own cell's dJ/dE (plane kk is the tile's); h_st: plane kk-1 is
// the tile's (not the plane below its first).
__device__ __forceinline__ void fz_step(const FPlane &p, const Tab3 &hu, const float4 *etab, bool fin, FCarry &c,
                                        float *const (&en)[3], float *const (&hn)[3], u32 o, u32 ob, bool e_st,
                                        bool h_st)
{
    const float hxy = hu.x.y, hxz = hu.x.z, hyy = hu.y.y, hyz = hu.y.z, hzy = hu.z.y, hzz = hu.z.z;
    // H quantities of the own cell, of (i, j-1) and (i-1, j) (the adjacent
    // lane's; the first lane's from its load)
    const float q0sy = hxy * p.h0, q0sz = hxz * p.h0, q0tz = hyy * p.h1, q0tx = hyz * p.h1;
    const float q0rx = hzy * p.h2, q0ry = hzz * p.h2;
    const float qjrx = hzy * p.g2, qjsz = hxz * p.g0;
    const float qitz = lane_prev(q0tz, hyy * p.m1), qiry = lane_prev(q0ry, hzz * p.m2);
    const Tab3 te = tab3(etab, p.cl);
    float ebx = p.e0, eby = p.e1, ebz = p.e2;
    if (fin) {
        ebx = te.x.x * ebx;
        eby = te.y.x * eby;
        ebz = te.z.x * ebz;
    }
    ebx += (c.qtx - q0tx) + (qjrx - q0rx);
    ebz += qjsz - q0sz;
    eby += (c.qsy - q0sy) + (qiry - q0ry);
    ebz += qitz - q0tz;
    c.qtx = q0tx;
    c.qsy = q0sy;
    if (e_st) {
        stf(en[0], o, ebx);
        stf(en[1], o, eby);
        stf(en[2], o, ebz);
    }
    const EQ q{te.x.y * ebx, te.x.z * ebx, te.y.y * eby, te.y.z * eby, te.z.y * ebz, te.z.z * ebz};
    // (i, j+1): its H quantities, (i-1, j+1)'s tz from the adjacent lane
    const float ytx = hyz * p.u1, yrx = hzy * p.u2, ysz = hxz * p.u0, ytz = hyy * p.u1;
    const float yitz = lane_prev(ytz, hyy * p.c1);
    const Tab3 ty = tab3(etab, p.cly);
    float ybx = p.y0, ybz = p.y2;
    if (fin) {
        ybx = ty.x.x * ybx;
        ybz = ty.z.x * ybz;
    }
    ybx += (c.ytx - ytx) + (q0rx - yrx);
    ybz += q0sz - ysz;
    ybz += yitz - ytz;
    c.ytx = ytx;
    // (i+1, j) (the last lane's): its H quantities, (i+1, j-1)'s sz
    const float xsy = hxy * p.w0, xsz = hxz * p.w0, xtz = hyy * p.w1, xry = hzz * p.w2;
    const float xjsz = hxz * p.d0;
    const Tab3 tx = tab3(etab, p.clx);
    float xby = p.x1, xbz = p.x2;
    if (fin) {
        xby = tx.y.x * xby;
        xbz = tx.z.x * xbz;
    }
    xbz += xjsz - xsz;
    xby += (c.xsy - xsy) + (q0ry - xry);
    xbz += q0tz - xtz;
    c.xsy = xsy;
    // H update of plane kk-1: (i+1, j)'s quantities from the adjacent lane,
    // the last lane's evaluated above
    if (h_st) {
        float hbx = hu.x.x * c.hx, hby = hu.y.x * c.hy, hbz = hu.z.x * c.hz;
        const float qiaz = lane_next(c.q.az, c.xaz), qiby = lane_next(c.q.by, c.xby);
        hbx += (c.q.ay - q.ay) + (c.q.bz - c.ybz);
        hbz += c.q.ax - c.yax;
        hby += (c.q.bx - q.bx) + (c.q.az - qiaz);
        hbz += c.q.by - qiby;
        stf(hn[0], o - ob, hbx);
        stf(hn[1], o - ob, hby);
        stf(hn[2], o - ob, hbz);
    }
    c.q = q;
    c.ybz = ty.z.z * ybz;
    c.yax = ty.x.y * ybx;
    c.xaz = tx.z.y * xbz;
    c.xby = tx.y.z * xby;
    c.hx = p.h0;
    c.hy = p.h1;
    c.hz = p.h2;
}

template <int CM>
__global__ void __launch_bounds__(TX * BY) adj_fused_kernel(const KArgs a, const AArgs b, const int fin, const int euni)
{
    int tx, ty, tz;
    tile_of_list(a, b.fl, b.nfl, tx, ty, tz);
    const int lane = (int)(threadIdx.x & (TX - 1)), row = (int)(threadIdx.x / TX);
    const int i = tx * TX + lane, j = ty * BY + row;
    const int k0 = tz * a.kchunk, k1 = min(a.nz, k0 + a.kchunk);   // box z-blocks end below the top slab
    const u32 ob = (u32)a.plane * 4u, cell0 = (u32)(j * a.px + i), o0 = cell0 * 4u + ob;
    const float *const eo[3] = {b.eb[0] - a.plane, b.eb[1] - a.plane, b.eb[2] - a.plane};
    const float *const ho[3] = {b.hb[0] - a.plane, b.hb[1] - a.plane, b.hb[2] - a.plane};
    float *const en[3] = {b.ebn[0] - a.plane, b.ebn[1] - a.plane, b.ebn[2] - a.plane};
    float *const hn[3] = {b.hbn[0] - a.plane, b.hbn[1] - a.plane, b.hbn[2] - a.plane};
    const Tab3 hu = uniform_tab3(a.htab);
    const bool fn = fin != 0;
    FCarry c;
    {
        // z-neighbour quantities of plane k0-1 (the E sweep's qkm)
        const u32 o = o0 + (u32)(k0 - 1) * ob;
        const u32 ox = (u32)__builtin_amdgcn_readlane((int)(o + 4u), TX - 1);
        c.qtx = hu.y.z * ldf(ho[1], o);
        c.qsy = hu.x.y * ldf(ho[0], o);
        c.ytx = hu.y.z * ldf(ho[1], o + (u32)a.px * 4u);
        c.xsy = hu.x.y * ldf(ho[0], ox);
        c.q = EQ{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        c.ybz = c.yax = c.xaz = c.xby = c.hx = c.hy = c.hz = 0.f;
    }
    FPlane A, B;
    fz_fetch<CM>(A, a, eo, ho, o0, cell0, ob, k0, euni);
    fz_fetch<CM>(B, a, eo, ho, o0, cell0, ob, k0 + 1, euni);
    fz_step(A, hu, a.etab, fn, c, en, hn, o0 + (u32)k0 * ob, ob, true, false);
    fz_fetch<CM>(A, a, eo, ho, o0, cell0, ob, min(k0 + 2, k1), euni);
    int kk = k0 + 1;
    // B holds plane kk, A plane kk+1 (the fetches past k1 re-read plane k1,
    // so that every path issues the same loads)
    for (; kk + 1 <= k1; kk += 2) {
        fz_step(B, hu, a.etab, fn, c, en, hn, o0 + (u32)kk * ob, ob, true, true);
        fz_fetch<CM>(B, a, eo, ho, o0, cell0, ob, min(kk + 2, k1), euni);
        fz_step(A, hu, a.etab, fn, c, en, hn, o0 + (u32)(kk + 1) * ob, ob, kk + 1 < k1, true);
        fz_fetch<CM>(A, a, eo, ho, o0, cell0, ob, min(kk + 3, k1), euni);
    }
    if (kk == k1) fz_step(B, hu, a.etab, fn, c, en, hn, o0 + (u32)kk * ob, ob, false, true);
}

// Interior tile box of the adjoint tiling (tiles tcols wide, rows BY, z-blocks
// of a.kchunk planes): the own cells of a tile and the neighbour quantities it
// evaluates (one cell beyond it on each side: the E gather reads H at i-1,
// j-1, k-1, the H gather E at i+1, j+1, k+1) all lie off the CPML slabs and
// the outer faces.  All zero when an axis has no such tile.
void adj_box(const KArgs &a, int tcols, int ib[6])
{
    const int n[3] = {a.nx, a.ny, a.nz}, w[3] = {tcols, BY, a.kchunk};
    const int nt[3] = {a.tiles_x, a.tiles_y, (a.nz + a.kchunk - 1) / a.kchunk};
    for (int q = 0; q < 6; ++q) ib[q] = 0;
    int box[6];
    for (int ax = 0; ax < 3; ++ax) {
        const int lo_c = (a.L[2 * ax] > 1 ? a.L[2 * ax] : 1) + 1;
        const int hi_c = n[ax] - (a.L[2 * ax + 1] > 1 ? a.L[2 * ax + 1] : 1) - 1;
        int lo = nt[ax], hi = 0;
        for (int t = 0; t < nt[ax]; ++t)
            if (t * w[ax] >= lo_c && (t + 1) * w[ax] <= hi_c) {
                lo = lo < t ? lo : t;
                hi = t + 1;
            }
        if (hi <= lo) return;
        box[ax] = lo;
        box[ax + 3] = hi;
    }
    for (int q = 0; q < 6; ++q) ib[q] = box[q];
}

// Interior tiles of the adjoint box, split into those the fused sweep takes
// (quiet: no event column -- source or probe -- among the tile's cells and
// the cells at i+1 and j+1 beyond it, and no design edge among its cells)
// and the others.  Each is its index in the box, x fastest.
void adj_fuse_tiles(const KArgs &a0, const std::vector<int> &e_first, const std::vector<int> &h_first,
                    const int sb[6], std::vector<int> &quiet, std::vector<int> &noisy)
{
    KArgs a = a0;
    a.tiles_x = (a.nx + TX - 1) / TX;
    adj_box(a, TX, a.ib);
    quiet.clear();
    noisy.clear();
    if (a.ib[5] == a.ib[2]) return;
    const int bx = a.ib[3] - a.ib[0], by = a.ib[4] - a.ib[1], bz = a.ib[5] - a.ib[2];
    std::vector<char> cq((size_t)bx * by, 1);   // column without events
    for (int ty = a.ib[1]; ty < a.ib[4]; ++ty)
        for (int tx = a.ib[0]; tx < a.ib[3]; ++tx) {
            const int i0 = tx * TX, j0 = ty * BY;
            bool q = true;
            for (int j = j0; j <= j0 + BY && q; ++j)
                for (int i = i0; i <= i0 + TX && q; ++i) {
                    const size_t col = (size_t)j * a.nx + i;
                    if ((col < e_first.size() && e_first[col] >= 0) || (col < h_first.size() && h_first[col] >= 0))
                        q = false;
                }
            cq[(size_t)(ty - a.ib[1]) * bx + (tx - a.ib[0])] = q;
        }
    for (int z = 0; z < bz; ++z) {
        const int k0 = (a.ib[2] + z) * a.kchunk, k1 = std::min(a.nz, k0 + a.kchunk);
        for (int y = 0; y < by; ++y)
            for (int x = 0; x < bx; ++x) {
                const int i0 = (a.ib[0] + x) * TX, j0 = (a.ib[1] + y) * BY;
                bool q = cq[(size_t)y * bx + x];
                // the (dilated) design box over the tile's own cells
                if (sb[3] > 0 && i0 < sb[0] + sb[3] && sb[0] < i0 + TX && j0 < sb[1] + sb[4] && sb[1] < j0 + BY &&
                    k0 < sb[2] + sb[5] && sb[2] < k1)
                    q = false;
                (q ? quiet : noisy).push_back((z * by + y) * bx + x);
            }
    }
}

// which: 0 adjoint E sweep, 1 adjoint H sweep, 2 the fused interior pass.  a0.ib[0] < 0 asks for the
// general variant alone (no interior split).
hipError_t launch_adj(int which, int mode, int euni, int huni, const KArgs &a0, const AArgs &b, int m, int fstep,
                      int sten, hipStream_t s)
{
    const bool stored = adj_stores_quantities(a0.nx);
    const int tcols = TX;
    KArgs a = a0;
    a.tiles_x = (a.nx + tcols - 1) / tcols;
    if (a0.ib[0] < 0) {
        for (int q = 0; q < 6; ++q) a.ib[q] = 0;
    } else {
        adj_box(a, tcols, a.ib);
    }
    // wide grids: the z-slab blocks of the box's columns and rows get a
    // variant with the z predicates only (the general variant costs 2.7x
    // the interior's time per cell there)
    a.zsplit = stored ? 0 : 1;
    a.sub[3] = 0;
    int n_in, n_x, n_z, n_gen;
    region_tiles(a, n_in, n_x, n_z, n_gen);
#define ADJ_E(CMV, V, NT, EU, HU)                                                                                \
    do {                                                                                                         \
        if (stored)                                                                                              \
            hipLaunchKernelGGL((adj_e_kernel<CMV, V>), dim3(NT), dim3(TX * BY), 0, s, a, bb, m, fstep, sten, EU); \
        else                                                                                                     \
            hipLaunchKernelGGL((adj_e_lds_kernel<CMV, V>), dim3(NT), dim3(TX * BY), 0, s, a, bb, m, fstep, sten, \
                               EU, HU);                                                                          \
    } while (0)
#define ADJ_H(CMV, V, NT, EU, HU)                                                                            \
    do {                                                                                                     \
        if (stored) hipLaunchKernelGGL((adj_h_kernel<CMV, V>), dim3(NT), dim3(TX * BY), 0, s, a, bb, EU, HU); \
        else hipLaunchKernelGGL((adj_h_lds_kernel<CMV, V>), dim3(NT), dim3(TX * BY), 0, s, a, bb, EU, HU);    \
    } while (0)
#define ADJ_V(MAC, V, NT)                                    \
    do {                                                     \
        if (mode == CLS_U32) MAC(CLS_U32, V, NT, euni, huni); \
        else if (mode == CLS_U16) MAC(CLS_U16, V, NT, euni, huni); \
        else MAC(CLS_UNIFORM, V, NT, 1, 1);                  \
    } while (0)
    // fused interior sweep (b.nfl > 0, which 2): the interior E and H sweeps
    // take the other tiles (b.tl).  The caller runs it after the previous
    // step's H sweep and before this step's, beside or after this step's E
    // sweep (they touch different cells).  An E-only pass (sten 0) covers the
    // whole box.
    const bool fuse = !stored && b.nfl > 0 && sten;
    AArgs bb = b;
    if (!fuse) bb.ntl = 0;
    if (fuse && n_in) n_in = b.ntl;
    if (which == 2) {
        if (fuse) {
            const int nt = b.nfl, fin = fstep >= 0 ? 1 : 0;
            if (mode == CLS_U32)
                hipLaunchKernelGGL((adj_fused_kernel<CLS_U32>), dim3(nt), dim3(TX * BY), 0, s, a, b, fin, euni);
            else if (mode == CLS_U16)
                hipLaunchKernelGGL((adj_fused_kernel<CLS_U16>), dim3(nt), dim3(TX * BY), 0, s, a, b, fin, euni);
            else hipLaunchKernelGGL((adj_fused_kernel<CLS_UNIFORM>), dim3(nt), dim3(TX * BY), 0, s, a, b, fin, 1);
        }
        return hipGetLastError();
    }
    if (which == 0) {
        if (n_in) ADJ_V(ADJ_E, V_IN, n_in);
        if (n_x) ADJ_V(ADJ_E, V_X, n_x);
        if (n_z) ADJ_V(ADJ_E, V_Z, n_z);
        if (n_gen) ADJ_V(ADJ_E, V_GEN, n_gen);
    } else {
        if (n_in) ADJ_V(ADJ_H, V_IN, n_in);
        if (n_x) ADJ_V(ADJ_H, V_X, n_x);
        if (n_z) ADJ_V(ADJ_H, V_Z, n_z);
        if (n_gen) ADJ_V(ADJ_H, V_GEN, n_gen);
    }
#undef ADJ_V
#undef ADJ_E
#undef ADJ_H
    return hipGetLastError();
}

hipError_t launch_snap(int storage, const KArgs &a, float *hist, int slot, hipStream_t s)
{
    const i64 nb = (i64)a.sb[3] * a.sb[4] * a.sb[5];
    dim3 grid((unsigned)((nb + 255) / 256)), block(256);
    switch (storage) {   // block-scaled storage has no adjoint (refused before any launch)
    case 0: hipLaunchKernelGGL((snap_kernel<float>), grid, block, 0, s, a, hist, slot); break;
    case 1: hipLaunchKernelGGL((snap_kernel<_Float16>), grid, block, 0, s, a, hist, slot); break;
    case 2: hipLaunchKernelGGL((snap_kernel<bf16_t>), grid, block, 0, s, a, hist, slot); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fdtd
