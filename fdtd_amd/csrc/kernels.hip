// kernels.hip -- the E and H half-step kernels for gfx950 (MI355X).
//
// One thread owns one (i, j) column of a 64 x BY tile and marches k over a
// z-chunk.  The previous (E sweep) or next (H sweep) plane of the two
// z-derivative operands is carried in registers, so each field component is
// fetched from HBM once per half-step; the i-1/j-1 (E) and i+1/j+1 (H)
// neighbours are re-reads of lines the neighbouring lanes and rows fetch in
// the same plane and hit in L1/L2.  Tiles are assigned to workgroups so that
// each of the 8 XCDs owns one contiguous band of tile rows, which keeps the
// halo lines of a tile in the L2 of the XCD that fetched them.
//
// The march is software-pipelined two planes deep: every value plane k reads
// (field operands, its own old values, the slab psi) is fetched into one of
// two register sets while plane k-1 is computed from the other, and the
// class index of plane k+1 rides along so that its coefficient-table entries
// are requested one plane ahead.  The loop is unrolled by two so the sets
// alternate without register copies (a copy of a register whose load is in
// flight would make the wave wait for it), and nothing is fetched behind a
// plane's own loads except the next coefficients: the in-order memory
// counter then lets plane k's compute start while plane k+1 is in flight.
// The per-plane z-profile is uniform across the workgroup and comes through
// the scalar data cache (read only).
//
// Per cell the kernels do exactly the reference's arithmetic, in its operand
// order, with no contraction (the library is built with -ffp-contract=off):
//   E: calc_e      fdtdrav.h:665-735, CPML fdtdrav.h:750-911
//   H: calc_h      fdtdrav.h:328-386, CPML fdtdrav.h:398-562
// The update coefficients come from a per-class table (one class index per
// cell, or none when the whole grid is one class) instead of the reference's
// 18 full-grid coefficient arrays; CPML kappa scaling and cpsi = Cb*d are
// applied in registers on the slab cells only; psi is stored only over the
// slabs.  Lumped-source injection, probe sampling and the FP64 DFT run inside
// the E (or H) kernel on the few cells that carry an event.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "device_common.h"
#include "internal.h"
#include "sweep_common.h"

// Occupancy floors of the two sweep kernels (waves per SIMD).
#ifdef FDTD_E_WAVES
#define E_OCC __attribute__((amdgpu_waves_per_eu(FDTD_E_WAVES, FDTD_E_WAVES)))
#else
#define E_OCC
#endif
#ifdef FDTD_H_WAVES
#define H_OCC __attribute__((amdgpu_waves_per_eu(FDTD_H_WAVES, FDTD_H_WAVES)))
#else
#define H_OCC
#endif

namespace fdtd {

// ---- E half-step ----------------------------------------------------------

// Column constants of one thread in the E sweep.
struct ECol {
    int i, j, sx, sy, mycol;
    i64 col;
    std::uint32_t o0;    // byte offset of (i, j, 0) from the guard-plane base
    std::uint32_t ob;    // bytes per plane
    bool ilo, ihi, jlo, jhi, xs, ys, in_snap;
    bool valid;          // i < nx (block-scaled storage keeps the lanes past nx alive)
    float4 pxe, pye;
};

// Everything the E update of plane k reads from memory.
struct EPlane {
    float hxc, hyc, hzc, hy_im, hz_im, hx_jm, hz_jm, ex0, ey0, ez0;
    float p_yx, p_zx, p_xy, p_zy, p_xz, p_yz;   // old psi of the slabs the cell is in
    int cln;                                      // class index of plane k+1
};

__device__ __forceinline__ bool e_zslab(const KArgs &a, int k)
{
    return (k > 0 && k < a.L[4]) || (k >= a.nz - a.L[5] && k < a.nz - 1);   // fdtdrav.h:853,879
}

template <class T, int CM, int V>
__device__ __forceinline__ void e_fetch(EPlane &p, const KArgs &a, const ECol &cc, const T *ex, const T *ey,
                                        const T *ez, const T *hx, const T *hy, const T *hz, const QPtrs &qp, int k,
                                        int k1)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    constexpr std::uint32_t S = sizeof(T);
    const std::uint32_t o = cc.o0 + (std::uint32_t)k * cc.ob;
    const std::uint32_t ojm = o - (std::uint32_t)a.px * S;
    const int px = a.px, ny = a.ny;
    // exponent words (block-scaled storage): H at the cell's block, at
    // i-1's and at row j-1's; E at the cell's block
    // H at i-1: the previous lane's own value (lane_prev in e_update); lane 0
    // fills from the column before the tile, one address for the wave.
    // Block-scaled storage loads it (the shift measured 3 % slower there).
    const std::uint32_t oim = is_bq<T>::value ? o - S : (std::uint32_t)__builtin_amdgcn_readfirstlane((int)o) - S;
    const QW<T> wh = ldq<T>(qp.hx, o), wim = ldq<T>(qp.hx, oim), wjm = ldq<T>(qp.hx, ojm);
    const QW<T> we = ldq<T>(qp.ex, o);
    p.hxc = fldw(hx, o, wh, 0);
    p.hyc = fldw(hy, o, wh, 1);
    p.hzc = fldw(hz, o, wh, 2);
    p.hy_im = fldw(hy, oim, wim, 1);
    p.hz_im = fldw(hz, oim, wim, 2);
    p.hx_jm = fldw(hx, ojm, wjm, 0);
    p.hz_jm = fldw(hz, ojm, wjm, 2);
    p.ex0 = fldnw(ex, o, we, 0);
    p.ey0 = fldnw(ey, o, we, 1);
    p.ez0 = fldnw(ez, o, we, 2);
    // class map (no guard planes): the entry of plane k+1
    p.cln = (CM != CLS_UNIFORM && k + 1 < k1)
                ? class_at<CM>(a.cls, (std::uint32_t)(k + 1) * (std::uint32_t)a.plane + (std::uint32_t)cc.col)
                : 0;
    const bool khi = !(V & V_Z) || k < a.nz - 1;
    p.p_yx = p.p_zx = p.p_xy = p.p_zy = p.p_xz = p.p_yz = 0.f;
    // Slab psi.  16-bit storage issues these loads on every path of the
    // variants with that axis's predicates (lanes off the slab read element
    // 0 of the array, which e_update does not use), so that the waits for a
    // plane count the same younger loads on every path; loaded only on the
    // slab lanes, the waits sized for the path without them also drained
    // part of the next plane's loads.  Measured on 512x512x256: FP16 general
    // E sweep 128.9 -> 108.3 us; FP32 0.35 % slower, so FP32 keeps the
    // conditional loads.
    constexpr bool ALL = sizeof(T) == 2;
    const bool nx_ = cc.xs && cc.jhi && khi;
    if ((V & V_X) && (ALL || nx_)) {
        const std::uint32_t q = nx_ ? (((std::uint32_t)k * ny + cc.j) * ra.Lx + cc.sx) * S : 0u;
        p.p_yx = ld(at((const PT *)ra.psi_eyx, q));
        p.p_zx = ld(at((const PT *)ra.psi_ezx, q));
    }
    const bool ny_ = cc.ys && cc.ihi && khi;
    if ((V & V_Y) && (ALL || ny_)) {
        const std::uint32_t q = ny_ ? (((std::uint32_t)k * ra.Ly + cc.sy) * px + cc.i) * S : 0u;
        p.p_xy = ld(at((const PT *)ra.psi_exy, q));
        p.p_zy = ld(at((const PT *)ra.psi_ezy, q));
    }
    const bool nz_ = (V & V_Z) && e_zslab(a, k) && cc.ihi && cc.jhi;
    if ((V & V_Z) && (ALL || nz_)) {
        const int sz = k < a.L[4] ? k : k - (a.nz - a.L[5]) + a.L[4];
        const std::uint32_t q = nz_ ? (((std::uint32_t)sz * ny + cc.j) * px + cc.i) * S : 0u;
        p.p_xz = ld(at((const PT *)ra.psi_exz, q));
        p.p_yz = ld(at((const PT *)ra.psi_eyz, q));
    }
}

template <class T, int CM, int V>
__device__ __forceinline__ void e_update(const EPlane &p, const Coef &t, const KArgs &a, const ECol &cc, T *ex,
                                         T *ey, T *ez, const QPtrs &qp, int k, int m, float &hx_km, float &hy_km,
                                         int &ev, int &ek)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    const int nz = a.nz;
    constexpr std::uint32_t S = sizeof(T);
    const std::uint32_t o = cc.o0 + (std::uint32_t)k * cc.ob;
    const float hxc = p.hxc, hyc = p.hyc, hzc = p.hzc;
    const float hy_im = is_bq<T>::value ? p.hy_im : lane_prev(hyc, p.hy_im);
    const float hz_im = is_bq<T>::value ? p.hz_im : lane_prev(hzc, p.hz_im);
    const float hx_jm = p.hx_jm, hz_jm = p.hz_jm;
    const float4 cx = t.x, cy = t.y, cz = t.z;
    const bool klo = !(V & V_Z) || k > 0, khi = !(V & V_Z) || k < nz - 1;
    const bool upd_x = cc.ihi && cc.jlo && cc.jhi && klo && khi;
    const bool upd_y = cc.ilo && cc.ihi && cc.jhi && klo && khi;
    const bool upd_z = cc.ilo && cc.ihi && cc.jlo && cc.jhi && khi;
    const bool zs = (V & V_Z) && e_zslab(a, k);

    // curl coefficients, divided by kappa on the slab cells (setABCPML)
    float cbx_hz = cx.y, cbx_hy = cx.z;
    float cby_hx = cy.y, cby_hz = cy.z;
    float cbz_hy = cz.y, cbz_hx = cz.z;
    float4 pze = make_float4(0.f, 0.f, 1.f, 0.f);
    if (cc.xs) { cby_hz = cby_hz / cc.pxe.z; cbz_hy = cbz_hy / cc.pxe.z; }
    if (cc.ys) { cbx_hz = cbx_hz / cc.pye.z; cbz_hx = cbz_hx / cc.pye.z; }
    if (zs) {
        pze = uload(a.pz_prof + k);
        cbx_hy = cbx_hy / pze.z;
        cby_hx = cby_hx / pze.z;
    }

    const float ex0 = p.ex0, ey0 = p.ey0, ez0 = p.ez0;
    float exn = ex0, eyn = ey0, ezn = ez0;
    if (upd_x) exn = (cx.x * ex0 + cbx_hz * (hzc - hz_jm)) + cbx_hy * (hyc - hy_km);
    if (upd_y) eyn = (cy.x * ey0 + cby_hx * (hxc - hx_km)) + cby_hz * (hzc - hz_im);
    if (upd_z) ezn = (cz.x * ez0 + cbz_hy * (hyc - hy_im)) + cbz_hx * (hxc - hx_jm);

    if (k == ek) {   // source edges and probes on this cell
        const EvE r = e_events(ra, ev, k, cc.mycol, m, upd_z, ez0, hyc, hy_im, hxc, hx_jm, exn, eyn, ezn);
        ev = r.ev;
        ek = r.ek;
        ezn = r.ezn;
    }

    // CPML psi recursions and corrections, fdtdrav.h:750-911
    const float dx = ra.d[0], dy = ra.d[1], dz = ra.d[2];
    if (cc.xs && cc.jhi && khi) {
        const std::uint32_t q = (((std::uint32_t)k * ra.ny + cc.j) * ra.Lx + cc.sx) * S;
        const float pey = cc.pxe.x * p.p_yx + cc.pxe.y * (hzc - hz_im);
        const float pez = cc.pxe.x * p.p_zx + cc.pxe.y * (hyc - hy_im);
        st(at((PT *)ra.psi_eyx, q), pey);
        st(at((PT *)ra.psi_ezx, q), pez);
        if (klo) eyn = eyn + (cy.z * dx) * pey;
        if (cc.jlo) ezn = ezn + (cz.y * dx) * pez;
    }
    if (cc.ys && cc.ihi && khi) {
        const std::uint32_t q = (((std::uint32_t)k * ra.Ly + cc.sy) * ra.px + cc.i) * S;
        const float pex = cc.pye.x * p.p_xy + cc.pye.y * (hzc - hz_jm);
        const float pez = cc.pye.x * p.p_zy + cc.pye.y * (hxc - hx_jm);
        st(at((PT *)ra.psi_exy, q), pex);
        st(at((PT *)ra.psi_ezy, q), pez);
        if (klo) exn = exn + (cx.y * dy) * pex;
        if (cc.ilo) ezn = ezn + (cz.z * dy) * pez;
    }
    if (zs && cc.ihi && cc.jhi) {
        const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
        const std::uint32_t q = (((std::uint32_t)sz * ra.ny + cc.j) * ra.px + cc.i) * S;
        const float pex = pze.x * p.p_xz + pze.y * (hyc - hy_km);
        const float pey = pze.x * p.p_yz + pze.y * (hxc - hx_km);
        st(at((PT *)ra.psi_exz, q), pex);
        st(at((PT *)ra.psi_eyz, q), pey);
        if (cc.jlo) exn = exn + (cx.z * dz) * pex;
        if (cc.ilo) eyn = eyn + (cy.y * dz) * pey;
    }
    const std::uint32_t bx = fst(ex, o, upd_x, exn, ex0, cc.valid);
    const std::uint32_t by = fst(ey, o, upd_y, eyn, ey0, cc.valid);
    const std::uint32_t bz = fst(ez, o, upd_z, ezn, ez0, cc.valid);
    stq<T>(qp.ex, o, bx, by, bz, cc.valid && (cc.i & (BQ_BLOCK - 1)) == 0);
    if (cc.in_snap && k >= a.sb[2] && k < a.sb[2] + a.sb[5])   // stored-precision value, as the next step reads it
        e_snapshot(ra, cc.i, cc.j, k, m, upd_x ? rnd(exn, ex) : ex0, upd_y ? rnd(eyn, ex) : ey0,
                   upd_z ? rnd(ezn, ex) : ez0);
    hx_km = carry(hxc);
    hy_km = carry(hyc);
}

template <class T, int CM, int V>
__global__ void __launch_bounds__(TX * BY) E_OCC e_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    ECol cc;
    cc.i = tx * TX + (int)(threadIdx.x & (TX - 1));
    cc.j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i = cc.i, j = cc.j;
    // block-scaled storage encodes 8-cell blocks across lanes: lanes past nx
    // (inside the padded pitch) stay, storing nothing
    if (j >= ny || (!is_bq<T>::value && i >= nx)) return;
    cc.valid = !is_bq<T>::value || i < nx;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    cc.col = (i64)j * a.px + i;
    cc.mycol = j * nx + i;
    cc.ob = (std::uint32_t)a.plane * (std::uint32_t)sizeof(T);
    cc.o0 = (std::uint32_t)cc.col * (std::uint32_t)sizeof(T) + cc.ob;
    const QPtrs qp = qptrs(a);

    // allocation bases (guard plane below k = 0)
    T *__restrict__ ex = (T *)a.ex - a.plane;
    T *__restrict__ ey = (T *)a.ey - a.plane;
    T *__restrict__ ez = (T *)a.ez - a.plane;
    const T *__restrict__ hx = (const T *)a.hx - a.plane;
    const T *__restrict__ hy = (const T *)a.hy - a.plane;
    const T *__restrict__ hz = (const T *)a.hz - a.plane;

    // column-constant predicates (reference region guards)
    cc.ilo = !(V & V_X) || i > 0;
    cc.ihi = !(V & V_X) || i < nx - 1;
    cc.jlo = !(V & V_Y) || j > 0;
    cc.jhi = !(V & V_Y) || j < ny - 1;
    cc.xs = (V & V_X) && ((cc.ilo && i < a.L[0]) || (i >= nx - a.L[1] && cc.ihi));   // E x-slab, fdtdrav.h:752,777
    cc.ys = (V & V_Y) && cc.valid && ((cc.jlo && j < a.L[2]) || (j >= ny - a.L[3] && cc.jhi));   // E y-slab, fdtdrav.h:803,827
    cc.sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
    cc.sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    cc.pxe = cc.xs ? a.px_prof[i] : make_float4(0.f, 0.f, 1.f, 0.f);
    cc.pye = cc.ys ? a.py_prof[j] : make_float4(0.f, 0.f, 1.f, 0.f);
    cc.in_snap = a.snap != nullptr && i >= a.sb[0] && i < a.sb[0] + a.sb[3] && j >= a.sb[1] &&
                 j < a.sb[1] + a.sb[4];

    int ev = cc.valid ? first_event(a, a.e_first, cc.mycol, k0) : -1;
    int ek = first_k(a, ev, cc.mycol);

    const std::uint32_t okm = cc.o0 + (std::uint32_t)(k0 - 1) * cc.ob;
    const QW<T> wkm = ldq<T>(qp.hx, okm);
    float hx_km = fldw(hx, okm, wkm, 0);
    float hy_km = fldw(hy, okm, wkm, 1);

    EPlane A, B;
    Coef tA, tB;
    if (CM == CLS_UNIFORM) {
        tA = coef_of<CM>(a.etab, 0);
        tB = tA;
    } else {
        tA = coef_of<CM>(a.etab, class_at<CM>(a.cls, (std::uint32_t)k0 * (std::uint32_t)a.plane + (std::uint32_t)cc.col));
    }
    e_fetch<T, CM, V>(A, a, cc, ex, ey, ez, hx, hy, hz, qp, k0, k1);
    int k = k0;
    // Steady state: the fetches are unconditional, so on every path the
    // next plane's loads are the younger ones when this plane's are awaited.
    for (; k + 2 < k1; k += 2) {
        e_fetch<T, CM, V>(B, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_of<CM>(a.etab, A.cln);
        e_update<T, CM, V>(A, tA, a, cc, ex, ey, ez, qp, k, m, hx_km, hy_km, ev, ek);
        e_fetch<T, CM, V>(A, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 2, k1);
        if (CM != CLS_UNIFORM) tA = coef_of<CM>(a.etab, B.cln);
        e_update<T, CM, V>(B, tB, a, cc, ex, ey, ez, qp, k + 1, m, hx_km, hy_km, ev, ek);
    }
    // last one or two planes of the chunk
    if (k + 1 < k1) {
        e_fetch<T, CM, V>(B, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_of<CM>(a.etab, A.cln);
        e_update<T, CM, V>(A, tA, a, cc, ex, ey, ez, qp, k, m, hx_km, hy_km, ev, ek);
        e_update<T, CM, V>(B, tB, a, cc, ex, ey, ez, qp, k + 1, m, hx_km, hy_km, ev, ek);
    } else {
        e_update<T, CM, V>(A, tA, a, cc, ex, ey, ez, qp, k, m, hx_km, hy_km, ev, ek);
    }
}

// ---- H half-step ----------------------------------------------------------

struct HCol {
    int i, j, sx, sy, mycol;
    i64 col;
    std::uint32_t o0;    // byte offset of (i, j, 0) from the guard-plane base
    std::uint32_t ob;    // bytes per plane
    bool ihi, jhi, xs, ys;
    bool valid;          // i < nx (block-scaled storage keeps the lanes past nx alive)
    float4 pxm, pym;
};

// Everything the H update of plane k reads (E of planes k and k+1 around
// it, its own H, slab psi).
struct HPlane {
    float ex_kp, ey_kp, ezc, ez_ip, ey_ip, ez_jp, ex_jp, hx0, hy0, hz0;
    float p_yx, p_zx, p_xy, p_zy, p_xz, p_yz;
    int cln;
};

__device__ __forceinline__ bool h_zslab(const KArgs &a, int k)
{
    return k < a.L[4] || (k >= a.nz - a.L[5] && k < a.nz - 1);   // fdtdrav.h:500,531
}

template <class T, int CM, int V>
__device__ __forceinline__ void h_fetch(HPlane &p, const KArgs &a, const HCol &cc, const T *ex, const T *ey,
                                        const T *ez, const T *hx, const T *hy, const T *hz, const QPtrs &qp, int k,
                                        int k1)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    constexpr std::uint32_t S = sizeof(T);
    const std::uint32_t o = cc.o0 + (std::uint32_t)k * cc.ob;
    const std::uint32_t okp = o + cc.ob;
    const std::uint32_t ojp = o + (std::uint32_t)a.px * S;
    const int px = a.px, ny = a.ny;
    // exponent words (block-scaled storage): E at plane k+1, at the cell's
    // block, at i+1's and at row j+1's; H at the cell's block
    // E at i+1: the next lane's own value (lane_next in h_update); lane 63
    // fills from the column after the tile, one address for the wave
    // (block-scaled storage loads it)
    const std::uint32_t oip =
        is_bq<T>::value ? o + S : (std::uint32_t)__builtin_amdgcn_readfirstlane((int)o) + (std::uint32_t)TX * S;
    const QW<T> wkp = ldq<T>(qp.ex, okp), we = ldq<T>(qp.ex, o), wip = ldq<T>(qp.ex, oip);
    const QW<T> wjp = ldq<T>(qp.ex, ojp), wh = ldq<T>(qp.hx, o);
    p.ex_kp = fldw(ex, okp, wkp, 0);
    p.ey_kp = fldw(ey, okp, wkp, 1);
    p.ezc = fldw(ez, o, we, 2);
    p.ez_ip = fldw(ez, oip, wip, 2);
    p.ey_ip = fldw(ey, oip, wip, 1);
    p.ez_jp = fldw(ez, ojp, wjp, 2);
    p.ex_jp = fldw(ex, ojp, wjp, 0);
    p.hx0 = fldnw(hx, o, wh, 0);
    p.hy0 = fldnw(hy, o, wh, 1);
    p.hz0 = fldnw(hz, o, wh, 2);
    p.cln = (CM != CLS_UNIFORM && k + 1 < k1)
                ? class_at<CM>(a.cls, (std::uint32_t)(k + 1) * (std::uint32_t)a.plane + (std::uint32_t)cc.col)
                : 0;
    p.p_yx = p.p_zx = p.p_xy = p.p_zy = p.p_xz = p.p_yz = 0.f;
    // slab psi: on every path for 16-bit storage (as in e_fetch)
    constexpr bool ALL = sizeof(T) == 2;
    if ((V & V_X) && (ALL || cc.xs)) {
        const std::uint32_t q = cc.xs ? (((std::uint32_t)k * ny + cc.j) * ra.Lx + cc.sx) * S : 0u;
        p.p_yx = ld(at((const PT *)ra.psi_hyx, q));
        p.p_zx = ld(at((const PT *)ra.psi_hzx, q));
    }
    if ((V & V_Y) && (ALL || cc.ys)) {
        const std::uint32_t q = cc.ys ? (((std::uint32_t)k * ra.Ly + cc.sy) * px + cc.i) * S : 0u;
        p.p_xy = ld(at((const PT *)ra.psi_hxy, q));
        p.p_zy = ld(at((const PT *)ra.psi_hzy, q));
    }
    const bool nz_ = (V & V_Z) && h_zslab(a, k) && cc.valid;
    if ((V & V_Z) && (ALL || nz_)) {
        const int sz = k < a.L[4] ? k : k - (a.nz - a.L[5]) + a.L[4];
        const std::uint32_t q = nz_ ? (((std::uint32_t)sz * ny + cc.j) * px + cc.i) * S : 0u;
        p.p_xz = ld(at((const PT *)ra.psi_hxz, q));
        p.p_yz = ld(at((const PT *)ra.psi_hyz, q));
    }
}

template <class T, int CM, int V>
__device__ __forceinline__ void h_update(const HPlane &p, const Coef &t, const KArgs &a, const HCol &cc, T *hx,
                                         T *hy, T *hz, const QPtrs &qp, int k, int m, float &ex_c, float &ey_c,
                                         int &ev, int &ek)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    const int nz = a.nz;
    constexpr std::uint32_t S = sizeof(T);
    const std::uint32_t o = cc.o0 + (std::uint32_t)k * cc.ob;
    const float ex_kp = p.ex_kp, ey_kp = p.ey_kp;
    const float ezc = p.ezc;
    const float ez_ip = is_bq<T>::value ? p.ez_ip : lane_next(ezc, p.ez_ip);
    const float ey_ip = is_bq<T>::value ? p.ey_ip : lane_next(ey_c, p.ey_ip);
    const float ez_jp = p.ez_jp, ex_jp = p.ex_jp;
    const float4 cx = t.x, cy = t.y, cz = t.z;
    const bool khi = !(V & V_Z) || k < nz - 1;
    const bool upd_x = cc.jhi && khi;
    const bool upd_y = cc.ihi && khi;
    const bool upd_z = cc.ihi && cc.jhi;
    const bool zs = (V & V_Z) && h_zslab(a, k) && cc.valid;

    float chx_ey = cx.y, chx_ez = cx.z;
    float chy_ez = cy.y, chy_ex = cy.z;
    float chz_ex = cz.y, chz_ey = cz.z;
    float4 pzm = make_float4(0.f, 0.f, 1.f, 0.f);
    if (cc.xs) { chy_ez = chy_ez / cc.pxm.z; chz_ey = chz_ey / cc.pxm.z; }
    if (cc.ys) { chx_ez = chx_ez / cc.pym.z; chz_ex = chz_ex / cc.pym.z; }
    if (zs) {
        pzm = uload(a.pz_prof + nz + k);
        chx_ey = chx_ey / pzm.z;
        chy_ex = chy_ex / pzm.z;
    }

    const float hx0 = p.hx0, hy0 = p.hy0, hz0 = p.hz0;
    float hxn = hx0, hyn = hy0, hzn = hz0;
    if (upd_x) hxn = (cx.x * hx0 + chx_ey * (ey_kp - ey_c)) + chx_ez * (ez_jp - ezc);
    if (upd_y) hyn = (cy.x * hy0 + chy_ez * (ez_ip - ezc)) + chy_ex * (ex_kp - ex_c);
    if (upd_z) hzn = (cz.x * hz0 + chz_ex * (ex_jp - ex_c)) + chz_ey * (ey_ip - ey_c);

    if (k == ek) {   // probes on this cell
        const EvH r = h_events(ra, ev, k, cc.mycol, m, hxn, hyn, hzn);
        ev = r.ev;
        ek = r.ek;
    }

    const float dx = ra.d[0], dy = ra.d[1], dz = ra.d[2];
    if (cc.xs) {
        const std::uint32_t q = (((std::uint32_t)k * ra.ny + cc.j) * ra.Lx + cc.sx) * S;
        const float phy = cc.pxm.x * p.p_yx + cc.pxm.y * (ez_ip - ezc);
        const float phz = cc.pxm.x * p.p_zx + cc.pxm.y * (ey_ip - ey_c);
        st(at((PT *)ra.psi_hyx, q), phy);
        st(at((PT *)ra.psi_hzx, q), phz);
        if (khi) hyn = hyn + (cy.y * dx) * phy;
        if (cc.jhi) hzn = hzn + (cz.z * dx) * phz;
    }
    if (cc.ys) {
        const std::uint32_t q = (((std::uint32_t)k * ra.Ly + cc.sy) * ra.px + cc.i) * S;
        const float phx = cc.pym.x * p.p_xy + cc.pym.y * (ez_jp - ezc);
        const float phz = cc.pym.x * p.p_zy + cc.pym.y * (ex_jp - ex_c);
        st(at((PT *)ra.psi_hxy, q), phx);
        st(at((PT *)ra.psi_hzy, q), phz);
        if (khi) hxn = hxn + (cx.z * dy) * phx;
        if (cc.ihi) hzn = hzn + (cz.y * dy) * phz;
    }
    if (zs) {
        const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
        const std::uint32_t q = (((std::uint32_t)sz * ra.ny + cc.j) * ra.px + cc.i) * S;
        const float phx = pzm.x * p.p_xz + pzm.y * (ey_kp - ey_c);
        const float phy = pzm.x * p.p_yz + pzm.y * (ex_kp - ex_c);
        st(at((PT *)ra.psi_hxz, q), phx);
        st(at((PT *)ra.psi_hyz, q), phy);
        if (cc.jhi) hxn = hxn + (cx.y * dz) * phx;
        if (cc.ihi) hyn = hyn + (cy.z * dz) * phy;
    }
    const std::uint32_t bx = fst(hx, o, upd_x, hxn, hx0, cc.valid);
    const std::uint32_t by = fst(hy, o, upd_y, hyn, hy0, cc.valid);
    const std::uint32_t bz = fst(hz, o, upd_z, hzn, hz0, cc.valid);
    stq<T>(qp.hx, o, bx, by, bz, cc.valid && (cc.i & (BQ_BLOCK - 1)) == 0);
    ex_c = carry(ex_kp);
    ey_c = carry(ey_kp);
}

template <class T, int CM, int V>
__global__ void __launch_bounds__(TX * BY) H_OCC h_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    HCol cc;
    cc.i = tx * TX + (int)(threadIdx.x & (TX - 1));
    cc.j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i = cc.i, j = cc.j;
    if (j >= ny || (!is_bq<T>::value && i >= nx)) return;
    cc.valid = !is_bq<T>::value || i < nx;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    cc.col = (i64)j * a.px + i;
    cc.mycol = j * nx + i;
    cc.ob = (std::uint32_t)a.plane * (std::uint32_t)sizeof(T);
    cc.o0 = (std::uint32_t)cc.col * (std::uint32_t)sizeof(T) + cc.ob;
    const QPtrs qp = qptrs(a);

    // allocation bases (guard plane below k = 0)
    const T *__restrict__ ex = (const T *)a.ex - a.plane;
    const T *__restrict__ ey = (const T *)a.ey - a.plane;
    const T *__restrict__ ez = (const T *)a.ez - a.plane;
    T *__restrict__ hx = (T *)a.hx - a.plane;
    T *__restrict__ hy = (T *)a.hy - a.plane;
    T *__restrict__ hz = (T *)a.hz - a.plane;

    cc.ihi = !(V & V_X) || i < nx - 1;
    cc.jhi = !(V & V_Y) || j < ny - 1;
    cc.xs = (V & V_X) && (i < a.L[0] || (i >= nx - a.L[1] && cc.ihi));   // H x-slab, fdtdrav.h:402,424
    cc.ys = (V & V_Y) && cc.valid && (j < a.L[2] || (j >= ny - a.L[3] && cc.jhi));   // H y-slab, fdtdrav.h:447,471
    cc.sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
    cc.sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    cc.pxm = cc.xs ? a.px_prof[nx + i] : make_float4(0.f, 0.f, 1.f, 0.f);
    cc.pym = cc.ys ? a.py_prof[ny + j] : make_float4(0.f, 0.f, 1.f, 0.f);

    int ev = cc.valid ? first_event(a, a.h_first, cc.mycol, k0) : -1;
    int ek = first_k(a, ev, cc.mycol);

    const std::uint32_t ok0 = cc.o0 + (std::uint32_t)k0 * cc.ob;
    const QW<T> wk0 = ldq<T>(qp.ex, ok0);
    float ex_c = fldw(ex, ok0, wk0, 0);
    float ey_c = fldw(ey, ok0, wk0, 1);

    HPlane A, B;
    Coef tA, tB;
    if (CM == CLS_UNIFORM) {
        tA = coef_of<CM>(a.htab, 0);
        tB = tA;
    } else {
        tA = coef_of<CM>(a.htab, class_at<CM>(a.cls, (std::uint32_t)k0 * (std::uint32_t)a.plane + (std::uint32_t)cc.col));
    }
    h_fetch<T, CM, V>(A, a, cc, ex, ey, ez, hx, hy, hz, qp, k0, k1);
    int k = k0;
    for (; k + 2 < k1; k += 2) {
        h_fetch<T, CM, V>(B, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_of<CM>(a.htab, A.cln);
        h_update<T, CM, V>(A, tA, a, cc, hx, hy, hz, qp, k, m, ex_c, ey_c, ev, ek);
        h_fetch<T, CM, V>(A, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 2, k1);
        if (CM != CLS_UNIFORM) tA = coef_of<CM>(a.htab, B.cln);
        h_update<T, CM, V>(B, tB, a, cc, hx, hy, hz, qp, k + 1, m, ex_c, ey_c, ev, ek);
    }
    if (k + 1 < k1) {
        h_fetch<T, CM, V>(B, a, cc, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_of<CM>(a.htab, A.cln);
        h_update<T, CM, V>(A, tA, a, cc, hx, hy, hz, qp, k, m, ex_c, ey_c, ev, ek);
        h_update<T, CM, V>(B, tB, a, cc, hx, hy, hz, qp, k + 1, m, ex_c, ey_c, ev, ek);
    } else {
        h_update<T, CM, V>(A, tA, a, cc, hx, hy, hz, qp, k, m, ex_c, ey_c, ev, ek);
    }
}

// ---- launchers --------------------------------------------------------------
template <class T, int V>
static void launch_v(bool efield, int mode, const KArgs &a, int m, int nt, hipStream_t s)
{
    dim3 grid(nt), block(TX * BY);
    if (efield) {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((e_kernel<T, CLS_UNIFORM, V>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((e_kernel<T, CLS_U16, V>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((e_kernel<T, CLS_U32, V>), grid, block, 0, s, a, m);
    } else {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((h_kernel<T, CLS_UNIFORM, V>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((h_kernel<T, CLS_U16, V>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((h_kernel<T, CLS_U32, V>), grid, block, 0, s, a, m);
    }
}

// One launch per non-empty region variant: interior box, x-edge tiles,
// general tiles.
template <class T>
static hipError_t launch_t(bool efield, int mode, const KArgs &a, int m, hipStream_t s)
{
    int nb, nx, nzt, ng;   // nzt = 0: the forward sweeps run without zsplit
    region_tiles(a, nb, nx, nzt, ng);
    if (nb > 0) launch_v<T, V_IN>(efield, mode, a, m, nb, s);
    if (nx > 0) launch_v<T, V_X>(efield, mode, a, m, nx, s);
    if (ng > 0) launch_v<T, V_GEN>(efield, mode, a, m, ng, s);
    return hipGetLastError();
}

// storage: 0 FP32, 1 FP16, 2 BF16, 3 block-scaled int16 (fields; psi FP16).  e_mode/h_mode: ClassMode of each sweep.
hipError_t launch_half_step(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s)
{
    switch (storage) {
    case 0: return launch_t<float>(efield, mode, a, m, s);
    case 1: return launch_t<_Float16>(efield, mode, a, m, s);
    case 2: return launch_t<bf16_t>(efield, mode, a, m, s);
    case 3: return launch_t<bq16_t>(efield, mode, a, m, s);
    default: return hipErrorInvalidValue;
    }
}

// The general variant alone (the rows and z-blocks outside the interior box;
// a in one-cell tile units).  The 16-bit pair sweeps hand it their slab rows
// and z-blocks: the one-cell kernel holds half the per-thread state there.
hipError_t launch_half_step_general(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s)
{
    int nb, nx, nzt, ng;   // nzt = 0: the forward sweeps run without zsplit
    region_tiles(a, nb, nx, nzt, ng);
    if (ng == 0) return hipSuccess;
    switch (storage) {
    case 0: launch_v<float, V_GEN>(efield, mode, a, m, ng, s); break;
    case 1: launch_v<_Float16, V_GEN>(efield, mode, a, m, ng, s); break;
    case 2: launch_v<bf16_t, V_GEN>(efield, mode, a, m, ng, s); break;
    case 3: launch_v<bq16_t, V_GEN>(efield, mode, a, m, ng, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int tile_rows() { return BY; }
int tile_cols() { return TX; }

}  // namespace fdtd
