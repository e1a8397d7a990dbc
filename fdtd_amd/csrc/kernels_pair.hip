// kernels_pair.hip -- the E and H half-step kernels for 16-bit storage
// (FP16 / BF16 / block-scaled int16), two adjacent x cells per thread.
//
// With 2-byte values a one-cell-per-thread sweep moves half the bytes per
// load instruction of the FP32 sweep and runs out of loads in flight long
// before HBM is busy.  Here a thread owns the cell pair (i, i+1), i even, of
// a 128 x BY tile: every field stream of a plane is one 4-byte load per lane
// holding both cells (the same bytes per instruction as the FP32 sweep), the
// pair stays packed in one register until it is used, and the two
// x-neighbour values that fall outside the pair (i-1 for the E sweep, i+2 for
// the H sweep) are single 2-byte loads.  The z-march, the two-deep software
// pipeline, the class-table prefetch, the XCD tile mapping and the per-cell
// arithmetic are those of kernels.hip; only the packing differs, so the
// stored bits equal the one-cell kernels' bit for bit.  Block-scaled storage
// adds the exponent byte of the pair's 8-cell block to every packed load;
// the four lanes of a block (one DPP quad) agree on the block's new exponent
// with two lane permutations before they store.
//   E: calc_e  fdtdrav.h:665-735, CPML fdtdrav.h:750-911
//   H: calc_h  fdtdrav.h:328-386, CPML fdtdrav.h:398-562
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>
#include <type_traits>

#include "device_common.h"
#include "internal.h"
#include "sweep_common.h"

namespace fdtd {

constexpr int TXP = 2 * TX;   // cells per tile row

// The exponent word of the previous / next lane's block (block-scaled storage;
// the other storage types carry none).
template <class T>
__device__ __forceinline__ QW<T> qw_prev(const QW<T> &q, const QW<T> &) { return q; }
template <>
__device__ __forceinline__ QW<bq16_t> qw_prev<bq16_t>(const QW<bq16_t> &q, const QW<bq16_t> &f)
{
    return QW<bq16_t>{lane_prev_u(q.w, f.w)};
}
template <class T>
__device__ __forceinline__ QW<T> qw_next(const QW<T> &q, const QW<T> &) { return q; }
template <>
__device__ __forceinline__ QW<bq16_t> qw_next<bq16_t>(const QW<bq16_t> &q, const QW<bq16_t> &f)
{
    return QW<bq16_t>{lane_next_u(q.w, f.w)};
}

// Occupancy floor of the pair sweeps (waves per SIMD), for tuning builds.
#ifdef FDTD_PAIR_WAVES
#define P_OCC __attribute__((amdgpu_waves_per_eu(FDTD_PAIR_WAVES)))
#else
#define P_OCC
#endif

hipError_t launch_half_step_general(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s);

typedef std::uint32_t u32;

// ---- packed pairs: cell i in the low half, cell i+1 in the high half -------
__device__ __forceinline__ float lo_f(u32 u, const _Float16 *)
{
    return (float)__builtin_bit_cast(_Float16, (std::uint16_t)(u & 0xFFFFu));
}
__device__ __forceinline__ float hi_f(u32 u, const _Float16 *)
{
    return (float)__builtin_bit_cast(_Float16, (std::uint16_t)(u >> 16));
}
__device__ __forceinline__ float lo_f(u32 u, const bf16_t *) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float hi_f(u32 u, const bf16_t *) { return __uint_as_float(u & 0xFFFF0000u); }

// storage bits of v (round to nearest even, as the one-cell kernels store)
__device__ __forceinline__ u32 bits16(float v, const _Float16 *)
{
    return (u32)__builtin_bit_cast(std::uint16_t, (_Float16)v);
}
__device__ __forceinline__ u32 bits16(float v, const bf16_t *)
{
    u32 u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}

template <class T>
__device__ __forceinline__ u32 ld2(const T *base, u32 off)
{
    return *(const u32 *)((const char *)base + off);
}
template <class T>
__device__ __forceinline__ u32 ldn2(const T *base, u32 off)
{
    return __builtin_nontemporal_load((const u32 *)((const char *)base + off));
}
template <class T>
__device__ __forceinline__ void stn2(T *base, u32 off, u32 v)
{
    __builtin_nontemporal_store(v, (u32 *)((char *)base + off));
}

// new value of a pair: updated halves replaced, the others keep their bits
template <class T>
__device__ __forceinline__ u32 merge2(u32 old, bool ua, float va, bool ub, float vb)
{
    const u32 lo = ua ? bits16(va, (const T *)nullptr) : (old & 0xFFFFu);
    const u32 hi = ub ? bits16(vb, (const T *)nullptr) : (old >> 16);
    return lo | (hi << 16);
}

// Values of a packed pair (cell i, cell i+1) of component c.
template <class T>
__device__ __forceinline__ float dlo(u32 m, const QW<T> &, int)
{
    return lo_f(m, (const T *)nullptr);
}
template <class T>
__device__ __forceinline__ float dhi(u32 m, const QW<T> &, int)
{
    return hi_f(m, (const T *)nullptr);
}
template <>
__device__ __forceinline__ float dlo<bq16_t>(u32 m, const QW<bq16_t> &q, int c)
{
    return ldexpf((float)(std::int16_t)(m & 0xFFFFu), qe(q, c));
}
template <>
__device__ __forceinline__ float dhi<bq16_t>(u32 m, const QW<bq16_t> &q, int c)
{
    return ldexpf((float)(std::int16_t)(m >> 16), qe(q, c));
}

// Store of a pair's new values (va / vb where ua / ub, else the old ones,
// oa / ob decoded and old_m as stored).  FP16 / BF16: only when a cell
// changes; the other half keeps its bits.  Block-scaled: every lane of the
// block re-encodes its cells (a cell past nx counts as 0 and keeps its bits)
// under the block's new exponent; returns that exponent's byte.
template <class T>
__device__ __forceinline__ u32 stp(T *base, u32 o, u32 old_m, float, float, bool ua, float va, bool ub, float vb,
                                   bool, bool)
{
    if (ua || ub) stn2(base, o, merge2<T>(old_m, ua, va, ub, vb));
    return 0;
}
template <>
__device__ __forceinline__ u32 stp<bq16_t>(bq16_t *base, u32 o, u32 old_m, float oa, float ob, bool ua, float va,
                                           bool ub, float vb, bool valid_a, bool valid_b)
{
    const float xa = !valid_a ? 0.f : ua ? va : oa;
    const float xb = !valid_b ? 0.f : ub ? vb : ob;
    float mx = fmaxf(fabsf(xa), fabsf(xb));
    mx = fmaxf(mx, dpp_f<DPP_QUAD_XOR1>(mx));
    mx = fmaxf(mx, dpp_f<DPP_QUAD_XOR2>(mx));
    const int e = bq_exponent(mx);
    if (valid_a) {
        const u32 wa = (u32)(std::uint16_t)(std::int16_t)bq_mantissa(xa, e);
        const u32 wb = valid_b ? (u32)(std::uint16_t)(std::int16_t)bq_mantissa(xb, e) : (old_m >> 16);
        stn2(base, o, wa | (wb << 16));
    }
    return (u32)(e + BQ_BIAS);
}

// class indices of the pair at cell index `cell` (even)
template <int CM>
__device__ __forceinline__ void class_pair(const void *cls, u32 cell, int &ca, int &cb)
{
    if (CM == CLS_U16) {
        const u32 u = *at((const u32 *)cls, cell * 2u);
        ca = (int)(u & 0xFFFFu);
        cb = (int)(u >> 16);
    } else if (CM == CLS_U32) {
        const uint2 u = *(const uint2 *)((const char *)cls + cell * 4u);
        ca = (int)u.x;
        cb = (int)u.y;
    } else {
        ca = cb = 0;
    }
}

struct CoefPair {
    Coef a, b;
};

template <int CM>
__device__ __forceinline__ CoefPair coef_pair(const float4 *tab, int ca, int cb)
{
    CoefPair t;
    t.a = coef_of<CM>(tab, ca);
    t.b = coef_of<CM>(tab, cb);
    return t;
}

// Per-cell column constants (cell 0 = i, cell 1 = i+1).
struct PCell {
    int i, sx, mycol;
    bool ilo, ihi, xs, valid, in_snap;   // valid: i < nx
    float4 px;   // CPML x profile of this cell (E or H nodes)
};

// Row constants of a thread.
struct PRow {
    int j, sy;
    i64 col;      // j * px + i (cell 0)
    u32 o0, ob;   // byte offset of cell 0 at k = 0 from the guard-plane base; bytes per plane
    bool jlo, jhi, ys;
    float4 py;
};

// ============================ E sweep ======================================

template <class T>
struct EPair {
    u32 hx, hy, hz, hx_jm, hz_jm, ex, ey, ez;   // packed pairs
    QW<T> qh, qh_jm, qe;                        // exponent words of H, H at row j-1, E
    u32 hy_fill, hz_fill;                       // H pair before the tile (lane 0's i-2, i-1)
    QW<T> qh_fill;                              // its exponent word
    float p_yx[2], p_zx[2], p_xy[2], p_zy[2], p_xz[2], p_yz[2];
    int cla, clb;                                // class indices of plane k+1
};

template <class T, int CM, int V>
__device__ __forceinline__ void ep_fetch(EPair<T> &p, const KArgs &a, const PRow &r, const PCell (&c)[2],
                                         const T *ex, const T *ey, const T *ez, const T *hx, const T *hy,
                                         const T *hz, const QPtrs &qp, int k, int k1)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    constexpr u32 S = sizeof(T);
    const u32 o = r.o0 + (u32)k * r.ob;
    const u32 ojm = o - (u32)a.px * S;
    p.hx = ld2(hx, o);
    p.hy = ld2(hy, o);
    p.hz = ld2(hz, o);
    p.hx_jm = ld2(hx, ojm);
    p.hz_jm = ld2(hz, ojm);
    p.ex = ldn2(ex, o);
    p.ey = ldn2(ey, o);
    p.ez = ldn2(ez, o);
    p.qh = ldq<T>(qp.hx, o);
    p.qh_jm = ldq<T>(qp.hx, ojm);
    p.qe = ldq<T>(qp.ex, o);
    {
        // H at i-1 is the previous lane's high half (lane_prev_u in
        // ep_update; block-scaled storage shifts the block's exponent word
        // with it); lane 0 fills from the pair before the tile, one address
        // for the wave
        const u32 ou = (u32)__builtin_amdgcn_readfirstlane((int)o) - 2u * S;
        p.hy_fill = ld2(hy, ou);
        p.hz_fill = ld2(hz, ou);
        p.qh_fill = ldq<T>(qp.hx, ou);
    }
    p.cla = p.clb = 0;
    if (CM != CLS_UNIFORM && k + 1 < k1)
        class_pair<CM>(a.cls, (u32)(k + 1) * (u32)a.plane + (u32)r.col, p.cla, p.clb);
    const bool khi = !(V & V_Z) || k < a.nz - 1;
    // Slab psi, issued on every path of the variants with that axis's
    // predicates: lanes off the slab read element 0 of the array, which
    // e_cell does not use, so that the waits for a plane count the same
    // younger loads on every path (loaded only on the slab lanes, the waits
    // sized for the path without them also drained part of the next plane's
    // loads)
    for (int q = 0; q < 2; ++q) {
        p.p_yx[q] = p.p_zx[q] = p.p_xy[q] = p.p_zy[q] = p.p_xz[q] = p.p_yz[q] = 0.f;
        if (V & V_X) {
            const bool n = c[q].xs && r.jhi && khi;
            const u32 qq = n ? (((u32)k * ra.ny + r.j) * ra.Lx + c[q].sx) * S : 0u;
            p.p_yx[q] = ld(at((const PT *)ra.psi_eyx, qq));
            p.p_zx[q] = ld(at((const PT *)ra.psi_ezx, qq));
        }
    }
    if (V & V_Y) {
        const u32 qq = (r.ys && khi) ? (((u32)k * ra.Ly + r.sy) * ra.px + c[0].i) * S : 0u;
        const u32 u = ld2((const PT *)ra.psi_exy, qq), v = ld2((const PT *)ra.psi_ezy, qq);
        p.p_xy[0] = lo_f(u, (const PT *)nullptr);
        p.p_xy[1] = hi_f(u, (const PT *)nullptr);
        p.p_zy[0] = lo_f(v, (const PT *)nullptr);
        p.p_zy[1] = hi_f(v, (const PT *)nullptr);
    }
    if (V & V_Z) {
        const bool n = ((k > 0 && k < a.L[4]) || (k >= a.nz - a.L[5] && khi)) && r.jhi;
        const int sz = k < a.L[4] ? k : k - (a.nz - a.L[5]) + a.L[4];
        const u32 qq = n ? (((u32)sz * ra.ny + r.j) * ra.px + c[0].i) * S : 0u;
        const u32 u = ld2((const PT *)ra.psi_exz, qq), v = ld2((const PT *)ra.psi_eyz, qq);
        p.p_xz[0] = lo_f(u, (const PT *)nullptr);
        p.p_xz[1] = hi_f(u, (const PT *)nullptr);
        p.p_yz[0] = lo_f(v, (const PT *)nullptr);
        p.p_yz[1] = hi_f(v, (const PT *)nullptr);
    }
}

// E update of one cell of the pair; the arithmetic of kernels.hip e_update.
template <class T, int V>
__device__ __forceinline__ void e_cell(const KArgs &a, const PRow &r, const PCell &c, int q, const EPair<T> &p,
                                       const Coef &t, int k, int m, float hxc, float hyc, float hzc, float hy_im,
                                       float hz_im, float hx_jm, float hz_jm, float ex0, float ey0, float ez0,
                                       float hx_km, float hy_km, int &ev, int &ek, bool zs, const float4 &pze,
                                       float &exn, float &eyn, float &ezn, bool &ux, bool &uy, bool &uz)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    const int nz = a.nz;
    const float4 cx = t.x, cy = t.y, cz = t.z;
    const bool klo = !(V & V_Z) || k > 0, khi = !(V & V_Z) || k < nz - 1;
    ux = c.valid && c.ihi && r.jlo && r.jhi && klo && khi;
    uy = c.valid && c.ilo && c.ihi && r.jhi && klo && khi;
    uz = c.valid && c.ilo && c.ihi && r.jlo && r.jhi && khi;

    float cbx_hz = cx.y, cbx_hy = cx.z;
    float cby_hx = cy.y, cby_hz = cy.z;
    float cbz_hy = cz.y, cbz_hx = cz.z;
    if (c.xs) { cby_hz = cby_hz / c.px.z; cbz_hy = cbz_hy / c.px.z; }
    if (r.ys) { cbx_hz = cbx_hz / r.py.z; cbz_hx = cbz_hx / r.py.z; }
    if (zs) { cbx_hy = cbx_hy / pze.z; cby_hx = cby_hx / pze.z; }

    exn = ex0;
    eyn = ey0;
    ezn = ez0;
    if (ux) exn = (cx.x * ex0 + cbx_hz * (hzc - hz_jm)) + cbx_hy * (hyc - hy_km);
    if (uy) eyn = (cy.x * ey0 + cby_hx * (hxc - hx_km)) + cby_hz * (hzc - hz_im);
    if (uz) ezn = (cz.x * ez0 + cbz_hy * (hyc - hy_im)) + cbz_hx * (hxc - hx_jm);

    if (k == ek) {   // source edges and probes on this cell
        const EvE rr = e_events(ra, ev, k, c.mycol, m, uz, ez0, hyc, hy_im, hxc, hx_jm, exn, eyn, ezn);
        ev = rr.ev;
        ek = rr.ek;
        ezn = rr.ezn;
    }

    constexpr u32 S = sizeof(T);
    const float dx = ra.d[0], dy = ra.d[1], dz = ra.d[2];
    if (c.xs && r.jhi && khi) {
        const u32 qq = (((u32)k * ra.ny + r.j) * ra.Lx + c.sx) * S;
        const float pey = c.px.x * p.p_yx[q] + c.px.y * (hzc - hz_im);
        const float pez = c.px.x * p.p_zx[q] + c.px.y * (hyc - hy_im);
        st(at((PT *)ra.psi_eyx, qq), pey);
        st(at((PT *)ra.psi_ezx, qq), pez);
        if (klo) eyn = eyn + (cy.z * dx) * pey;
        if (r.jlo) ezn = ezn + (cz.y * dx) * pez;
    }
    if (r.ys && c.valid && c.ihi && khi) {
        const u32 qq = (((u32)k * ra.Ly + r.sy) * ra.px + c.i) * S;
        const float pex = r.py.x * p.p_xy[q] + r.py.y * (hzc - hz_jm);
        const float pez = r.py.x * p.p_zy[q] + r.py.y * (hxc - hx_jm);
        st(at((PT *)ra.psi_exy, qq), pex);
        st(at((PT *)ra.psi_ezy, qq), pez);
        if (klo) exn = exn + (cx.y * dy) * pex;
        if (c.ilo) ezn = ezn + (cz.z * dy) * pez;
    }
    if (zs && c.valid && c.ihi && r.jhi) {
        const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
        const u32 qq = (((u32)sz * ra.ny + r.j) * ra.px + c.i) * S;
        const float pex = pze.x * p.p_xz[q] + pze.y * (hyc - hy_km);
        const float pey = pze.x * p.p_yz[q] + pze.y * (hxc - hx_km);
        st(at((PT *)ra.psi_exz, qq), pex);
        st(at((PT *)ra.psi_eyz, qq), pey);
        if (r.jlo) exn = exn + (cx.z * dz) * pex;
        if (c.ilo) eyn = eyn + (cy.y * dz) * pey;
    }
    if (c.in_snap && k >= a.sb[2] && k < a.sb[2] + a.sb[5]) {
        const T *tp = nullptr;
        e_snapshot(ra, c.i, r.j, k, m, ux ? rnd(exn, tp) : ex0, uy ? rnd(eyn, tp) : ey0, uz ? rnd(ezn, tp) : ez0);
    }
}

template <class T, int CM, int V>
__device__ __forceinline__ void ep_update(const EPair<T> &p, const CoefPair &t, const KArgs &a, const PRow &r,
                                          const PCell (&c)[2], T *ex, T *ey, T *ez, const QPtrs &qp, int k, int m,
                                          u32 &hx_km, u32 &hy_km, QW<T> &q_km, int (&ev)[2], int (&ek)[2])
{
    const u32 o = r.o0 + (u32)k * r.ob;
    const bool zs = (V & V_Z) && ((k > 0 && k < a.L[4]) || (k >= a.nz - a.L[5] && k < a.nz - 1));   // fdtdrav.h:853,879
    const float4 pze = zs ? uload(a.pz_prof + k) : make_float4(0.f, 0.f, 1.f, 0.f);
    const float hx0 = dlo(p.hx, p.qh, 0), hx1 = dhi(p.hx, p.qh, 0);
    const float hy0 = dlo(p.hy, p.qh, 1), hy1 = dhi(p.hy, p.qh, 1);
    const float hz0 = dlo(p.hz, p.qh, 2), hz1 = dhi(p.hz, p.qh, 2);
    const float ex0 = dlo(p.ex, p.qe, 0), ex1 = dhi(p.ex, p.qe, 0);
    const float ey0 = dlo(p.ey, p.qe, 1), ey1 = dhi(p.ey, p.qe, 1);
    const float ez0 = dlo(p.ez, p.qe, 2), ez1 = dhi(p.ez, p.qe, 2);
    float ex_a, ey_a, ez_a, ex_b, ey_b, ez_b;
    bool uxa, uya, uza, uxb, uyb, uzb;
    const QW<T> qim = qw_prev(p.qh, p.qh_fill);
    const float hy_im = dhi(lane_prev_u(p.hy, p.hy_fill), qim, 1);
    const float hz_im = dhi(lane_prev_u(p.hz, p.hz_fill), qim, 2);
    e_cell<T, V>(a, r, c[0], 0, p, t.a, k, m, hx0, hy0, hz0, hy_im, hz_im, dlo(p.hx_jm, p.qh_jm, 0),
                 dlo(p.hz_jm, p.qh_jm, 2), ex0, ey0, ez0, dlo(hx_km, q_km, 0), dlo(hy_km, q_km, 1), ev[0], ek[0], zs,
                 pze, ex_a, ey_a, ez_a, uxa, uya, uza);
    e_cell<T, V>(a, r, c[1], 1, p, t.b, k, m, hx1, hy1, hz1, hy0, hz0, dhi(p.hx_jm, p.qh_jm, 0),
                 dhi(p.hz_jm, p.qh_jm, 2), ex1, ey1, ez1, dhi(hx_km, q_km, 0), dhi(hy_km, q_km, 1), ev[1], ek[1], zs,
                 pze, ex_b, ey_b, ez_b, uxb, uyb, uzb);
    const bool va = c[0].valid, vb = c[1].valid;
    const u32 bx = stp(ex, o, p.ex, ex0, ex1, uxa, ex_a, uxb, ex_b, va, vb);
    const u32 by = stp(ey, o, p.ey, ey0, ey1, uya, ey_a, uyb, ey_b, va, vb);
    const u32 bz = stp(ez, o, p.ez, ez0, ez1, uza, ez_a, uzb, ez_b, va, vb);
    stq<T>(qp.ex, o, bx, by, bz, va && (c[0].i & (BQ_BLOCK - 1)) == 0);
    hx_km = carry(p.hx);
    hy_km = carry(p.hy);
    q_km = carryq(p.qh);
}

template <class T, int CM, int V>
__global__ void __launch_bounds__(TX * BY) P_OCC e_pair_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = tx * TXP + 2 * (int)(threadIdx.x & (TX - 1));
    PRow r;
    r.j = ty * BY + (int)(threadIdx.x / TX);
    // block-scaled storage encodes 8-cell blocks across the four lanes of a
    // quad: the lanes of a block that reaches past nx stay (the block lies
    // inside the row's padded pitch), storing nothing past nx
    if (r.j >= ny || (is_bq<T>::value ? (i0 & ~(BQ_BLOCK - 1)) : i0) >= nx) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const int j = r.j;
    r.col = (i64)j * a.px + i0;
    r.ob = (u32)a.plane * (u32)sizeof(T);
    r.o0 = (u32)r.col * (u32)sizeof(T) + r.ob;
    r.jlo = !(V & V_Y) || j > 0;
    r.jhi = !(V & V_Y) || j < ny - 1;
    r.ys = (V & V_Y) && ((r.jlo && j < a.L[2]) || (j >= ny - a.L[3] && r.jhi));   // E y-slab, fdtdrav.h:803,827
    r.sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    r.py = r.ys ? a.py_prof[j] : make_float4(0.f, 0.f, 1.f, 0.f);

    PCell c[2];
    for (int q = 0; q < 2; ++q) {
        const int i = i0 + q;
        c[q].i = i;
        c[q].valid = !(V & V_X) || i < nx;
        c[q].ilo = !(V & V_X) || i > 0;
        c[q].ihi = !(V & V_X) || i < nx - 1;
        c[q].xs = (V & V_X) && c[q].valid && ((c[q].ilo && i < a.L[0]) || (i >= nx - a.L[1] && c[q].ihi));   // fdtdrav.h:752,777
        c[q].sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
        c[q].px = c[q].xs ? a.px_prof[i] : make_float4(0.f, 0.f, 1.f, 0.f);
        c[q].mycol = j * nx + i;
        c[q].in_snap = c[q].valid && a.snap != nullptr && i >= a.sb[0] && i < a.sb[0] + a.sb[3] &&
                       j >= a.sb[1] && j < a.sb[1] + a.sb[4];
    }

    T *__restrict__ ex = (T *)a.ex - a.plane;
    T *__restrict__ ey = (T *)a.ey - a.plane;
    T *__restrict__ ez = (T *)a.ez - a.plane;
    const T *__restrict__ hx = (const T *)a.hx - a.plane;
    const T *__restrict__ hy = (const T *)a.hy - a.plane;
    const T *__restrict__ hz = (const T *)a.hz - a.plane;
    const QPtrs qp = qptrs(a);

    int ev[2], ek[2];
    for (int q = 0; q < 2; ++q) {
        ev[q] = c[q].valid ? first_event(a, a.e_first, c[q].mycol, k0) : -1;
        ek[q] = first_k(a, ev[q], c[q].mycol);
    }

    u32 hx_km = ld2(hx, r.o0 + (u32)(k0 - 1) * r.ob);
    u32 hy_km = ld2(hy, r.o0 + (u32)(k0 - 1) * r.ob);
    QW<T> q_km = ldq<T>(qp.hx, r.o0 + (u32)(k0 - 1) * r.ob);

    EPair<T> A, B;
    CoefPair tA, tB;
    if (CM == CLS_UNIFORM) {
        tA = coef_pair<CM>(a.etab, 0, 0);
        tB = tA;
    } else {
        int ca, cb;
        class_pair<CM>(a.cls, (u32)k0 * (u32)a.plane + (u32)r.col, ca, cb);
        tA = coef_pair<CM>(a.etab, ca, cb);
    }
    ep_fetch<T, CM, V>(A, a, r, c, ex, ey, ez, hx, hy, hz, qp, k0, k1);
    int k = k0;
    for (; k + 2 < k1; k += 2) {
        ep_fetch<T, CM, V>(B, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_pair<CM>(a.etab, A.cla, A.clb);
        ep_update<T, CM, V>(A, tA, a, r, c, ex, ey, ez, qp, k, m, hx_km, hy_km, q_km, ev, ek);
        ep_fetch<T, CM, V>(A, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 2, k1);
        if (CM != CLS_UNIFORM) tA = coef_pair<CM>(a.etab, B.cla, B.clb);
        ep_update<T, CM, V>(B, tB, a, r, c, ex, ey, ez, qp, k + 1, m, hx_km, hy_km, q_km, ev, ek);
    }
    if (k + 1 < k1) {
        ep_fetch<T, CM, V>(B, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_pair<CM>(a.etab, A.cla, A.clb);
        ep_update<T, CM, V>(A, tA, a, r, c, ex, ey, ez, qp, k, m, hx_km, hy_km, q_km, ev, ek);
        ep_update<T, CM, V>(B, tB, a, r, c, ex, ey, ez, qp, k + 1, m, hx_km, hy_km, q_km, ev, ek);
    } else {
        ep_update<T, CM, V>(A, tA, a, r, c, ex, ey, ez, qp, k, m, hx_km, hy_km, q_km, ev, ek);
    }
}

// ============================ H sweep ======================================

template <class T>
struct HPair {
    u32 ex_kp, ey_kp, ez, ez_jp, ex_jp, hx, hy, hz;   // packed pairs
    QW<T> qe_kp, qe_jp, qh;                           // exponent words of E at k+1, E at row j+1, H
    float ez_i2, ey_i2;                               // E at i+2 (x neighbour of cell 1; loaded)
    u32 ez_fill, ey_fill;                             // E pair after the tile (lane 63's i+2, i+3; shifted)
    QW<T> qe_fill;                                    // its exponent word
    float p_yx[2], p_zx[2], p_xy[2], p_zy[2], p_xz[2], p_yz[2];
    int cla, clb;
};

// E at i+2 from the next lane (lane_next_u) or loaded: the x-edge variant
// loads it (there the shift's two registers cost a wave per SIMD: 116.6
// against 113.9 us on 512x512x256 FP16).
template <class T, int V>
struct HpShift {
    static constexpr bool value = V != V_X;
};

__device__ __forceinline__ bool hp_zslab(const KArgs &a, int k)
{
    return k < a.L[4] || (k >= a.nz - a.L[5] && k < a.nz - 1);   // fdtdrav.h:500,531
}

template <class T, int CM, int V>
__device__ __forceinline__ void hp_fetch(HPair<T> &p, const KArgs &a, const PRow &r, const PCell (&c)[2],
                                         const T *ex, const T *ey, const T *ez, const T *hx, const T *hy,
                                         const T *hz, const QPtrs &qp, int k, int k1)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    constexpr u32 S = sizeof(T);
    const u32 o = r.o0 + (u32)k * r.ob;
    const u32 okp = o + r.ob;
    const u32 ojp = o + (u32)a.px * S;
    p.ex_kp = ld2(ex, okp);
    p.ey_kp = ld2(ey, okp);
    p.ez = ld2(ez, o);
    p.ez_jp = ld2(ez, ojp);
    p.ex_jp = ld2(ex, ojp);
    p.hx = ldn2(hx, o);
    p.hy = ldn2(hy, o);
    p.hz = ldn2(hz, o);
    p.qe_kp = ldq<T>(qp.ex, okp);
    p.qe_jp = ldq<T>(qp.ex, ojp);
    p.qh = ldq<T>(qp.hx, o);
    if (!HpShift<T, V>::value) {
        const QW<T> qe_i2 = ldq<T>(qp.ex, o + 2 * S);
        p.ez_i2 = fldw(ez, o + 2 * S, qe_i2, 2);
        p.ey_i2 = fldw(ey, o + 2 * S, qe_i2, 1);
    } else {
        // E at i+2 is the next lane's low half (lane_next_u in hp_update);
        // lane 63 fills from the pair after the tile, one address for the wave
        const u32 ou = (u32)__builtin_amdgcn_readfirstlane((int)o) + (u32)TXP * S;
        p.ez_fill = ld2(ez, ou);
        p.ey_fill = ld2(ey, ou);
        p.qe_fill = ldq<T>(qp.ex, ou);
    }
    p.cla = p.clb = 0;
    if (CM != CLS_UNIFORM && k + 1 < k1)
        class_pair<CM>(a.cls, (u32)(k + 1) * (u32)a.plane + (u32)r.col, p.cla, p.clb);
    // slab psi on every path of the variants with that axis's predicates
    // (as in ep_fetch)
    for (int q = 0; q < 2; ++q) {
        p.p_yx[q] = p.p_zx[q] = p.p_xy[q] = p.p_zy[q] = p.p_xz[q] = p.p_yz[q] = 0.f;
        if (V & V_X) {
            const u32 qq = c[q].xs ? (((u32)k * ra.ny + r.j) * ra.Lx + c[q].sx) * S : 0u;
            p.p_yx[q] = ld(at((const PT *)ra.psi_hyx, qq));
            p.p_zx[q] = ld(at((const PT *)ra.psi_hzx, qq));
        }
    }
    if (V & V_Y) {
        const u32 qq = r.ys ? (((u32)k * ra.Ly + r.sy) * ra.px + c[0].i) * S : 0u;
        const u32 u = ld2((const PT *)ra.psi_hxy, qq), v = ld2((const PT *)ra.psi_hzy, qq);
        p.p_xy[0] = lo_f(u, (const PT *)nullptr);
        p.p_xy[1] = hi_f(u, (const PT *)nullptr);
        p.p_zy[0] = lo_f(v, (const PT *)nullptr);
        p.p_zy[1] = hi_f(v, (const PT *)nullptr);
    }
    if (V & V_Z) {
        const int sz = k < a.L[4] ? k : k - (a.nz - a.L[5]) + a.L[4];
        const u32 qq = hp_zslab(a, k) ? (((u32)sz * ra.ny + r.j) * ra.px + c[0].i) * S : 0u;
        const u32 u = ld2((const PT *)ra.psi_hxz, qq), v = ld2((const PT *)ra.psi_hyz, qq);
        p.p_xz[0] = lo_f(u, (const PT *)nullptr);
        p.p_xz[1] = hi_f(u, (const PT *)nullptr);
        p.p_yz[0] = lo_f(v, (const PT *)nullptr);
        p.p_yz[1] = hi_f(v, (const PT *)nullptr);
    }
}

// H update of one cell of the pair; the arithmetic of kernels.hip h_update.
template <class T, int V>
__device__ __forceinline__ void h_cell(const KArgs &a, const PRow &r, const PCell &c, int q, const HPair<T> &p,
                                       const Coef &t, int k, int m, float ex_kp, float ey_kp, float ezc,
                                       float ez_ip, float ey_ip, float ez_jp, float ex_jp, float hx0, float hy0,
                                       float hz0, float ex_c, float ey_c, int &ev, int &ek, bool zs,
                                       const float4 &pzm, float &hxn, float &hyn, float &hzn, bool &ux, bool &uy,
                                       bool &uz)
{
    typedef typename psi_type<T>::type PT;
    ConstKArgs &ra = rare_args(a);
    const int nz = a.nz;
    const float4 cx = t.x, cy = t.y, cz = t.z;
    const bool khi = !(V & V_Z) || k < nz - 1;
    ux = c.valid && r.jhi && khi;
    uy = c.valid && c.ihi && khi;
    uz = c.valid && c.ihi && r.jhi;

    float chx_ey = cx.y, chx_ez = cx.z;
    float chy_ez = cy.y, chy_ex = cy.z;
    float chz_ex = cz.y, chz_ey = cz.z;
    if (c.xs) { chy_ez = chy_ez / c.px.z; chz_ey = chz_ey / c.px.z; }
    if (r.ys) { chx_ez = chx_ez / r.py.z; chz_ex = chz_ex / r.py.z; }
    if (zs) { chx_ey = chx_ey / pzm.z; chy_ex = chy_ex / pzm.z; }

    hxn = hx0;
    hyn = hy0;
    hzn = hz0;
    if (ux) hxn = (cx.x * hx0 + chx_ey * (ey_kp - ey_c)) + chx_ez * (ez_jp - ezc);
    if (uy) hyn = (cy.x * hy0 + chy_ez * (ez_ip - ezc)) + chy_ex * (ex_kp - ex_c);
    if (uz) hzn = (cz.x * hz0 + chz_ex * (ex_jp - ex_c)) + chz_ey * (ey_ip - ey_c);

    if (k == ek) {   // probes on this cell
        const EvH rr = h_events(ra, ev, k, c.mycol, m, hxn, hyn, hzn);
        ev = rr.ev;
        ek = rr.ek;
    }

    constexpr u32 S = sizeof(T);
    const float dx = ra.d[0], dy = ra.d[1], dz = ra.d[2];
    if (c.xs) {
        const u32 qq = (((u32)k * ra.ny + r.j) * ra.Lx + c.sx) * S;
        const float phy = c.px.x * p.p_yx[q] + c.px.y * (ez_ip - ezc);
        const float phz = c.px.x * p.p_zx[q] + c.px.y * (ey_ip - ey_c);
        st(at((PT *)ra.psi_hyx, qq), phy);
        st(at((PT *)ra.psi_hzx, qq), phz);
        if (khi) hyn = hyn + (cy.y * dx) * phy;
        if (r.jhi) hzn = hzn + (cz.z * dx) * phz;
    }
    if (r.ys && c.valid) {
        const u32 qq = (((u32)k * ra.Ly + r.sy) * ra.px + c.i) * S;
        const float phx = r.py.x * p.p_xy[q] + r.py.y * (ez_jp - ezc);
        const float phz = r.py.x * p.p_zy[q] + r.py.y * (ex_jp - ex_c);
        st(at((PT *)ra.psi_hxy, qq), phx);
        st(at((PT *)ra.psi_hzy, qq), phz);
        if (khi) hxn = hxn + (cx.z * dy) * phx;
        if (c.ihi) hzn = hzn + (cz.y * dy) * phz;
    }
    if (zs && c.valid) {
        const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
        const u32 qq = (((u32)sz * ra.ny + r.j) * ra.px + c.i) * S;
        const float phx = pzm.x * p.p_xz[q] + pzm.y * (ey_kp - ey_c);
        const float phy = pzm.x * p.p_yz[q] + pzm.y * (ex_kp - ex_c);
        st(at((PT *)ra.psi_hxz, qq), phx);
        st(at((PT *)ra.psi_hyz, qq), phy);
        if (r.jhi) hxn = hxn + (cx.y * dz) * phx;
        if (c.ihi) hyn = hyn + (cy.z * dz) * phy;
    }
}

template <class T, int CM, int V>
__device__ __forceinline__ void hp_update(const HPair<T> &p, const CoefPair &t, const KArgs &a, const PRow &r,
                                          const PCell (&c)[2], T *hx, T *hy, T *hz, const QPtrs &qp, int k, int m,
                                          u32 &ex_c, u32 &ey_c, QW<T> &q_c, int (&ev)[2], int (&ek)[2])
{
    const u32 o = r.o0 + (u32)k * r.ob;
    const bool zs = (V & V_Z) && hp_zslab(a, k);
    const float4 pzm = zs ? uload(a.pz_prof + a.nz + k) : make_float4(0.f, 0.f, 1.f, 0.f);
    // q_c: the exponent word of E at plane k (carried from the previous plane)
    const float ez0 = dlo(p.ez, q_c, 2), ez1 = dhi(p.ez, q_c, 2);
    const float exc0 = dlo(ex_c, q_c, 0), exc1 = dhi(ex_c, q_c, 0);
    const float eyc0 = dlo(ey_c, q_c, 1), eyc1 = dhi(ey_c, q_c, 1);
    const float hx0 = dlo(p.hx, p.qh, 0), hx1 = dhi(p.hx, p.qh, 0);
    const float hy0 = dlo(p.hy, p.qh, 1), hy1 = dhi(p.hy, p.qh, 1);
    const float hz0 = dlo(p.hz, p.qh, 2), hz1 = dhi(p.hz, p.qh, 2);
    float hx_a, hy_a, hz_a, hx_b, hy_b, hz_b;
    bool uxa, uya, uza, uxb, uyb, uzb;
    h_cell<T, V>(a, r, c[0], 0, p, t.a, k, m, dlo(p.ex_kp, p.qe_kp, 0), dlo(p.ey_kp, p.qe_kp, 1), ez0, ez1, eyc1,
                 dlo(p.ez_jp, p.qe_jp, 2), dlo(p.ex_jp, p.qe_jp, 0), hx0, hy0, hz0, exc0, eyc0, ev[0], ek[0], zs, pzm,
                 hx_a, hy_a, hz_a, uxa, uya, uza);
    const QW<T> qi2 = qw_next(q_c, p.qe_fill);
    const float ez_i2 = !HpShift<T, V>::value ? p.ez_i2 : dlo(lane_next_u(p.ez, p.ez_fill), qi2, 2);
    const float ey_i2 = !HpShift<T, V>::value ? p.ey_i2 : dlo(lane_next_u(ey_c, p.ey_fill), qi2, 1);
    h_cell<T, V>(a, r, c[1], 1, p, t.b, k, m, dhi(p.ex_kp, p.qe_kp, 0), dhi(p.ey_kp, p.qe_kp, 1), ez1, ez_i2,
                 ey_i2, dhi(p.ez_jp, p.qe_jp, 2), dhi(p.ex_jp, p.qe_jp, 0), hx1, hy1, hz1, exc1, eyc1, ev[1], ek[1],
                 zs, pzm, hx_b, hy_b, hz_b, uxb, uyb, uzb);
    const bool va = c[0].valid, vb = c[1].valid;
    const u32 bx = stp(hx, o, p.hx, hx0, hx1, uxa, hx_a, uxb, hx_b, va, vb);
    const u32 by = stp(hy, o, p.hy, hy0, hy1, uya, hy_a, uyb, hy_b, va, vb);
    const u32 bz = stp(hz, o, p.hz, hz0, hz1, uza, hz_a, uzb, hz_b, va, vb);
    stq<T>(qp.hx, o, bx, by, bz, va && (c[0].i & (BQ_BLOCK - 1)) == 0);
    ex_c = carry(p.ex_kp);
    ey_c = carry(p.ey_kp);
    q_c = carryq(p.qe_kp);
}

template <class T, int CM, int V>
__global__ void __launch_bounds__(TX * BY) P_OCC h_pair_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of_region<V>(a, tx, ty, tz);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int i0 = tx * TXP + 2 * (int)(threadIdx.x & (TX - 1));
    PRow r;
    r.j = ty * BY + (int)(threadIdx.x / TX);
    // block-scaled storage encodes 8-cell blocks across the four lanes of a
    // quad: the lanes of a block that reaches past nx stay (the block lies
    // inside the row's padded pitch), storing nothing past nx
    if (r.j >= ny || (is_bq<T>::value ? (i0 & ~(BQ_BLOCK - 1)) : i0) >= nx) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const int j = r.j;
    r.col = (i64)j * a.px + i0;
    r.ob = (u32)a.plane * (u32)sizeof(T);
    r.o0 = (u32)r.col * (u32)sizeof(T) + r.ob;
    r.jlo = !(V & V_Y) || j > 0;
    r.jhi = !(V & V_Y) || j < ny - 1;
    r.ys = (V & V_Y) && (j < a.L[2] || (j >= ny - a.L[3] && r.jhi));   // H y-slab, fdtdrav.h:447,471
    r.sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    r.py = r.ys ? a.py_prof[ny + j] : make_float4(0.f, 0.f, 1.f, 0.f);

    PCell c[2];
    for (int q = 0; q < 2; ++q) {
        const int i = i0 + q;
        c[q].i = i;
        c[q].valid = !(V & V_X) || i < nx;
        c[q].ilo = !(V & V_X) || i > 0;
        c[q].ihi = !(V & V_X) || i < nx - 1;
        c[q].xs = (V & V_X) && c[q].valid && (i < a.L[0] || (i >= nx - a.L[1] && c[q].ihi));   // fdtdrav.h:402,424
        c[q].sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
        c[q].px = c[q].xs ? a.px_prof[nx + i] : make_float4(0.f, 0.f, 1.f, 0.f);
        c[q].mycol = j * nx + i;
        c[q].in_snap = false;
    }

    const T *__restrict__ ex = (const T *)a.ex - a.plane;
    const T *__restrict__ ey = (const T *)a.ey - a.plane;
    const T *__restrict__ ez = (const T *)a.ez - a.plane;
    T *__restrict__ hx = (T *)a.hx - a.plane;
    T *__restrict__ hy = (T *)a.hy - a.plane;
    T *__restrict__ hz = (T *)a.hz - a.plane;
    const QPtrs qp = qptrs(a);

    int ev[2], ek[2];
    for (int q = 0; q < 2; ++q) {
        ev[q] = c[q].valid ? first_event(a, a.h_first, c[q].mycol, k0) : -1;
        ek[q] = first_k(a, ev[q], c[q].mycol);
    }

    u32 ex_c = ld2(ex, r.o0 + (u32)k0 * r.ob);
    u32 ey_c = ld2(ey, r.o0 + (u32)k0 * r.ob);
    QW<T> q_c = ldq<T>(qp.ex, r.o0 + (u32)k0 * r.ob);

    HPair<T> A, B;
    CoefPair tA, tB;
    if (CM == CLS_UNIFORM) {
        tA = coef_pair<CM>(a.htab, 0, 0);
        tB = tA;
    } else {
        int ca, cb;
        class_pair<CM>(a.cls, (u32)k0 * (u32)a.plane + (u32)r.col, ca, cb);
        tA = coef_pair<CM>(a.htab, ca, cb);
    }
    hp_fetch<T, CM, V>(A, a, r, c, ex, ey, ez, hx, hy, hz, qp, k0, k1);
    int k = k0;
    for (; k + 2 < k1; k += 2) {
        hp_fetch<T, CM, V>(B, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_pair<CM>(a.htab, A.cla, A.clb);
        hp_update<T, CM, V>(A, tA, a, r, c, hx, hy, hz, qp, k, m, ex_c, ey_c, q_c, ev, ek);
        hp_fetch<T, CM, V>(A, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 2, k1);
        if (CM != CLS_UNIFORM) tA = coef_pair<CM>(a.htab, B.cla, B.clb);
        hp_update<T, CM, V>(B, tB, a, r, c, hx, hy, hz, qp, k + 1, m, ex_c, ey_c, q_c, ev, ek);
    }
    if (k + 1 < k1) {
        hp_fetch<T, CM, V>(B, a, r, c, ex, ey, ez, hx, hy, hz, qp, k + 1, k1);
        if (CM != CLS_UNIFORM) tB = coef_pair<CM>(a.htab, A.cla, A.clb);
        hp_update<T, CM, V>(A, tA, a, r, c, hx, hy, hz, qp, k, m, ex_c, ey_c, q_c, ev, ek);
        hp_update<T, CM, V>(B, tB, a, r, c, hx, hy, hz, qp, k + 1, m, ex_c, ey_c, q_c, ev, ek);
    } else {
        hp_update<T, CM, V>(A, tA, a, r, c, hx, hy, hz, qp, k, m, ex_c, ey_c, q_c, ev, ek);
    }
}

// ---- launchers --------------------------------------------------------------
template <class T, int V>
static void launch_pair_v(bool efield, int mode, const KArgs &a, int m, int nt, hipStream_t s)
{
    dim3 grid(nt), block(TX * BY);
    if (efield) {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((e_pair_kernel<T, CLS_UNIFORM, V>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((e_pair_kernel<T, CLS_U16, V>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((e_pair_kernel<T, CLS_U32, V>), grid, block, 0, s, a, m);
    } else {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((h_pair_kernel<T, CLS_UNIFORM, V>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((h_pair_kernel<T, CLS_U16, V>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((h_pair_kernel<T, CLS_U32, V>), grid, block, 0, s, a, m);
    }
}

template <class T>
static hipError_t launch_pair_t(bool efield, int mode, const KArgs &a, int m, hipStream_t s, bool gen_cells)
{
    int nb, nx, nzt, ng;   // nzt = 0: the forward sweeps run without zsplit
    region_tiles(a, nb, nx, nzt, ng);
    if (nb > 0) launch_pair_v<T, V_IN>(efield, mode, a, m, nb, s);
    if (nx > 0) launch_pair_v<T, V_X>(efield, mode, a, m, nx, s);
    if (ng > 0 && !gen_cells) launch_pair_v<T, V_GEN>(efield, mode, a, m, ng, s);
    if (ng > 0 && gen_cells) {
        // the same rows and z-blocks in one-cell tiles (TX columns): the
        // general region depends on the box's y and z range only
        KArgs c = a;
        c.tiles_x = (a.nx + TX - 1) / TX;
        c.ib[0] = a.ib[0] * (TXP / TX);
        c.ib[3] = a.ib[3] * (TXP / TX);
        if (c.sub[3] > 0) {   // light-cone range in one-cell tile columns
            c.sub[0] = a.sub[0] * (TXP / TX);
            c.sub[3] = min(a.sub[3] * (TXP / TX), c.tiles_x - c.sub[0]);
        }
        const int st = std::is_same<T, bf16_t>::value ? 2 : is_bq<T>::value ? 3 : 1;
        const hipError_t e = launch_half_step_general(efield, st, mode, c, m, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

// storage: 1 FP16, 2 BF16, 3 block-scaled int16.  a.tiles_x counts
// 128-cell tile columns.  gen_cells: the general variant runs as the
// one-cell kernel.
hipError_t launch_pair_half_step(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s,
                                 bool gen_cells)
{
    switch (storage) {
    case 1: return launch_pair_t<_Float16>(efield, mode, a, m, s, gen_cells);
    case 2: return launch_pair_t<bf16_t>(efield, mode, a, m, s, gen_cells);
    case 3: return launch_pair_t<bq16_t>(efield, mode, a, m, s, gen_cells);
    default: return hipErrorInvalidValue;
    }
}

int pair_tile_cols() { return TXP; }

}  // namespace fdtd
