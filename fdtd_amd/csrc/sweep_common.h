// sweep_common.h -- device helpers shared by the E/H sweep kernels
// (kernels.hip: one cell per thread; kernels_pair.hip: two cells per thread
// for 16-bit storage).
#pragma once
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "device_common.h"
#include "internal.h"

namespace fdtd {

// Copy of a carried z-operand (the previous plane's H for E, the next
// plane's E for H) into a register of its own.  Left to itself the compiler
// keeps the carried value in its load register and renames the next fetch
// instead, which costs a register copy -- and a wait for that fetch -- at the
// loop back-edge.  The source is already consumed when this runs.
__device__ __forceinline__ float carry(float v)
{
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ std::uint32_t carry(std::uint32_t v)
{
    std::uint32_t r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// Non-temporal access for the streams only the owning thread touches (the
// swept field's own values): they are read once and written once per sweep.
__device__ __forceinline__ float ldn(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stn(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ float ldn(const _Float16 *p) { return (float)__builtin_nontemporal_load(p); }
__device__ __forceinline__ void stn(_Float16 *p, float v) { __builtin_nontemporal_store((_Float16)v, p); }
__device__ __forceinline__ float ldn(const bf16_t *p)
{
    const std::uint16_t u = __builtin_nontemporal_load((const std::uint16_t *)p);
    return __uint_as_float((std::uint32_t)u << 16);
}
__device__ __forceinline__ void stn(bf16_t *p, float v)
{
    std::uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    __builtin_nontemporal_store((std::uint16_t)(u >> 16), (std::uint16_t *)p);
}

// ---- block-scaled 16-bit fields (bq16_t) ----------------------------------
// Exponent bytes of the six field components, from the start of the
// (guard-plane) allocation: ex = the E exponent array, ey = ex + 1, ez =
// ex + 2, likewise for H (null for the other storage types).
struct QPtrs {
    std::uint8_t *ex, *ey, *ez, *hx, *hy, *hz;
};

// The exponent pointers of the kernel arguments (which start at k = 0),
// moved back by the guard plane below k = 0.
__device__ __forceinline__ QPtrs qptrs(const KArgs &a)
{
    const i64 g = a.plane / BQ_BLOCK * BQ_STRIDE;
    return QPtrs{a.qex - g, a.qey - g, a.qez - g, a.qhx - g, a.qhy - g, a.qhz - g};
}

// Byte index of the exponent of the value at byte offset o, from the
// component's QPtrs entry.
__device__ __forceinline__ std::uint32_t bq_at(std::uint32_t o) { return (o / (2 * BQ_BLOCK)) * BQ_STRIDE; }

// Exponent word of block-scaled storage: the exponent bytes of the three
// components of E (or H) at one 8-cell block, one 4-byte load (empty for the
// other storage types, which then carry no extra registers).
template <class T>
struct QW {
};
template <>
struct QW<bq16_t> {
    std::uint32_t w;
};

template <class T>
__device__ __forceinline__ QW<T> ldq(const std::uint8_t *, std::uint32_t)
{
    return QW<T>{};
}
template <>
__device__ __forceinline__ QW<bq16_t> ldq<bq16_t>(const std::uint8_t *q, std::uint32_t off)
{
    QW<bq16_t> r;
    r.w = *(const std::uint32_t *)(q + bq_at(off));
    return r;
}
template <class T>
__device__ __forceinline__ QW<T> carryq(const QW<T> &q)
{
    return q;
}
template <>
__device__ __forceinline__ QW<bq16_t> carryq<bq16_t>(const QW<bq16_t> &q)
{
    QW<bq16_t> r;
    r.w = carry(q.w);
    return r;
}
// exponent of component c (bias removed)
__device__ __forceinline__ int qe(const QW<bq16_t> &q, int c) { return (int)((q.w >> (8 * c)) & 0xFFu) - BQ_BIAS; }

// Value of component c of a field at byte offset o from the start of its
// allocation (w: the exponent word of o's block for block-scaled storage).
template <class T>
__device__ __forceinline__ float fldw(const T *base, std::uint32_t o, const QW<T> &, int)
{
    return ld((const T *)((const char *)base + o));
}
template <class T>
__device__ __forceinline__ float fldnw(const T *base, std::uint32_t o, const QW<T> &, int)
{
    return ldn((const T *)((const char *)base + o));
}
template <>
__device__ __forceinline__ float fldw<bq16_t>(const bq16_t *base, std::uint32_t o, const QW<bq16_t> &w, int c)
{
    return ldexpf((float)*(const std::int16_t *)((const char *)base + o), qe(w, c));
}
template <>
__device__ __forceinline__ float fldnw<bq16_t>(const bq16_t *base, std::uint32_t o, const QW<bq16_t> &w, int c)
{
    return ldexpf((float)__builtin_nontemporal_load((const std::int16_t *)((const char *)base + o)), qe(w, c));
}

// Lane permutations within a 16-lane row (DPP, no memory access): another
// lane of the same quad, or the mirrored lane of the same 8-lane half row.
// A lane whose source lane is inactive keeps its own value.
constexpr int DPP_QUAD_XOR1 = 0xB1;     // quad_perm [1, 0, 3, 2]
constexpr int DPP_QUAD_XOR2 = 0x4E;     // quad_perm [2, 3, 0, 1]
constexpr int DPP_HALF_MIRROR = 0x141;  // row_half_mirror: lane i <- lane 7 - i of its 8
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Block exponent from the block's largest magnitude mx: the smallest e with
// 2^e > mx / 32767 (e = -BQ_BIAS for an all-zero block), and the int16
// mantissa of x under it.
__device__ __forceinline__ int bq_exponent(float mx)
{
    int e = -BQ_BIAS;
    if (mx > 0.f) {
        (void)frexpf(mx * (1.0f / 32767.0f), &e);   // 2^(e-1) <= max/32767 < 2^e
        e = max(e, -BQ_BIAS);
    }
    return e;
}
__device__ __forceinline__ int bq_mantissa(float x, int e)
{
    return (int)fminf(fmaxf(rintf(ldexpf(x, -e)), -32767.f), 32767.f);
}

// Store of a swept field's new value v (when upd) at byte offset o.  For
// block-scaled storage every lane of the 8-cell block takes part (cell i
// sits on lane i % 64 of its wave): the block is re-encoded with one
// exponent from its largest magnitude, each valid cell stores upd ? v : old
// (its current value), and the exponent's byte is returned (for stq).
template <class T>
__device__ __forceinline__ std::uint32_t fst(T *base, std::uint32_t o, bool upd, float v, float, bool)
{
    if (upd) stn((T *)((char *)base + o), v);
    return 0;
}
template <>
__device__ __forceinline__ std::uint32_t fst<bq16_t>(bq16_t *base, std::uint32_t o, bool upd, float v, float old,
                                                     bool valid)
{
    const float x = !valid ? 0.f : upd ? v : old;
    float mx = fabsf(x);
    mx = fmaxf(mx, dpp_f<DPP_QUAD_XOR1>(mx));
    mx = fmaxf(mx, dpp_f<DPP_QUAD_XOR2>(mx));
    mx = fmaxf(mx, dpp_f<DPP_HALF_MIRROR>(mx));
    const int e = bq_exponent(mx);
    if (valid) __builtin_nontemporal_store((std::int16_t)bq_mantissa(x, e), (std::int16_t *)((char *)base + o));
    return (std::uint32_t)(e + BQ_BIAS);
}

// The three exponent bytes of a swept field at the block of byte offset o
// (q: the field's exponent array, component x), written by the lane of the
// block's first cell.
template <class T>
__device__ __forceinline__ void stq(std::uint8_t *q, std::uint32_t o, std::uint32_t bx, std::uint32_t by,
                                    std::uint32_t bz, bool first)
{
    if (is_bq<T>::value && first) *(std::uint32_t *)(q + bq_at(o)) = bx | (by << 8) | (bz << 16);
}

// Byte-offset addressing.  Every array a sweep touches is smaller than 4 GiB
// (checked when the scene is finalized), so a 32-bit byte offset in one
// vector register, added to a scalar base, addresses every stream of a
// plane; field offsets are taken from the start of the allocation (the guard
// plane below k = 0) so that they are never negative.
template <class T>
__device__ __forceinline__ const T *at(const T *base, std::uint32_t off)
{
    return (const T *)((const char *)base + off);
}
template <class T>
__device__ __forceinline__ T *at(T *base, std::uint32_t off)
{
    return (T *)((char *)base + off);
}

// Read-only, workgroup-uniform table entry through the scalar data cache.
typedef __attribute__((address_space(4))) const float const_float;
__device__ __forceinline__ float4 uload(const float4 *p)
{
    const const_float *q = (const const_float *)p;
    return make_float4(q[0], q[1], q[2], q[3]);
}

// Coefficient-table entries of one cell: (c, cb_a, cb_b, -) per component.
struct Coef {
    float4 x, y, z;
};

template <int CM>
__device__ __forceinline__ Coef coef_of(const float4 *tab, int cl)
{
    Coef t;
    t.x = tab[cl * 3 + 0];
    t.y = tab[cl * 3 + 1];
    t.z = tab[cl * 3 + 2];
    return t;
}

template <int CM>
__device__ __forceinline__ int class_at(const void *cls, std::uint32_t cell)
{
    if (CM == CLS_U16) return *at((const std::uint16_t *)cls, cell * 2u);
    if (CM == CLS_U32) return (int)*at((const std::uint32_t *)cls, cell * 4u);
    return 0;
}

template <class A>
__device__ __forceinline__ int first_k(const A &a, int ev, int col)
{
    return (ev >= 0 && a.events[ev].col == col) ? a.events[ev].k : INT_MAX;
}

template <class A>
__device__ __forceinline__ void record_probe(const A &a, int p, int comp, float stored, int m)
{
    const float v = stored * pick6(a.probe_unscale, comp);
    a.probe_hist[(i64)p * a.steps_cap + m] = v;
    const double dv = (double)v;
    for (int f = 0; f < a.nfreq; ++f) {
        const double2 ph = a.phasor[(i64)m * a.nfreq + f];
        double2 acc = a.dft[(i64)p * a.nfreq + f];
        acc.x = acc.x + dv * ph.x;
        acc.y = acc.y + dv * ph.y;
        a.dft[(i64)p * a.nfreq + f] = acc;
    }
}


// ---- rare paths ------------------------------------------------------------
// Cells carrying a source edge or a probe are a few hundred per sweep.  Their
// handling reads a dozen further kernel arguments (event, source, waveform,
// probe and DFT tables).  Read through `a` directly, the compiler hoists those
// loads out of the z-march and holds the addresses in scalar registers for
// the whole loop, spilling the loop's own scalars to vector lanes; read
// through rare_args() they are loaded inside the rare branch each time.

// The argument block seen through a pointer the compiler cannot follow.  It
// is the first parameter of every sweep kernel, so it sits at the start of
// the kernel-argument segment (taking the parameter's own address instead
// would make the compiler copy the block to scratch).
typedef __attribute__((address_space(4))) const KArgs ConstKArgs;
__device__ __forceinline__ ConstKArgs &rare_args(const KArgs &)
{
    ConstKArgs *p = (ConstKArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *p;
}

struct EvE {
    int ev, ek;
    float ezn;
};

// Source edges and E probes of one cell at plane k (calc_e's source term,
// fdtdrav.h:726-729, and sampleField, fdtdrav.h:740).  Returns the next
// event, its plane and Ez after the source term.
template <class A>
__device__ __forceinline__ EvE e_events(const A &a, int ev, int k, int col, int m, bool upd_z,
                                                  float ez0, float hyc, float hy_im, float hxc, float hx_jm,
                                                  float exn, float eyn, float ezn)
{
    while (ev >= 0 && a.events[ev].k == k && a.events[ev].col == col) {
        const Event e = a.events[ev];
        if (e.kind == EV_SOURCE) {
            if (upd_z) {
                const SrcEdge s = a.srcs[e.id];
                ezn = ((s.ceze * ez0 + s.cezhy * (hyc - hy_im)) + s.cezhx * (hxc - hx_jm)) +
                      s.cezj * a.vpn[(i64)s.wave * a.steps_cap + m];
            }
        } else if (a.rec) {
            const int comp = e.kind - EV_PROBE;
            record_probe(a, e.id, comp, comp == 0 ? exn : comp == 1 ? eyn : ezn, m);
        }
        ++ev;
    }
    EvE r;
    r.ev = ev;
    r.ek = first_k(a, ev, col);
    r.ezn = ezn;
    return r;
}

struct EvH {
    int ev, ek;
};

// H probes of one cell at plane k (sampleField in calc_h, fdtdrav.h:347-391).
template <class A>
__device__ __forceinline__ EvH h_events(const A &a, int ev, int k, int col, int m, float hxn,
                                                  float hyn, float hzn)
{
    while (ev >= 0 && a.events[ev].k == k && a.events[ev].col == col) {
        const Event e = a.events[ev];
        const int comp = e.kind - EV_PROBE;
        if (a.rec) record_probe(a, e.id, comp, comp == 3 ? hxn : comp == 4 ? hyn : hzn, m);
        ++ev;
    }
    EvH r;
    r.ev = ev;
    r.ek = first_k(a, ev, col);
    return r;
}

// Design-region snapshot of E^{m+1} (stored precision, SI units) for the
// adjoint; only runs when the adjoint's forward pass asks for it.
template <class A>
__device__ __forceinline__ void e_snapshot(const A &a, int i, int j, int k, int m, float vx, float vy,
                                                     float vz)
{
    const i64 nb = (i64)a.sb[3] * a.sb[4] * a.sb[5];
    const i64 q = (i - a.sb[0]) + (i64)a.sb[3] * ((j - a.sb[1]) + (i64)a.sb[4] * (k - a.sb[2]));
    float *o = a.snap + (i64)(m + 1 - a.snap_base) * 3 * nb + q;
    o[0] = vx * a.probe_unscale[0];
    o[nb] = vy * a.probe_unscale[1];
    o[2 * nb] = vz * a.probe_unscale[2];
}

}  // namespace fdtd
