// device_common.h -- storage conversions and workgroup mapping shared by the
// forward kernels (kernels.hip) and the adjoint kernels (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "internal.h"

namespace fdtd {

constexpr int TX = 64;   // tile width (one wavefront row)
#ifndef FDTD_TILE_ROWS
#define FDTD_TILE_ROWS 4
#endif
constexpr int BY = FDTD_TILE_ROWS;   // tile rows per workgroup

// ---- storage types ------------------------------------------------------
struct bf16_t {
    std::uint16_t v;
};

__device__ __forceinline__ float ld(const float *p) { return *p; }
__device__ __forceinline__ void st(float *p, float v) { *p = v; }
__device__ __forceinline__ float ld(const _Float16 *p) { return (float)*p; }
__device__ __forceinline__ void st(_Float16 *p, float v) { *p = (_Float16)v; }
__device__ __forceinline__ float ld(const bf16_t *p) { return __uint_as_float((std::uint32_t)p->v << 16); }
__device__ __forceinline__ void st(bf16_t *p, float v)
{
    std::uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    p->v = (std::uint16_t)(u >> 16);
}

// Block-scaled 16-bit storage of the fields: an int16 mantissa per value and
// one power-of-two exponent byte per aligned run of BQ_BLOCK cells along x,
// value = mantissa * 2^(byte - BQ_BIAS).  The exponents of the three
// components of E (and of H) share one array with the field's layout, x
// divided by BQ_BLOCK and BQ_STRIDE bytes per block: bytes 0, 1, 2 hold the
// x, y, z components' exponents (byte 3 is unused), so one 4-byte load
// brings all three.  The byte of component c of the value at byte offset o
// (2-byte elements) is at (o / (2 * BQ_BLOCK)) * BQ_STRIDE + c.  psi is
// stored in FP16 (psi_type).
struct bq16_t {
    std::int16_t v;
};
// BQ_BLOCK, BQ_BIAS, BQ_STRIDE: internal.h (the host code shares them)

template <class T> struct psi_type { typedef T type; };
template <> struct psi_type<bq16_t> { typedef _Float16 type; };
template <class T> struct is_bq { static constexpr bool value = false; };
template <> struct is_bq<bq16_t> { static constexpr bool value = true; };

// v rounded to the storage type and back (what the next read will see)
__device__ __forceinline__ float rnd(float v, const float *) { return v; }
__device__ __forceinline__ float rnd(float v, const _Float16 *) { return (float)(_Float16)v; }
// block-scaled storage: the stored value depends on the whole block, which
// one lane does not see here (the adjoint, the only caller, refuses this
// storage type)
__device__ __forceinline__ float rnd(float v, const bq16_t *) { return v; }
__device__ __forceinline__ float rnd(float v, const bf16_t *)
{
    std::uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}

// per-component scale from a kernel-argument array without dynamic indexing
template <class V>
__device__ __forceinline__ float pick6(const V &v, int c)
{
    float r = v[0];
    r = c == 1 ? v[1] : r;
    r = c == 2 ? v[2] : r;
    r = c == 3 ? v[3] : r;
    r = c == 4 ? v[4] : r;
    r = c == 5 ? v[5] : r;
    return r;
}

template <int CM>
__device__ __forceinline__ int class_of(const void *cls, i64 c)
{
    if (CM == CLS_U16) return ((const std::uint16_t *)cls)[c];
    if (CM == CLS_U32) return (int)((const std::uint32_t *)cls)[c];
    return 0;
}

// Workgroup -> tile: XCD x (= blockIdx % 8 under round-robin dispatch) gets
// the contiguous tile range [start_x, start_x + count_x).  With a light-cone
// sub-range (a.sub), the tiles of that range only.
__device__ __forceinline__ void tile_of(const KArgs &a, int &tx, int &ty, int &tz)
{
    if (a.sub[3] > 0) {
        const int b = blockIdx.x, nt = a.sub[3] * a.sub[4] * a.sub[5];
        const int q = nt >> 3, r = nt & 7, x = b & 7;
        const int t = x * q + (x < r ? x : r) + (b >> 3);
        tx = a.sub[0] + t % a.sub[3];
        ty = a.sub[1] + (t / a.sub[3]) % a.sub[4];
        tz = a.sub[2] + t / (a.sub[3] * a.sub[4]);
        return;
    }
    const int nt = a.tiles_x * a.tiles_y * ((a.nz + a.kchunk - 1) / a.kchunk);
    const int b = blockIdx.x;
    const int q = nt >> 3, r = nt & 7, x = b & 7;
    const int t = x * q + (x < r ? x : r) + (b >> 3);
    tx = t % a.tiles_x;
    ty = (t / a.tiles_x) % a.tiles_y;
    tz = t / (a.tiles_x * a.tiles_y);
}


// Workgroup -> tile of the interior box, with the same XCD-aware remap.
__device__ __forceinline__ void tile_of_box(const KArgs &a, int &tx, int &ty, int &tz)
{
    const int bx = a.ib[3] - a.ib[0], by = a.ib[4] - a.ib[1];
    const int nt = bx * by * (a.ib[5] - a.ib[2]);
    const int b = blockIdx.x;
    const int q = nt >> 3, r = nt & 7, x = b & 7;
    const int t = x * q + (x < r ? x : r) + (b >> 3);
    tx = a.ib[0] + t % bx;
    ty = a.ib[1] + (t / bx) % by;
    tz = a.ib[2] + t / (bx * by);
}

// XCD-aware remap of workgroup b over nt tiles (as in tile_of).
__device__ __forceinline__ int remap_tile(int nt)
{
    const int b = blockIdx.x;
    const int q = nt >> 3, r = nt & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Workgroup -> tile of region variant V, computed from the interior box (no
// tile lists: the many short workgroups then start without a dependent
// load).  V_X: tile columns outside [ib0, ib3) of the box's rows and
// z-blocks.  V_Z (with zsplit): the box's tile columns and rows in every
// z-block outside [ib2, ib5).  V_GEN: every z-block outside [ib2, ib5) (with
// zsplit only its tiles off the box's columns and rows), then the rows
// outside [ib1, ib4) of the box's z-blocks; the whole grid (or the light-cone
// range) when there is no box.
template <int V>
__device__ __forceinline__ void tile_of_region(const KArgs &a, int &tx, int &ty, int &tz)
{
    const int *ib = a.ib;
    if (V == V_IN) {
        tile_of_box(a, tx, ty, tz);
    } else if (V == V_X) {
        const int ex = ib[0] + a.tiles_x - ib[3], by = ib[4] - ib[1];
        const int t = remap_tile(ex * by * (ib[5] - ib[2]));
        const int xe = t % ex;
        tx = xe < ib[0] ? xe : ib[3] + (xe - ib[0]);
        ty = ib[1] + (t / ex) % by;
        tz = ib[2] + t / (ex * by);
    } else if (V == V_Z) {
        const int bx = ib[3] - ib[0], by = ib[4] - ib[1], bz = ib[5] - ib[2];
        const int tzs = (a.nz + a.kchunk - 1) / a.kchunk;
        const int t = remap_tile(bx * by * (tzs - bz));
        const int ze = t / (bx * by);
        tx = ib[0] + t % bx;
        ty = ib[1] + (t / bx) % by;
        tz = ze < ib[2] ? ze : ib[5] + (ze - ib[2]);
    } else if (ib[5] == ib[2]) {
        tile_of(a, tx, ty, tz);
    } else {
        const int txy = a.tiles_x * a.tiles_y, tzs = (a.nz + a.kchunk - 1) / a.kchunk;
        const int bx = ib[3] - ib[0], by = ib[4] - ib[1], bz = ib[5] - ib[2], ey = a.tiles_y - by;
        const int per = a.zsplit ? txy - bx * by : txy;   // tiles per z-block outside the box's z range
        const int n1 = per * (tzs - bz);
        const int t = remap_tile(n1 + a.tiles_x * ey * bz);
        if (t < n1) {
            const int ze = t / per, u = t % per;
            tz = ze < ib[2] ? ze : ib[5] + (ze - ib[2]);
            if (!a.zsplit || u < ey * a.tiles_x) {
                // every tile of the plane, or (zsplit) the rows outside the box's
                const int w = u / a.tiles_x;
                tx = u % a.tiles_x;
                ty = !a.zsplit ? w : (w < ib[1] ? w : ib[4] + (w - ib[1]));
            } else {
                // (zsplit) the box's rows, columns outside it
                const int v = u - ey * a.tiles_x, ex = a.tiles_x - bx, xe = v % ex;
                tx = xe < ib[0] ? xe : ib[3] + (xe - ib[0]);
                ty = ib[1] + v / ex;
            }
        } else {
            const int u = t - n1;
            const int ye = (u / a.tiles_x) % ey;
            tx = u % a.tiles_x;
            ty = ye < ib[1] ? ye : ib[4] + (ye - ib[1]);
            tz = ib[2] + u / (a.tiles_x * ey);
        }
    }
}

__device__ __forceinline__ int first_event(const KArgs &a, const int *first, int col, int k0)
{
    int ev = first[col];
    if (ev < 0) return -1;
    while (a.events[ev].col == col && a.events[ev].k < k0) ++ev;
    return a.events[ev].col == col ? ev : -1;
}

// The x-neighbour of a lane's cell as the adjacent lane's own value (a wave
// is one 64-cell tile row): a DPP wave shift, no memory access.  The tile's
// edge lane, which has no source lane, takes `fill`.  Loading the neighbour
// instead reads each row at a 4-byte offset, three 128-byte lines where the
// wave's own loads touch two.
__device__ __forceinline__ float lane_prev(float v, float fill)   // lane l <- lane l-1; lane 0 <- fill
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_next(float v, float fill)   // lane l <- lane l+1; lane 63 <- fill
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), 0x130, 0xF, 0xF, false));
}
// the same for packed 16-bit pairs (two cells per lane)
__device__ __forceinline__ std::uint32_t lane_prev_u(std::uint32_t v, std::uint32_t fill)
{
    return (std::uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ std::uint32_t lane_next_u(std::uint32_t v, std::uint32_t fill)
{
    return (std::uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x130, 0xF, 0xF, false);
}

}  // namespace fdtd
