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

// v rounded to the storage type and back (what the next read will see)
__device__ __forceinline__ float rnd(float v, const float *) { return v; }
__device__ __forceinline__ float rnd(float v, const _Float16 *) { return (float)(_Float16)v; }
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
// the contiguous tile range [start_x, start_x + count_x).
__device__ __forceinline__ void tile_of(const KArgs &a, int &tx, int &ty, int &tz)
{
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

// Workgroup -> tile of a tile list, with the same XCD-aware remap.
__device__ __forceinline__ void tile_of_list(const KArgs &a, const int *list, int nt, int &tx, int &ty, int &tz)
{
    const int b = blockIdx.x;
    const int q = nt >> 3, r = nt & 7, x = b & 7;
    const int t = list[x * q + (x < r ? x : r) + (b >> 3)];
    tx = t % a.tiles_x;
    ty = (t / a.tiles_x) % a.tiles_y;
    tz = t / (a.tiles_x * a.tiles_y);
}

// Sweep-kernel variants by the region of their tiles.
enum SweepRegion { V_IN = 0, V_XE = 1, V_GEN = 2 };

// Workgroup -> tile of variant V: the interior box, the x-edge list, or the
// general list (the whole grid when there is no box).
template <int V>
__device__ __forceinline__ void tile_of_region(const KArgs &a, int &tx, int &ty, int &tz)
{
    if (V == V_IN) tile_of_box(a, tx, ty, tz);
    else if (V == V_XE) tile_of_list(a, a.xlist, a.nxlist, tx, ty, tz);
    else if (a.nglist == 0) tile_of(a, tx, ty, tz);
    else tile_of_list(a, a.glist, a.nglist, tx, ty, tz);
}

__device__ __forceinline__ int first_event(const KArgs &a, const int *first, int col, int k0)
{
    int ev = first[col];
    if (ev < 0) return -1;
    while (a.events[ev].col == col && a.events[ev].k < k0) ++ev;
    return a.events[ev].col == col ? ev : -1;
}

}  // namespace fdtd
