// fused.hip -- one kernel per time step: E^{n+1} and H^{n+1/2} in a single
// z-march, for gfx950 (MI355X).
//
// The reference runs calc_e then calc_h over the whole grid
// (fdtdrav.h:3222-3341), so every field component crosses HBM once per
// half-step: E and H are each read twice and written once per step (18
// values).  Here each wavefront marches k over a strip of the grid and
// produces, per plane, E^{n+1} of plane k+1 and then H^{n+1/2} of plane k,
// so each component is read once and written once per step (12 values).
//
// Strip: lane l of a wavefront owns column i = 62*w - 1 + l, so consecutive
// wavefronts overlap by two columns (lanes 0 and 63 are halo columns whose
// results are not stored); each thread carries STRIP_R rows j0..j0+R-1 plus
// the halo row j0+R in registers.  The -x/+x neighbours come from the
// adjacent lanes (cross-lane shuffles), the -y/+y neighbours from the
// adjacent register rows, the z neighbours from the planes carried from the
// previous iteration.  E^{n+1} is computed on the halo row and column too
// (recomputed rather than exchanged), so wavefronts never wait on each other:
// no LDS staging, no workgroup barriers, and latency is hidden the way the
// half-step kernels hide it, by many independent wavefronts.  Reads come from
// one buffer set and writes go to the other (ping-pong), because a
// neighbouring wavefront may still need the old value of a halo cell; E-side
// psi is ping-ponged for the same reason, H-side psi is only touched by its
// owner and stays in place.
//
// The arithmetic of every cell is that of the half-step kernels in
// kernels.hip (the reference's operand order, no contraction), so the FP32
// form reproduces the CPU restatement bit for bit.
#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "device_common.h"
#include "internal.h"

namespace fdtd {

constexpr int SO = 62;    // owned columns per wavefront
constexpr int WPB = 4;    // wavefronts per workgroup (consecutive strips)
constexpr int R = STRIP_R;
#ifndef STEP_MIN_WAVES
#define STEP_MIN_WAVES 4   // wavefronts per SIMD the register budget must allow
#endif

namespace {

struct Cur {
    int ev, ek;
};

__device__ __forceinline__ void cur_init(const KArgs &a, int key, int k0, Cur &c)
{
    c.ev = -1;
    c.ek = INT_MAX;
    if (key < 0) return;
    int e = a.s_first[key];
    if (e < 0) return;
    while (a.s_events[e].col == key && a.s_events[e].k < k0) ++e;
    c.ev = e;
    c.ek = a.s_events[e].col == key ? a.s_events[e].k : INT_MAX;
}

__device__ __forceinline__ void cur_next(const KArgs &a, int key, int k, Cur &c)
{
    if (c.ek != k) return;
    while (a.s_events[c.ev].col == key && a.s_events[c.ev].k <= k) ++c.ev;
    c.ek = a.s_events[c.ev].col == key ? a.s_events[c.ev].k : INT_MAX;
}

__device__ __forceinline__ void record(const KArgs &a, int p, int comp, float stored, int m)
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

// Cross-lane neighbours with DPP whole-wavefront shifts: lane l receives
// the value of lane l-1 (wave_shr:1) or l+1 (wave_shl:1); the lane that has
// no source receives 0 (its result is never used: lanes 0 and 63 are halo).
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}

template <class T>
__device__ __forceinline__ float ld0(const void *p, bool ok, i64 c)
{
    return ok ? ld((const T *)p + c) : 0.f;
}

struct F3 {
    float x, y, z;
};

// Inputs of one E cell beyond its own (i, j, k): the H values its update
// reads and the loaded E^n and class index.
struct EIn {
    float hxc, hyc, hzc, hy_im, hz_im, hx_jm, hz_jm, hx_km, hy_km;
    float e0[3];
    int cls;
};

// E^{n+1} of (i, j, k) -> out[3] (rounded to the storage precision).
// Same arithmetic as e_kernel (fdtdrav.h:665-911).  store: write HBM, new
// psi_e, probes and the design snapshot (owned cell inside the chunk).
template <class T, int CM>
__device__ __forceinline__ F3 e_cell(const KArgs &a, int i, int j, int k, int m, int r, bool store, int key,
                                     const Cur &cur, const EIn in)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    const i64 c = (i64)k * a.plane + (i64)j * px + i;
    float4 cx = a.etab[0], cy = a.etab[1], cz = a.etab[2];
    if (CM != CLS_UNIFORM && !a.e_uni) {
        cx = a.etab[in.cls * 3 + 0];
        cy = a.etab[in.cls * 3 + 1];
        cz = a.etab[in.cls * 3 + 2];
    }
    const bool ilo = i > 0, ihi = i < nx - 1, jlo = j > 0, jhi = j < ny - 1, klo = k > 0, khi = k < nz - 1;
    const bool upd_x = ihi && jlo && jhi && klo && khi;
    const bool upd_y = ilo && ihi && jhi && klo && khi;
    const bool upd_z = ilo && ihi && jlo && jhi && khi;
    const bool xs = (ilo && i < a.L[0]) || (i >= nx - a.L[1] && ihi);
    const bool ys = (jlo && j < a.L[2]) || (j >= ny - a.L[3] && jhi);
    const bool zs = (klo && k < a.L[4]) || (k >= nz - a.L[5] && khi);
    float cbx_hz = cx.y, cbx_hy = cx.z;
    float cby_hx = cy.y, cby_hz = cy.z;
    float cbz_hy = cz.y, cbz_hx = cz.z;
    float4 pxe = make_float4(0.f, 0.f, 1.f, 0.f), pye = pxe, pze = pxe;
    if (xs) { pxe = a.px_prof[i]; cby_hz = cby_hz / pxe.z; cbz_hy = cbz_hy / pxe.z; }
    if (ys) { pye = a.py_prof[j]; cbx_hz = cbx_hz / pye.z; cbz_hx = cbz_hx / pye.z; }
    if (zs) { pze = a.pz_prof[k]; cbx_hy = cbx_hy / pze.z; cby_hx = cby_hx / pze.z; }

    const float ex0 = in.e0[0], ey0 = in.e0[1], ez0 = in.e0[2];
    float exn = ex0, eyn = ey0, ezn = ez0;
    if (upd_x) exn = (cx.x * ex0 + cbx_hz * (in.hzc - in.hz_jm)) + cbx_hy * (in.hyc - in.hy_km);
    if (upd_y) eyn = (cy.x * ey0 + cby_hx * (in.hxc - in.hx_km)) + cby_hz * (in.hzc - in.hz_im);
    if (upd_z) ezn = (cz.x * ez0 + cbz_hy * (in.hyc - in.hy_im)) + cbz_hx * (in.hxc - in.hx_jm);

    if (k == cur.ek) {   // lumped-source edges and E probes of this cell
        for (int e = cur.ev; a.s_events[e].col == key && a.s_events[e].k == k; ++e) {
            const Event ev = a.s_events[e];
            if ((ev.kind >> 8) != r) continue;
            const int kind = ev.kind & 0xFF;
            if (kind == EV_SOURCE) {
                if (upd_z) {
                    const SrcEdge s = a.srcs[ev.id];
                    ezn = ((s.ceze * ez0 + s.cezhy * (in.hyc - in.hy_im)) + s.cezhx * (in.hxc - in.hx_jm)) +
                          s.cezj * a.vpn[(i64)s.wave * a.steps_cap + m];
                }
            } else if (kind <= EV_PROBE + 2 && store && a.rec) {
                const int comp = kind - EV_PROBE;
                record(a, ev.id, comp, comp == 0 ? exn : comp == 1 ? eyn : ezn, m);
            }
        }
    }

    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    if (xs && jhi && khi) {
        const i64 q = ((i64)k * ny + j) * a.Lx + (i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0]);
        const float pey = pxe.x * ld((const T *)a.psi_eyx + q) + pxe.y * (in.hzc - in.hz_im);
        const float pez = pxe.x * ld((const T *)a.psi_ezx + q) + pxe.y * (in.hyc - in.hy_im);
        if (store) {
            st((T *)a.npsi_eyx + q, pey);
            st((T *)a.npsi_ezx + q, pez);
        }
        if (klo) eyn = eyn + (cy.z * dx) * pey;
        if (jlo) ezn = ezn + (cz.y * dx) * pez;
    }
    if (ys && ihi && khi) {
        const i64 q = ((i64)k * a.Ly + (j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2])) * px + i;
        const float pex = pye.x * ld((const T *)a.psi_exy + q) + pye.y * (in.hzc - in.hz_jm);
        const float pez = pye.x * ld((const T *)a.psi_ezy + q) + pye.y * (in.hxc - in.hx_jm);
        if (store) {
            st((T *)a.npsi_exy + q, pex);
            st((T *)a.npsi_ezy + q, pez);
        }
        if (klo) exn = exn + (cx.y * dy) * pex;
        if (ilo) ezn = ezn + (cz.z * dy) * pez;
    }
    if (zs && ihi && jhi) {
        const i64 q = ((i64)(k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4]) * ny + j) * px + i;
        const float pex = pze.x * ld((const T *)a.psi_exz + q) + pze.y * (in.hyc - in.hy_km);
        const float pey = pze.x * ld((const T *)a.psi_eyz + q) + pze.y * (in.hxc - in.hx_km);
        if (store) {
            st((T *)a.npsi_exz + q, pex);
            st((T *)a.npsi_eyz + q, pey);
        }
        if (jlo) exn = exn + (cx.z * dz) * pex;
        if (ilo) eyn = eyn + (cy.y * dz) * pey;
    }
    // what the next read of this cell returns: the stored precision
    F3 out;
    out.x = rnd(exn, (const T *)nullptr);
    out.y = rnd(eyn, (const T *)nullptr);
    out.z = rnd(ezn, (const T *)nullptr);
    if (store) {
        if (upd_x) st((T *)a.nex + c, exn);
        if (upd_y) st((T *)a.ney + c, eyn);
        if (upd_z) st((T *)a.nez + c, ezn);
        if (a.snap && i >= a.sb[0] && i < a.sb[0] + a.sb[3] && j >= a.sb[1] && j < a.sb[1] + a.sb[4] &&
            k >= a.sb[2] && k < a.sb[2] + a.sb[5]) {
            const i64 nb = (i64)a.sb[3] * a.sb[4] * a.sb[5];
            const i64 q = (i - a.sb[0]) + (i64)a.sb[3] * ((j - a.sb[1]) + (i64)a.sb[4] * (k - a.sb[2]));
            float *o = a.snap + (i64)(m + 1 - a.snap_base) * 3 * nb + q;
            o[0] = out.x * a.probe_unscale[0];
            o[nb] = out.y * a.probe_unscale[1];
            o[2 * nb] = out.z * a.probe_unscale[2];
        }
    }
    return out;
}

// Inputs of one H cell: E^{n+1} of planes k and k+1 around it, H^{n-1/2}.
struct HIn {
    float ex_c, ey_c, ezc, ex_kp, ey_kp, ez_ip, ey_ip, ez_jp, ex_jp, h0[3];
};

// H^{n+1/2} of (i, j, k); same arithmetic as h_kernel (fdtdrav.h:328-562).
template <class T, int CM>
__device__ __forceinline__ void h_cell(const KArgs &a, int i, int j, int k, int m, int r, int key, const Cur &cur,
                                       const HIn in)
{
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    const i64 c = (i64)k * a.plane + (i64)j * px + i;
    float4 cx = a.htab[0], cy = a.htab[1], cz = a.htab[2];
    if (CM != CLS_UNIFORM && !a.h_uni) {
        const int cl = class_of<CM>(a.cls, c);
        cx = a.htab[cl * 3 + 0];
        cy = a.htab[cl * 3 + 1];
        cz = a.htab[cl * 3 + 2];
    }
    const bool ihi = i < nx - 1, jhi = j < ny - 1, khi = k < nz - 1;
    const bool upd_x = jhi && khi, upd_y = ihi && khi, upd_z = ihi && jhi;
    const bool xs = i < a.L[0] || (i >= nx - a.L[1] && ihi);
    const bool ys = j < a.L[2] || (j >= ny - a.L[3] && jhi);
    const bool zs = k < a.L[4] || (k >= nz - a.L[5] && khi);
    float chx_ey = cx.y, chx_ez = cx.z;
    float chy_ez = cy.y, chy_ex = cy.z;
    float chz_ex = cz.y, chz_ey = cz.z;
    float4 pxm = make_float4(0.f, 0.f, 1.f, 0.f), pym = pxm, pzm = pxm;
    if (xs) { pxm = a.px_prof[nx + i]; chy_ez = chy_ez / pxm.z; chz_ey = chz_ey / pxm.z; }
    if (ys) { pym = a.py_prof[ny + j]; chx_ez = chx_ez / pym.z; chz_ex = chz_ex / pym.z; }
    if (zs) { pzm = a.pz_prof[nz + k]; chx_ey = chx_ey / pzm.z; chy_ex = chy_ex / pzm.z; }

    const float hx0 = in.h0[0], hy0 = in.h0[1], hz0 = in.h0[2];
    float hxn = hx0, hyn = hy0, hzn = hz0;
    if (upd_x) hxn = (cx.x * hx0 + chx_ey * (in.ey_kp - in.ey_c)) + chx_ez * (in.ez_jp - in.ezc);
    if (upd_y) hyn = (cy.x * hy0 + chy_ez * (in.ez_ip - in.ezc)) + chy_ex * (in.ex_kp - in.ex_c);
    if (upd_z) hzn = (cz.x * hz0 + chz_ex * (in.ex_jp - in.ex_c)) + chz_ey * (in.ey_ip - in.ey_c);

    if (k == cur.ek && a.rec) {   // H probes of this cell
        for (int e = cur.ev; a.s_events[e].col == key && a.s_events[e].k == k; ++e) {
            const Event ev = a.s_events[e];
            const int kind = ev.kind & 0xFF;
            if ((ev.kind >> 8) != r || kind < EV_PROBE + 3) continue;
            const int comp = kind - EV_PROBE;
            record(a, ev.id, comp, comp == 3 ? hxn : comp == 4 ? hyn : hzn, m);
        }
    }
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];
    if (xs) {
        const i64 q = ((i64)k * ny + j) * a.Lx + (i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0]);
        T *pa = (T *)a.psi_hyx + q;
        T *pb = (T *)a.psi_hzx + q;
        const float phy = pxm.x * ld(pa) + pxm.y * (in.ez_ip - in.ezc);
        const float phz = pxm.x * ld(pb) + pxm.y * (in.ey_ip - in.ey_c);
        st(pa, phy);
        st(pb, phz);
        if (khi) hyn = hyn + (cy.y * dx) * phy;
        if (jhi) hzn = hzn + (cz.z * dx) * phz;
    }
    if (ys) {
        const i64 q = ((i64)k * a.Ly + (j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2])) * px + i;
        T *pa = (T *)a.psi_hxy + q;
        T *pb = (T *)a.psi_hzy + q;
        const float phx = pym.x * ld(pa) + pym.y * (in.ez_jp - in.ezc);
        const float phz = pym.x * ld(pb) + pym.y * (in.ex_jp - in.ex_c);
        st(pa, phx);
        st(pb, phz);
        if (khi) hxn = hxn + (cx.z * dy) * phx;
        if (ihi) hzn = hzn + (cz.y * dy) * phz;
    }
    if (zs) {
        const i64 q = ((i64)(k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4]) * ny + j) * px + i;
        T *pa = (T *)a.psi_hxz + q;
        T *pb = (T *)a.psi_hyz + q;
        const float phx = pzm.x * ld(pa) + pzm.y * (in.ey_kp - in.ey_c);
        const float phy = pzm.x * ld(pb) + pzm.y * (in.ex_kp - in.ex_c);
        st(pa, phx);
        st(pb, phy);
        if (jhi) hxn = hxn + (cx.y * dz) * phx;
        if (ihi) hyn = hyn + (cy.z * dz) * phy;
    }
    if (upd_x) st((T *)a.nhx + c, hxn);
    if (upd_y) st((T *)a.nhy + c, hyn);
    if (upd_z) st((T *)a.nhz + c, hzn);
}

}  // namespace

template <class T, int CM>
__global__ void __launch_bounds__(64 * WPB, STEP_MIN_WAVES) step_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of(a, tx, ty, tz);
    const int lane = (int)(threadIdx.x & 63);
    const int s = ty * WPB + (int)(threadIdx.x >> 6);   // strip
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    const int nstrip = (ny + R - 1) / R;
    if (s >= nstrip) return;   // a whole wavefront: there are no barriers here
    const int i = tx * SO - 1 + lane;
    const int j0 = s * R;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const bool col_ok = i >= 0 && i < nx;
    const bool owned = col_ok && lane >= 1 && lane <= SO;
    const int key = col_ok ? s * nx + i : -1;
    bool rok[R + 2];   // rows j0-1 .. j0+R inside the grid (and a real column)
#pragma unroll
    for (int r = 0; r < R + 2; ++r) rok[r] = col_ok && j0 - 1 + r >= 0 && j0 - 1 + r < ny;
    Cur ce, ch;
    cur_init(a, key, k0, ce);
    cur_init(a, key, k0, ch);

    // carried: H^{n-1/2} of plane k and E^{n+1} of plane k, rows j0..j0+R
    F3 h[R + 1], e[R + 1];
    // this plane: H^{n-1/2} of plane k+1 (rows j0-1..j0+R), E^{n+1} of plane k+1
    F3 nh[R + 2], ne[R + 1];

#define LOAD_H_PLANE(kp)                                                                   \
    _Pragma("unroll") for (int r = 0; r < R + 2; ++r)                                      \
    {                                                                                      \
        const bool ok_ = rok[r] && (kp) >= 0 && (kp) < nz;                                 \
        const i64 c_ = (i64)(kp) * a.plane + (i64)(j0 - 1 + r) * px + i;                  \
        nh[r].x = ld0<T>(a.hx, ok_, c_);                                                   \
        nh[r].y = ld0<T>(a.hy, ok_, c_);                                                   \
        nh[r].z = ld0<T>(a.hz, ok_, c_);                                                   \
    }

    // E^{n+1} of plane kp, rows j0..j0+R, from H planes kp (nh) and kp-1 (h)
#define E_PLANE(kp, in_chunk)                                                              \
    {                                                                                      \
        F3 e0_[R + 1];                                                                     \
        int cls_[R + 1];                                                                   \
        _Pragma("unroll") for (int r = 0; r <= R; ++r)                                     \
        {                                                                                  \
            const bool ok_ = rok[r + 1];                                                   \
            const i64 c_ = (i64)(kp) * a.plane + (i64)(j0 + r) * px + i;                   \
            e0_[r].x = ld0<T>(a.ex, ok_, c_);                                              \
            e0_[r].y = ld0<T>(a.ey, ok_, c_);                                              \
            e0_[r].z = ld0<T>(a.ez, ok_, c_);                                              \
            cls_[r] = (CM != CLS_UNIFORM && !a.e_uni && ok_) ? class_of<CM>(a.cls, c_) : 0; \
        }                                                                                  \
        _Pragma("unroll") for (int r = 0; r <= R; ++r)                                     \
        {                                                                                  \
            EIn in_;                                                                       \
            in_.hxc = nh[r + 1].x;                                                         \
            in_.hyc = nh[r + 1].y;                                                         \
            in_.hzc = nh[r + 1].z;                                                         \
            in_.hy_im = from_left(nh[r + 1].y);                                            \
            in_.hz_im = from_left(nh[r + 1].z);                                            \
            in_.hx_jm = nh[r].x;                                                           \
            in_.hz_jm = nh[r].z;                                                           \
            in_.hx_km = h[r].x;                                                            \
            in_.hy_km = h[r].y;                                                            \
            in_.e0[0] = e0_[r].x;                                                          \
            in_.e0[1] = e0_[r].y;                                                          \
            in_.e0[2] = e0_[r].z;                                                          \
            in_.cls = cls_[r];                                                             \
            ne[r] = rok[r + 1] ? e_cell<T, CM>(a, i, j0 + r, (kp), m, r, owned && r < R && (in_chunk), key, ce, in_) \
                               : F3{0.f, 0.f, 0.f};                                         \
        }                                                                                  \
        cur_next(a, key, (kp), ce);                                                        \
    }

    // prologue: H planes k0-1 (x, y) and k0, E^{n+1} of plane k0
    LOAD_H_PLANE(k0 - 1)
#pragma unroll
    for (int r = 0; r <= R; ++r) h[r] = nh[r + 1];
    LOAD_H_PLANE(k0)
    E_PLANE(k0, true)
#pragma unroll
    for (int r = 0; r <= R; ++r) {
        h[r] = nh[r + 1];
        e[r] = ne[r];
    }

    for (int k = k0; k < k1; ++k) {
        const int kk = k + 1;
        if (kk < nz) {
            LOAD_H_PLANE(kk)
            E_PLANE(kk, kk < k1)
        } else {
#pragma unroll
            for (int r = 0; r <= R; ++r) {
                ne[r] = F3{0.f, 0.f, 0.f};
                nh[r + 1] = F3{0.f, 0.f, 0.f};
            }
        }
        // H^{n+1/2} of plane k, rows j0..j0+R-1
#pragma unroll
        for (int r = 0; r < R; ++r) {
            HIn in;
            in.ex_c = e[r].x;
            in.ey_c = e[r].y;
            in.ezc = e[r].z;
            in.ex_kp = ne[r].x;
            in.ey_kp = ne[r].y;
            in.ez_ip = from_right(e[r].z);
            in.ey_ip = from_right(e[r].y);
            in.ez_jp = e[r + 1].z;
            in.ex_jp = e[r + 1].x;
            in.h0[0] = h[r].x;
            in.h0[1] = h[r].y;
            in.h0[2] = h[r].z;
            if (owned && rok[r + 1]) h_cell<T, CM>(a, i, j0 + r, k, m, r, key, ch, in);
        }
        cur_next(a, key, k, ch);
#pragma unroll
        for (int r = 0; r <= R; ++r) {
            h[r] = nh[r + 1];
            e[r] = ne[r];
        }
    }
#undef LOAD_H_PLANE
#undef E_PLANE
}

hipError_t launch_step(int storage, int mode, const KArgs &a, int m, hipStream_t s)
{
    const int nt = a.tiles_x * a.tiles_y * ((a.nz + a.kchunk - 1) / a.kchunk);
    dim3 grid(nt), block(64 * WPB);
    if (storage == 1) {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((step_kernel<_Float16, CLS_UNIFORM>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((step_kernel<_Float16, CLS_U16>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((step_kernel<_Float16, CLS_U32>), grid, block, 0, s, a, m);
    } else if (storage == 2) {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((step_kernel<bf16_t, CLS_UNIFORM>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((step_kernel<bf16_t, CLS_U16>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((step_kernel<bf16_t, CLS_U32>), grid, block, 0, s, a, m);
    } else {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((step_kernel<float, CLS_UNIFORM>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((step_kernel<float, CLS_U16>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((step_kernel<float, CLS_U32>), grid, block, 0, s, a, m);
    }
    return hipGetLastError();
}

// Tile geometry of the step kernel: a tile is WPB consecutive strips of one
// wavefront column block (lanes 1..62 of block w own columns 62w..62w+61).
void step_tiles(int nx, int ny, int &tiles_x, int &tiles_y)
{
    tiles_x = (nx + SO - 1) / SO;
    const int nstrip = (ny + R - 1) / R;
    tiles_y = (nstrip + WPB - 1) / WPB;
}

}  // namespace fdtd
