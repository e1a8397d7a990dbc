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

namespace fdtd {

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

// Everything the E update of one cell reads from memory, loaded a plane
// ahead of its use so that the next plane's loads are in flight while this
// plane is computed.
struct EPlane {
    float hxc, hyc, hzc, hy_im, hz_im, hx_jm, hz_jm, ex0, ey0, ez0;
    int cl;
};

template <class T, int CM>
__device__ __forceinline__ EPlane e_load(const KArgs &a, const T *hx, const T *hy, const T *hz, const T *ex,
                                         const T *ey, const T *ez, i64 c, int px)
{
    EPlane p;
    p.hxc = ld(hx + c);
    p.hyc = ld(hy + c);
    p.hzc = ld(hz + c);
    p.hy_im = ld(hy + c - 1);
    p.hz_im = ld(hz + c - 1);
    p.hx_jm = ld(hx + c - px);
    p.hz_jm = ld(hz + c - px);
    p.ex0 = ldn(ex + c);
    p.ey0 = ldn(ey + c);
    p.ez0 = ldn(ez + c);
    p.cl = CM != CLS_UNIFORM ? class_of<CM>(a.cls, c) : 0;
    return p;
}

// Everything the H update of one cell reads (E of planes k and k+1 around
// it, its own H), loaded a plane ahead.
struct HPlane {
    float ex_kp, ey_kp, ezc, ez_ip, ey_ip, ez_jp, ex_jp, hx0, hy0, hz0;
    int cl;
};

template <class T, int CM>
__device__ __forceinline__ HPlane h_load(const KArgs &a, const T *ex, const T *ey, const T *ez, const T *hx,
                                         const T *hy, const T *hz, i64 c, i64 plane, int px)
{
    HPlane p;
    p.ex_kp = ld(ex + c + plane);
    p.ey_kp = ld(ey + c + plane);
    p.ezc = ld(ez + c);
    p.ez_ip = ld(ez + c + 1);
    p.ey_ip = ld(ey + c + 1);
    p.ez_jp = ld(ez + c + px);
    p.ex_jp = ld(ex + c + px);
    p.hx0 = ldn(hx + c);
    p.hy0 = ldn(hy + c);
    p.hz0 = ldn(hz + c);
    p.cl = CM != CLS_UNIFORM ? class_of<CM>(a.cls, c) : 0;
    return p;
}

__device__ __forceinline__ void record_probe(const KArgs &a, int p, int comp, float stored, int m)
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

// ---- E half-step ----------------------------------------------------------
template <class T, int CM>
__global__ void __launch_bounds__(TX * BY) e_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of(a, tx, ty, tz);
    const int i = tx * TX + (int)(threadIdx.x & (TX - 1));
    const int j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    if (i >= nx || j >= ny) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const i64 plane = a.plane;
    const i64 col = (i64)j * px + i;

    T *__restrict__ ex = (T *)a.ex;
    T *__restrict__ ey = (T *)a.ey;
    T *__restrict__ ez = (T *)a.ez;
    const T *__restrict__ hx = (const T *)a.hx;
    const T *__restrict__ hy = (const T *)a.hy;
    const T *__restrict__ hz = (const T *)a.hz;

    // column-constant predicates (reference region guards)
    const bool ilo = i > 0, ihi = i < nx - 1, jlo = j > 0, jhi = j < ny - 1;
    const bool xs = (ilo && i < a.L[0]) || (i >= nx - a.L[1] && ihi);   // E x-slab, fdtdrav.h:752,777
    const bool ys = (jlo && j < a.L[2]) || (j >= ny - a.L[3] && jhi);   // E y-slab, fdtdrav.h:803,827
    const int sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
    const int sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    const float4 pxe = xs ? a.px_prof[i] : make_float4(0.f, 0.f, 1.f, 0.f);
    const float4 pye = ys ? a.py_prof[j] : make_float4(0.f, 0.f, 1.f, 0.f);
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];

    float4 cx = a.etab[0], cy = a.etab[1], cz = a.etab[2];

    const int mycol = j * nx + i;
    int ev = first_event(a, a.e_first, mycol, k0);
    int ek = ev >= 0 ? a.events[ev].k : INT_MAX;

    float hx_km = ld(hx + col + (i64)(k0 - 1) * plane);
    float hy_km = ld(hy + col + (i64)(k0 - 1) * plane);
    const bool in_snap = a.snap != nullptr && i >= a.sb[0] && i < a.sb[0] + a.sb[3] && j >= a.sb[1] &&
                         j < a.sb[1] + a.sb[4];

    EPlane nxt = e_load<T, CM>(a, hx, hy, hz, ex, ey, ez, col + (i64)k0 * plane, px);
    for (int k = k0; k < k1; ++k) {
        const i64 c = col + (i64)k * plane;
        const EPlane cur = nxt;
        if (k + 1 < k1) nxt = e_load<T, CM>(a, hx, hy, hz, ex, ey, ez, c + plane, px);
        const float hxc = cur.hxc, hyc = cur.hyc, hzc = cur.hzc;
        const float hy_im = cur.hy_im, hz_im = cur.hz_im;
        const float hx_jm = cur.hx_jm, hz_jm = cur.hz_jm;
        if (CM != CLS_UNIFORM) {
            cx = a.etab[cur.cl * 3 + 0];
            cy = a.etab[cur.cl * 3 + 1];
            cz = a.etab[cur.cl * 3 + 2];
        }
        const bool klo = k > 0, khi = k < nz - 1;
        const bool upd_x = ihi && jlo && jhi && klo && khi;
        const bool upd_y = ilo && ihi && jhi && klo && khi;
        const bool upd_z = ilo && ihi && jlo && jhi && khi;
        const bool zs = (klo && k < a.L[4]) || (k >= nz - a.L[5] && khi);   // fdtdrav.h:853,879

        // curl coefficients, divided by kappa on the slab cells (setABCPML)
        float cbx_hz = cx.y, cbx_hy = cx.z;
        float cby_hx = cy.y, cby_hz = cy.z;
        float cbz_hy = cz.y, cbz_hx = cz.z;
        float4 pze = make_float4(0.f, 0.f, 1.f, 0.f);
        if (xs) { cby_hz = cby_hz / pxe.z; cbz_hy = cbz_hy / pxe.z; }
        if (ys) { cbx_hz = cbx_hz / pye.z; cbz_hx = cbz_hx / pye.z; }
        if (zs) {
            pze = a.pz_prof[k];
            cbx_hy = cbx_hy / pze.z;
            cby_hx = cby_hx / pze.z;
        }

        const float ex0 = cur.ex0, ey0 = cur.ey0, ez0 = cur.ez0;
        float exn = ex0, eyn = ey0, ezn = ez0;
        if (upd_x) exn = (cx.x * ex0 + cbx_hz * (hzc - hz_jm)) + cbx_hy * (hyc - hy_km);
        if (upd_y) eyn = (cy.x * ey0 + cby_hx * (hxc - hx_km)) + cby_hz * (hzc - hz_im);
        if (upd_z) ezn = (cz.x * ez0 + cbz_hy * (hyc - hy_im)) + cbz_hx * (hxc - hx_jm);

        if (k == ek) {   // source edges and probes on this cell
            while (ev >= 0 && a.events[ev].k == k && a.events[ev].col == mycol) {
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
            ek = (a.events[ev].col == mycol) ? a.events[ev].k : INT_MAX;
        }

        // CPML psi recursions and corrections, fdtdrav.h:750-911
        if (xs && jhi && khi) {
            const i64 q = ((i64)k * ny + j) * a.Lx + sx;
            T *pa = (T *)a.psi_eyx + q;
            T *pb = (T *)a.psi_ezx + q;
            const float pey = pxe.x * ld(pa) + pxe.y * (hzc - hz_im);
            const float pez = pxe.x * ld(pb) + pxe.y * (hyc - hy_im);
            st(pa, pey);
            st(pb, pez);
            if (klo) eyn = eyn + (cy.z * dx) * pey;
            if (jlo) ezn = ezn + (cz.y * dx) * pez;
        }
        if (ys && ihi && khi) {
            const i64 q = ((i64)k * a.Ly + sy) * px + i;
            T *pa = (T *)a.psi_exy + q;
            T *pb = (T *)a.psi_ezy + q;
            const float pex = pye.x * ld(pa) + pye.y * (hzc - hz_jm);
            const float pez = pye.x * ld(pb) + pye.y * (hxc - hx_jm);
            st(pa, pex);
            st(pb, pez);
            if (klo) exn = exn + (cx.y * dy) * pex;
            if (ilo) ezn = ezn + (cz.z * dy) * pez;
        }
        if (zs && ihi && jhi) {
            const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
            const i64 q = ((i64)sz * ny + j) * px + i;
            T *pa = (T *)a.psi_exz + q;
            T *pb = (T *)a.psi_eyz + q;
            const float pex = pze.x * ld(pa) + pze.y * (hyc - hy_km);
            const float pey = pze.x * ld(pb) + pze.y * (hxc - hx_km);
            st(pa, pex);
            st(pb, pey);
            if (jlo) exn = exn + (cx.z * dz) * pex;
            if (ilo) eyn = eyn + (cy.y * dz) * pey;
        }
        if (upd_x) stn(ex + c, exn);
        if (upd_y) stn(ey + c, eyn);
        if (upd_z) stn(ez + c, ezn);
        if (in_snap && k >= a.sb[2] && k < a.sb[2] + a.sb[5]) {
            // stored-precision value, as the next step reads it, in SI units
            const i64 nb = (i64)a.sb[3] * a.sb[4] * a.sb[5];
            const i64 q = (i - a.sb[0]) + (i64)a.sb[3] * ((j - a.sb[1]) + (i64)a.sb[4] * (k - a.sb[2]));
            float *o = a.snap + (i64)(m + 1 - a.snap_base) * 3 * nb + q;
            o[0] = ld(ex + c) * a.probe_unscale[0];
            o[nb] = ld(ey + c) * a.probe_unscale[1];
            o[2 * nb] = ld(ez + c) * a.probe_unscale[2];
        }
        hx_km = hxc;
        hy_km = hyc;
    }
}

// ---- H half-step ----------------------------------------------------------
template <class T, int CM>
__global__ void __launch_bounds__(TX * BY) h_kernel(const KArgs a, const int m)
{
    int tx, ty, tz;
    tile_of(a, tx, ty, tz);
    const int i = tx * TX + (int)(threadIdx.x & (TX - 1));
    const int j = ty * BY + (int)(threadIdx.x / TX);
    const int nx = a.nx, ny = a.ny, nz = a.nz, px = a.px;
    if (i >= nx || j >= ny) return;
    const int k0 = tz * a.kchunk;
    const int k1 = min(nz, k0 + a.kchunk);
    const i64 plane = a.plane;
    const i64 col = (i64)j * px + i;

    const T *__restrict__ ex = (const T *)a.ex;
    const T *__restrict__ ey = (const T *)a.ey;
    const T *__restrict__ ez = (const T *)a.ez;
    T *__restrict__ hx = (T *)a.hx;
    T *__restrict__ hy = (T *)a.hy;
    T *__restrict__ hz = (T *)a.hz;

    const bool ihi = i < nx - 1, jhi = j < ny - 1;
    const bool xs = i < a.L[0] || (i >= nx - a.L[1] && ihi);   // H x-slab, fdtdrav.h:402,424
    const bool ys = j < a.L[2] || (j >= ny - a.L[3] && jhi);   // H y-slab, fdtdrav.h:447,471
    const int sx = i < a.L[0] ? i : i - (nx - a.L[1]) + a.L[0];
    const int sy = j < a.L[2] ? j : j - (ny - a.L[3]) + a.L[2];
    const float4 pxm = xs ? a.px_prof[nx + i] : make_float4(0.f, 0.f, 1.f, 0.f);
    const float4 pym = ys ? a.py_prof[ny + j] : make_float4(0.f, 0.f, 1.f, 0.f);
    const float dx = a.d[0], dy = a.d[1], dz = a.d[2];

    float4 cx = a.htab[0], cy = a.htab[1], cz = a.htab[2];

    const int mycol = j * nx + i;
    int ev = first_event(a, a.h_first, mycol, k0);
    int ek = ev >= 0 ? a.events[ev].k : INT_MAX;

    float ex_c = ld(ex + col + (i64)k0 * plane);
    float ey_c = ld(ey + col + (i64)k0 * plane);
    HPlane nxt = h_load<T, CM>(a, ex, ey, ez, hx, hy, hz, col + (i64)k0 * plane, plane, px);

    for (int k = k0; k < k1; ++k) {
        const i64 c = col + (i64)k * plane;
        const HPlane cur = nxt;
        if (k + 1 < k1) nxt = h_load<T, CM>(a, ex, ey, ez, hx, hy, hz, c + plane, plane, px);
        const float ex_kp = cur.ex_kp, ey_kp = cur.ey_kp;
        const float ezc = cur.ezc;
        const float ez_ip = cur.ez_ip, ey_ip = cur.ey_ip;
        const float ez_jp = cur.ez_jp, ex_jp = cur.ex_jp;
        if (CM != CLS_UNIFORM) {
            cx = a.htab[cur.cl * 3 + 0];
            cy = a.htab[cur.cl * 3 + 1];
            cz = a.htab[cur.cl * 3 + 2];
        }
        const bool khi = k < nz - 1;
        const bool upd_x = jhi && khi;
        const bool upd_y = ihi && khi;
        const bool upd_z = ihi && jhi;
        const bool zs = k < a.L[4] || (k >= nz - a.L[5] && khi);   // fdtdrav.h:500,531

        float chx_ey = cx.y, chx_ez = cx.z;
        float chy_ez = cy.y, chy_ex = cy.z;
        float chz_ex = cz.y, chz_ey = cz.z;
        float4 pzm = make_float4(0.f, 0.f, 1.f, 0.f);
        if (xs) { chy_ez = chy_ez / pxm.z; chz_ey = chz_ey / pxm.z; }
        if (ys) { chx_ez = chx_ez / pym.z; chz_ex = chz_ex / pym.z; }
        if (zs) {
            pzm = a.pz_prof[nz + k];
            chx_ey = chx_ey / pzm.z;
            chy_ex = chy_ex / pzm.z;
        }

        const float hx0 = cur.hx0, hy0 = cur.hy0, hz0 = cur.hz0;
        float hxn = hx0, hyn = hy0, hzn = hz0;
        if (upd_x) hxn = (cx.x * hx0 + chx_ey * (ey_kp - ey_c)) + chx_ez * (ez_jp - ezc);
        if (upd_y) hyn = (cy.x * hy0 + chy_ez * (ez_ip - ezc)) + chy_ex * (ex_kp - ex_c);
        if (upd_z) hzn = (cz.x * hz0 + chz_ex * (ex_jp - ex_c)) + chz_ey * (ey_ip - ey_c);

        if (k == ek) {   // probes on this cell
            while (ev >= 0 && a.events[ev].k == k && a.events[ev].col == mycol) {
                const Event e = a.events[ev];
                const int comp = e.kind - EV_PROBE;
                if (a.rec) record_probe(a, e.id, comp, comp == 3 ? hxn : comp == 4 ? hyn : hzn, m);
                ++ev;
            }
            ek = (a.events[ev].col == mycol) ? a.events[ev].k : INT_MAX;
        }

        if (xs) {
            const i64 q = ((i64)k * ny + j) * a.Lx + sx;
            T *pa = (T *)a.psi_hyx + q;
            T *pb = (T *)a.psi_hzx + q;
            const float phy = pxm.x * ld(pa) + pxm.y * (ez_ip - ezc);
            const float phz = pxm.x * ld(pb) + pxm.y * (ey_ip - ey_c);
            st(pa, phy);
            st(pb, phz);
            if (khi) hyn = hyn + (cy.y * dx) * phy;
            if (jhi) hzn = hzn + (cz.z * dx) * phz;
        }
        if (ys) {
            const i64 q = ((i64)k * a.Ly + sy) * px + i;
            T *pa = (T *)a.psi_hxy + q;
            T *pb = (T *)a.psi_hzy + q;
            const float phx = pym.x * ld(pa) + pym.y * (ez_jp - ezc);
            const float phz = pym.x * ld(pb) + pym.y * (ex_jp - ex_c);
            st(pa, phx);
            st(pb, phz);
            if (khi) hxn = hxn + (cx.z * dy) * phx;
            if (ihi) hzn = hzn + (cz.y * dy) * phz;
        }
        if (zs) {
            const int sz = k < a.L[4] ? k : k - (nz - a.L[5]) + a.L[4];
            const i64 q = ((i64)sz * ny + j) * px + i;
            T *pa = (T *)a.psi_hxz + q;
            T *pb = (T *)a.psi_hyz + q;
            const float phx = pzm.x * ld(pa) + pzm.y * (ey_kp - ey_c);
            const float phy = pzm.x * ld(pb) + pzm.y * (ex_kp - ex_c);
            st(pa, phx);
            st(pb, phy);
            if (jhi) hxn = hxn + (cx.y * dz) * phx;
            if (ihi) hyn = hyn + (cy.z * dz) * phy;
        }
        if (upd_x) stn(hx + c, hxn);
        if (upd_y) stn(hy + c, hyn);
        if (upd_z) stn(hz + c, hzn);
        ex_c = ex_kp;
        ey_c = ey_kp;
    }
}

// ---- launchers --------------------------------------------------------------
template <class T>
static hipError_t launch_t(bool efield, int mode, const KArgs &a, int m, hipStream_t s)
{
    const int nt = a.tiles_x * a.tiles_y * ((a.nz + a.kchunk - 1) / a.kchunk);
    dim3 grid(nt), block(TX * BY);
    if (efield) {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((e_kernel<T, CLS_UNIFORM>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((e_kernel<T, CLS_U16>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((e_kernel<T, CLS_U32>), grid, block, 0, s, a, m);
    } else {
        if (mode == CLS_UNIFORM) hipLaunchKernelGGL((h_kernel<T, CLS_UNIFORM>), grid, block, 0, s, a, m);
        else if (mode == CLS_U16) hipLaunchKernelGGL((h_kernel<T, CLS_U16>), grid, block, 0, s, a, m);
        else hipLaunchKernelGGL((h_kernel<T, CLS_U32>), grid, block, 0, s, a, m);
    }
    return hipGetLastError();
}

// storage: 0 FP32, 1 FP16, 2 BF16.  e_mode/h_mode: ClassMode of each sweep.
hipError_t launch_half_step(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s)
{
    if (storage == 1) return launch_t<_Float16>(efield, mode, a, m, s);
    if (storage == 2) return launch_t<bf16_t>(efield, mode, a, m, s);
    return launch_t<float>(efield, mode, a, m, s);
}

int tile_rows() { return BY; }
int tile_cols() { return TX; }

}  // namespace fdtd
