// port.cpp -- port voltage and current of a lumped port, and its spectrum.
//
// The reference's output is the sampled port voltage and current of every
// step, written as results1003.dat / results1004.dat (fdtdrav.h:4362-4367):
//   voltMgpu   fdtdrav.h:2185-2298   V = -d / (cross-section count) * sum E
//   currMgpu   fdtdrav.h:2301-2487   I = loop integral of H around the box
//   correcI    fdtdrav.h:2490-2507   I[m] <- mean of neighbouring samples
//   c_zinS11   fdtdrav.h:2510-2546   a, b, S11 = b/a, Zin = z0 (1+S11)/(1-S11)
//   saveTDomain fdtdrav.h:2703-2717  one "%e \n" line per step
// The reference samples those fields into full per-step boxes
// (sampleField, fdtdrav.h:234-246) and reduces them after the run.  Here
// each field value a reduction reads is an ordinary probe (sampled at the
// same point of the update), and the reduction runs on the host over the
// probe histories, in the reference's summation order and in FP32, so the
// series match the reference's arithmetic.  The spectrum is a direct FP64
// DFT of the series at the caller's frequencies (at f = k / (n dt) it equals
// the R2C FFT bin k that calc_otm, fdtdrav.h:2988-3051, would take).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fdtd.h"
#include "internal.h"

namespace fdtd {

namespace {

void add(PortPlan &p, std::vector<int> &list, int comp, int i, int j, int k)
{
    list.push_back((int)p.samples.size());
    p.samples.push_back(Probe{comp, i, j, k});
}

bool inside(int i, int j, int k, int nx, int ny, int nz)
{
    return i >= 0 && j >= 0 && k >= 0 && i < nx && j < ny && k < nz;
}

// sum of the listed samples at step m, in list order (FP32, as the
// reference's loops accumulate)
float line_sum(const std::vector<int> &list, const float *hist, size_t ns, size_t m)
{
    float s = 0.f;
    for (int q : list) s = s + hist[(size_t)q * ns + m];
    return s;
}

}  // namespace

int port_plan(const PortDef &d, int nx, int ny, int nz, float dx, float dy, float dz, PortPlan &p,
              std::string &err)
{
    p = PortPlan{};
    p.def = d;
    const int *v = d.vbox, *c = d.ibox;
    // voltage: voltMgpu case 3 (fdtdrav.h:2252-2281): Ez over
    // is<=i<=ie, js<=j<=je, ks<=k<ke, loops k, j, i.  The reference samples
    // Ez only (sampleField, fdtdrav.h:740), so its x and y voltage cases sum
    // Ez over an x or y box; they are not offered.
    if (d.vdir != 3) {
        err = "port voltage direction must be 3 (z): the reference samples only Ez";
        return FDTD_E_UNSUPPORTED;
    }
    if (v[0] > v[3] || v[1] > v[4] || v[2] >= v[5] || !inside(v[0], v[1], v[2], nx, ny, nz) ||
        !inside(v[3], v[4], v[5] - 1, nx, ny, nz)) {
        err = "port voltage box outside the grid or empty";
        return FDTD_E_RANGE;
    }
    for (int k = v[2]; k < v[5]; ++k)
        for (int j = v[1]; j <= v[4]; ++j)
            for (int i = v[0]; i <= v[3]; ++i) add(p, p.v, FDTD_EZ, i, j, k);
    // csuf, with the reference's float extents (fdtdrav.h:2257)
    const float ext = ((float)v[3] - (float)v[0] + 1.f) * ((float)v[4] - (float)v[1] + 1.f);
    p.vscale = (-dz) / ext;

    // current: currMgpu, the loop of H around the box.  Local index 0 of the
    // sampled H box is is-1 (js-1, ks-1), so the reference's local ranges
    // 1..N-1 are is..ie (js..je) and local N-1 is ie (je).
    const int is = c[0], js = c[1], ks = c[2], ie = c[3], je = c[4], ke = c[5];
    if (is > ie || js > je || ks > ke || !inside(is - 1, js - 1, ks - 1, nx, ny, nz) || !inside(ie, je, ke, nx, ny, nz)) {
        err = "port current box (with its ring of H one cell below) outside the grid";
        return FDTD_E_RANGE;
    }
    switch (d.idir) {
    case 1: {   // fdtdrav.h:2329-2380: plane i = ie-1, y sums at k = ks-1 / ke, z sums over ks..ke
        const int ip = ie - 1;
        if (ip < 0) { err = "port current box: x-directed loop needs ie >= 1"; return FDTD_E_RANGE; }
        for (int j = js; j <= je; ++j) add(p, p.ia, FDTD_HY, ip, j, ks - 1);   // sum_yp
        for (int k = ks; k <= ke; ++k) add(p, p.ib, FDTD_HZ, ip, je, k);       // sum_zp
        for (int j = js; j <= je; ++j) add(p, p.ic, FDTD_HY, ip, j, ke);       // sum_yn
        for (int k = ks; k <= ke; ++k) add(p, p.id, FDTD_HZ, ip, js - 1, k);   // sum_zn
        p.ca = dy; p.cb = dz; p.cc = dy; p.cd = dz;
        break;
    }
    case 2: {   // fdtdrav.h:2382-2431: plane j = je-1, z sums over ks..ke, x sums at k = ke / ks-1
        const int jp = je - 1;
        if (jp < 0) { err = "port current box: y-directed loop needs je >= 1"; return FDTD_E_RANGE; }
        for (int k = ks; k <= ke; ++k) add(p, p.ia, FDTD_HZ, is - 1, jp, k);   // sum_zp
        for (int i = is; i <= ie; ++i) add(p, p.ib, FDTD_HX, i, jp, ke);       // sum_xp
        for (int k = ks; k <= ke; ++k) add(p, p.ic, FDTD_HZ, ie, jp, k);       // sum_zn
        for (int i = is; i <= ie; ++i) add(p, p.id, FDTD_HX, i, jp, ks - 1);   // sum_xn
        p.ca = dz; p.cb = dx; p.cc = dz; p.cd = dx;
        break;
    }
    case 3: {   // fdtdrav.h:2433-2481: plane k = ke-1
        const int kp = ke - 1;
        if (kp < 0) { err = "port current box: z-directed loop needs ke >= 1"; return FDTD_E_RANGE; }
        for (int j = js; j <= je; ++j) add(p, p.ia, FDTD_HY, ie, j, kp);       // sum_yp
        for (int i = is; i <= ie; ++i) add(p, p.ib, FDTD_HX, i, js - 1, kp);   // sum_xp
        for (int j = js; j <= je; ++j) add(p, p.ic, FDTD_HY, is - 1, j, kp);   // sum_yn
        for (int i = is; i <= ie; ++i) add(p, p.id, FDTD_HX, i, je, kp);       // sum_xn
        p.ca = dy; p.cb = dx; p.cc = dy; p.cd = dx;
        break;
    }
    default:
        err = "port current direction must be 1, 2 or 3";
        return FDTD_E_INVALID;
    }
    return FDTD_OK;
}

void port_series(const PortPlan &p, const float *hist, size_t ns, float *v, float *cur)
{
    std::vector<float> raw(ns);
    for (size_t m = 0; m < ns; ++m) {
        v[m] = line_sum(p.v, hist, ns, m) * p.vscale;
        // d_current_tran = ca*sa + cb*sb - cc*sc - cd*sd, left to right
        raw[m] = ((p.ca * line_sum(p.ia, hist, ns, m) + p.cb * line_sum(p.ib, hist, ns, m)) -
                  p.cc * line_sum(p.ic, hist, ns, m)) -
                 p.cd * line_sum(p.id, hist, ns, m);
    }
    // correcI with the samples as they were before the pass (the reference's
    // in-place update races across warps; its per-warp order reads old values)
    for (size_t m = 0; m < ns; ++m) {
        if (ns == 1) cur[m] = raw[m];
        else if (m + 1 < ns) cur[m] = (raw[m] + raw[m + 1]) / 2.f;
        else cur[m] = (raw[m] + raw[m - 1]) / 2.f;
    }
}

void port_spectrum(const float *v, const float *cur, size_t ns, double dt, double z0, int nfreq,
                   const double *freqs, double *out)
{
    const double two_pi = 6.283185307179586476925286766559;
    const std::complex<double> z(z0, 0.0), one(1.0, 0.0);
    for (int q = 0; q < nfreq; ++q) {
        std::complex<double> V(0.0, 0.0), I(0.0, 0.0);
        for (size_t m = 0; m < ns; ++m) {
            const double ph = -two_pi * freqs[q] * (double)m * dt;
            const std::complex<double> e(std::cos(ph), std::sin(ph));
            V += (double)v[m] * e;
            I += (double)cur[m] * e;
        }
        const double den = 2.0 * std::sqrt(z0);
        const std::complex<double> a = (V + z * I) / den, b = (V - z * I) / den;
        const std::complex<double> s11 = b / a;
        const std::complex<double> zin = z * ((one + s11) / (one - s11));
        double *o = out + 8 * (size_t)q;
        o[0] = V.real(); o[1] = V.imag();
        o[2] = I.real(); o[3] = I.imag();
        o[4] = s11.real(); o[5] = s11.imag();
        o[6] = zin.real(); o[7] = zin.imag();
    }
}

int port_save(const float *x, size_t ns, const char *path)
{
    FILE *f = std::fopen(path, "w");
    if (!f) return FDTD_E_INVALID;
    for (size_t m = 0; m < ns; ++m) std::fprintf(f, "%e \n", (double)x[m]);
    return std::fclose(f) == 0 ? FDTD_OK : FDTD_E_INVALID;
}

}  // namespace fdtd
