// scene.cpp -- host-side build of everything the gfx950 kernels read:
// coefficient classes, CPML profiles, source edges, probe events, waveform
// and DFT-phasor tables.
//
// The reference keeps 84 full-grid FP32 arrays (fdtdrav.h:3886-4014), 18 of
// them update coefficients.  Here the coefficients are derived once per cell
// on the host with the reference's formulas and deduplicated by value into a
// small class table; the device keeps one class index per cell (or none, when
// every cell shares one class).  The formulas and their FP32 operation order
// follow the reference:
//   edge averaging      setSigmaEpsSigmaMu   fdtdrav.h:1127-1214
//   coefficients        setCoefficients      fdtdrav.h:1247-1305
//   CPML profiles       setABCPML            fdtdrav.h:1774-2155
//   source loading      defineSourceZ        fdtdrav.h:1391-1421
//   conducting plates   definePlateX/Y/Z     fdtdrav.h:1574-1678
//   lumped resistors    defineResistor       fdtdrav.h:1429-1484
//   conducting sectors  defineSec, tricheck  fdtdrav.h:1681-1744, 56-128
//   x-directed source   defineSourceX        fdtdrav.h:1311-1366 (loaded Ex
//                       edges; calc_e's Ex injection is commented out, fdtdrav.h:681)
//   per-node waveform   setGauss / setSine   fdtdrav.h:1486-1520
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <math.h>   // sincos (glibc)
#include <cstring>
#include <map>
#include <unordered_map>

#include "internal.h"

namespace fdtd {

namespace {

struct Mat {
    float v[12];  // eps_r xyz, sigma_e xyz, mu_r xyz, sigma_m xyz
};

// Waveform value at time t (FP64, rounded by the caller).
double waveform(int type, double amp, double t0, double tau, double f, double t)
{
    if (type == 1) {
        double u = (t - t0) / tau;
        return amp * std::exp(-(u * u));
    }
    if (type == 2) {
        double u = (t - t0) / tau;
        return amp * std::sqrt(2.0 * std::exp(1.0)) * u * std::exp(-(u * u));
    }
    if (type == 3) return amp * std::sin(2.0 * M_PI * f * t);
    return 0.0;
}

// 4-cell arithmetic mean, reference operand order (fdtdrav.h:1140-1177).
inline float avg4(float a, float b, float c, float d) { return 0.25f * (((a + b) + c) + d); }

// 2-cell harmonic mean (fdtdrav.h:1185-1213).  The reference keeps a+b
// nonzero by initialising sigma to 1.2e-38, where 2ab underflows to 0.
inline float harm(float a, float b)
{
    float den = a + b;
    if (den == 0.0f) return 0.0f;
    return (2.0f * a * b) / den;
}

struct Consts {
    float dx, dy, dz, dt, eps0, mu0;
};

// tricheck (fdtdrav.h:56-128): 0 when the point (px, py) lies in the sector
// with apex (x1, y1) between the rays towards (x0, y0) and (x2, y2), within
// radius r of the apex; -1 otherwise.  As there, the coordinate differences
// and the squared distance are float operations widened to double, and the
// two side tests are products of double cross products.
inline int sector_test(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float r)
{
    const double ax = x1 - x0, ay = y1 - y0, bx = x2 - x1, by = y2 - y1;
    const double cx = x2 - x0, cy = y2 - y0, ex = x0 - x1, ey = y0 - y1;
    const double zx1 = px - x0, zy1 = py - y0, zx2 = px - x1, zy2 = py - y1;
    const double dist = std::sqrt((double)(((x1 - px) * (x1 - px)) + ((y1 - py) * (y1 - py))));
    auto cross = [](double ux, double uy, double vx, double vy) { return (ux * vy) - (uy * vx); };
    if (cross(ax, ay, cx, cy) * cross(ax, ay, zx1, zy1) < 0) return -1;
    if (cross(bx, by, ex, ey) * cross(bx, by, zx2, zy2) < 0) return -1;
    if (dist > (double)r) return -1;
    return 0;
}

// (ce, cb_a, cb_b) of one E component (fdtdrav.h:1271-1305).
inline void ecoef(const Consts &c, float eps, float sig, float da, float db, float *o)
{
    const float two = 2.0f;
    float den = two * eps * c.eps0 + c.dt * sig;
    o[0] = (two * eps * c.eps0 - c.dt * sig) / den;
    o[1] = (two * c.dt) / (den * da);
    o[2] = (-2.0f * c.dt) / (den * db);
}

// (ch, cb_a, cb_b) of one H component (fdtdrav.h:1247-1268).
inline void hcoef(const Consts &c, float mu, float sig, float da, float db, float *o)
{
    const float two = 2.0f;
    float den = two * mu * c.mu0 + c.dt * sig;
    o[0] = (two * mu * c.mu0 - c.dt * sig) / den;
    o[1] = (two * c.dt) / (den * da);
    o[2] = (-2.0f * c.dt) / (den * db);
}

struct KeyHash {
    size_t operator()(const std::array<std::uint32_t, 13> &k) const
    {
        std::uint64_t h = 1469598103934665603ull;
        for (std::uint32_t v : k) { h ^= v; h *= 1099511628211ull; h ^= h >> 29; }
        return (size_t)h;
    }
};
struct TupHash {
    size_t operator()(const std::array<std::uint32_t, 18> &k) const
    {
        std::uint64_t h = 1469598103934665603ull;
        for (std::uint32_t v : k) { h ^= v; h *= 1099511628211ull; h ^= h >> 31; }
        return (size_t)h;
    }
};

}  // namespace

int build_scene(const SceneIn &in, Built &out, std::string &err)
{
    const int nx = in.nx, ny = in.ny, nz = in.nz, px = in.px;
    const i64 nxy = (i64)nx * ny, N = nxy * nz;
    const Consts C{(float)in.dx, (float)in.dy, (float)in.dz, (float)in.dt, (float)in.eps0, (float)in.mu0};

    // ---- raw materials and the label grid (setZero + setBrick order)
    std::vector<Mat> mats;
    mats.push_back(Mat{{1, 1, 1, 0, 0, 0, 1, 1, 1, 0, 0, 0}});
    std::vector<std::uint32_t> lab((size_t)N, 0u);
    for (size_t b = 0; b < in.boxes.size(); ++b) {
        const Box &bx = in.boxes[b];
        Mat m;
        std::memcpy(m.v, bx.mat, sizeof m.v);
        mats.push_back(m);
        const std::uint32_t id = (std::uint32_t)(mats.size() - 1);
        const int i0 = std::max(0, bx.b[0]), i1 = std::min(nx, bx.b[3]);
        const int j0 = std::max(0, bx.b[1]), j1 = std::min(ny, bx.b[4]);
        const int k0 = std::max(0, bx.b[2]), k1 = std::min(nz, bx.b[5]);
        for (int k = k0; k < k1; ++k)
            for (int j = j0; j < j1; ++j)
                for (int i = i0; i < i1; ++i) lab[(size_t)(i + nx * (j + (i64)ny * k))] = id;
    }
    {
        const int *d = in.design;
        if (!in.design_eps.empty()) {
            const int dnx = d[3] - d[0], dny = d[4] - d[1];
            for (int k = d[2]; k < d[5]; ++k)
                for (int j = d[1]; j < d[4]; ++j)
                    for (int i = d[0]; i < d[3]; ++i) {
                        size_t x = (size_t)(i + nx * (j + (i64)ny * k));
                        Mat m = mats[lab[x]];
                        float e = in.design_eps[(size_t)((i - d[0]) + dnx * ((j - d[1]) + (i64)dny * (k - d[2])))];
                        m.v[0] = m.v[1] = m.v[2] = e;
                        mats.push_back(m);
                        lab[x] = (std::uint32_t)(mats.size() - 1);
                    }
        }
    }
    if (mats.size() >= 0xFFFFFFF0u) { err = "too many materials"; return -1; }

    // ---- plates: averaged sigma_e overrides of the edges in each plate, in
    // call order (definePlateX/Y/Z, fdtdrav.h:1574-1678, applied between
    // setSigmaEpsSigmaMu and setCoefficients, fdtdrav.h:4183-4210)
    std::unordered_map<i64, float> sig_over[3];
    // sectors first: defineSec (fdtdrav.h:1681-1744) on the node points
    // (i*dx, j*dy) of the plane kz, in its marchingLoop slot before the plates
    for (const Sector &sc : in.sectors) {
        const float dx = C.dx, dy = C.dy;
        const float x0 = (float)sc.p[0] * dx, y0 = (float)sc.p[1] * dy, x1 = (float)sc.p[2] * dx;
        const float y1 = (float)sc.p[3] * dy, x2 = (float)sc.p[4] * dx, y2 = (float)sc.p[5] * dy;
        const i64 kb = (i64)ny * sc.kz;
        for (int j = 0; j <= sc.max_y; ++j)
            for (int i = 0; i <= sc.max_x; ++i) {
                if (sector_test((float)i * dx, (float)j * dy, x0, y0, x1, y1, x2, y2, sc.radius) != 0) continue;
                const i64 cell = i + nx * (j + kb);
                if (i < sc.max_x) sig_over[0][cell] = sc.sigma[0];
                if (j < sc.max_y) sig_over[1][cell] = sc.sigma[1];
            }
    }
    for (const Plate &pl : in.plates) {
        const int *b = pl.b;
        auto put = [&](int comp, int i0, int i1, int j0, int j1, int k0, int k1) {   // inclusive ranges
            for (int k = k0; k <= k1; ++k)
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i) sig_over[comp][i + nx * (j + (i64)ny * k)] = pl.sigma[comp];
        };
        if (pl.axis == 1) {          // plane i = is
            put(1, b[0], b[0], b[1], b[4] - 1, b[2], b[5]);
            put(2, b[0], b[0], b[1], b[4], b[2], b[5] - 1);
        } else if (pl.axis == 2) {   // plane j = js
            put(2, b[0], b[3], b[1], b[1], b[2], b[5] - 1);
            put(0, b[0], b[3] - 1, b[1], b[1], b[2], b[5]);
        } else {                     // plane k = ks
            put(0, b[0], b[3] - 1, b[1], b[4], b[2], b[2]);
            put(1, b[0], b[3], b[1], b[4] - 1, b[2], b[2]);
        }
    }
    auto edge_sigma = [&](int comp, size_t x, float avg) {
        if (sig_over[comp].empty()) return avg;
        auto it = sig_over[comp].find((i64)x);
        return it == sig_over[comp].end() ? avg : it->second;
    };
    // ---- resistors: Ez edges with the resistor-loaded coefficients
    // (defineResistor, fdtdrav.h:1429-1484; after the source and the CPML,
    // so a resistor edge keeps its curl coefficients undivided by kappa --
    // resistors are therefore kept off the CPML slabs)
    std::unordered_map<i64, int> res_of;
    {
        const int *L = in.cpml.thick;
        for (size_t r = 0; r < in.resistors.size(); ++r) {
            const int *b = in.resistors[r].b;
            if (b[0] < L[0] || b[3] >= nx - L[1] || b[1] < L[2] || b[4] >= ny - L[3] || b[2] < L[4] ||
                b[5] > nz - L[5]) {
                err = "resistor inside a CPML slab";
                return -1;
            }
            for (int k = b[2]; k < b[5]; ++k)
                for (int j = b[1]; j <= b[4]; ++j)
                    for (int i = b[0]; i <= b[3]; ++i) res_of[i + nx * (j + (i64)ny * k)] = (int)r;
        }
        for (const Source &sr : in.sources)
            for (int k = sr.b[2]; k < sr.b[5]; ++k)
                for (int j = sr.b[1]; j <= sr.b[4]; ++j)
                    for (int i = sr.b[0]; i <= sr.b[3]; ++i)
                        if (res_of.count(i + nx * (j + (i64)ny * k))) {
                            err = "a resistor overlaps a lumped source";
                            return -1;
                        }
        if (!in.design_eps.empty() && (!in.plates.empty() || !in.sectors.empty() || !in.resistors.empty())) {
            // the adjoint's edge derivatives assume the averaged materials
            const int *d = in.design;
            auto near = [&](i64 cell) {
                const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / nxy);
                return i >= d[0] - 1 && i <= d[3] && j >= d[1] - 1 && j <= d[4] && k >= d[2] - 1 && k <= d[5];
            };
            for (int c = 0; c < 3; ++c)
                for (const auto &kv : sig_over[c])
                    if (near(kv.first)) { err = "a plate or sector touches the design region"; return -1; }
            for (const auto &kv : res_of)
                if (near(kv.first)) { err = "a resistor touches the design region"; return -1; }
        }
    }
    // ---- x-directed sources: Ex edges with defineSourceX's loaded
    // coefficients (fdtdrav.h:1311-1366; its j range ends at ie), after the
    // CPML like the z-sources -- kept off the CPML slabs for the same reason
    // as resistors
    std::unordered_map<i64, int> xs_of;
    {
        const int *L = in.cpml.thick;
        for (size_t q = 0; q < in.sources.size(); ++q) {
            if (in.sources[q].dir != 1) continue;
            const int *b = in.sources[q].b;
            if (b[0] < L[0] || b[3] > nx - L[1] - 1 || b[1] < L[2] || std::max(b[3], b[4]) > ny - L[3] - 1 ||
                b[2] < L[4] || b[5] > nz - L[5] - 1) {
                err = "x-directed source inside a CPML slab";
                return -1;
            }
            for (int k = b[2]; k <= b[5]; ++k)
                for (int j = b[1]; j <= b[3]; ++j)
                    for (int i = b[0]; i < b[3]; ++i) {
                        const i64 cell = i + nx * (j + (i64)ny * k);
                        if (xs_of.count(cell)) { err = "x-directed sources overlap"; return -1; }
                        xs_of[cell] = (int)q;
                    }
        }
        if (!in.design_eps.empty()) {
            const int *d = in.design;
            for (const auto &kv : xs_of) {
                const i64 cell = kv.first;
                const int i = (int)(cell % nx), j = (int)((cell / nx) % ny), k = (int)(cell / nxy);
                if (i >= d[0] - 1 && i <= d[3] && j >= d[1] - 1 && j <= d[4] && k >= d[2] - 1 && k <= d[5]) {
                    err = "an x-directed source touches the design region";
                    return -1;
                }
            }
        }
    }
    auto xcoef = [&](const Source &xs, float e, float sg, float *o) {
        const float _is = (float)xs.b[0], _js = (float)xs.b[1], _ks = (float)xs.b[2];
        const float _ie = (float)xs.b[3], _je = (float)xs.b[4], _ke = (float)xs.b[5];
        float r_comp = ((1.0f + _je) - _js) * ((1.0f + _ke) - _ks) / (_ie - _is);
        r_comp = r_comp * (float)xs.R;
        const float rfactor = (C.dt * C.dx) / (r_comp * C.dz * C.dy);
        const float two = 2.0f;
        const float D = two * e * C.eps0 + C.dt * sg + rfactor;
        o[0] = (two * e * C.eps0 - C.dt * sg - rfactor) / D;
        o[1] = (two * C.dt / C.dy) / D;
        o[2] = (-2.0f * C.dt / C.dz) / D;
    };
    // resistor-loaded Ez coefficients (defineResistor's operand order)
    auto rcoef = [&](const Resistor &rs, float e, float sg, float *o) {
        const float _is = (float)rs.b[0], _js = (float)rs.b[1], _ks = (float)rs.b[2];
        const float _ie = (float)rs.b[3], _je = (float)rs.b[4], _ke = (float)rs.b[5];
        float r_comp = ((1.0f + _ie) - _is) * ((1.0f + _je) - _js) / (_ke - _ks);
        r_comp = r_comp * (float)rs.R;
        const float rfactor = (C.dt * C.dz) / (r_comp * C.dx * C.dy);
        const float two = 2.0f;
        const float D = two * e * C.eps0 + C.dt * sg + rfactor;
        o[0] = (two * e * C.eps0 - C.dt * sg - rfactor) / D;
        o[1] = (two * C.dt / C.dx) / D;
        o[2] = (-2.0f * C.dt / C.dy) / D;
    };

    // ---- per-cell coefficient tuples, deduplicated by value
    // The tuple of a cell depends on 7 labels -- (i,j,k) and its -x, -y, -z,
    // -x-y, -y-z, -x-z neighbours -- on which faces the cell lies on, and on
    // its plate sigma overrides and resistor.  Components not updated at a
    // cell get the vacuum values (never read).
    const std::uint32_t NONE = 0xFFFFFFFFu;
    std::unordered_map<std::array<std::uint32_t, 13>, int, KeyHash> key2cls;
    std::unordered_map<std::array<std::uint32_t, 18>, int, TupHash> tup2cls;
    std::vector<std::array<float, 18>> tuples;
    std::vector<std::uint32_t> cellcls((size_t)N);
    std::array<std::uint32_t, 13> prev;
    prev.fill(NONE - 1);
    int prev_cls = -1;
    float vac[18];
    {
        ecoef(C, 1.0f, 0.0f, C.dy, C.dz, vac + 0);
        ecoef(C, 1.0f, 0.0f, C.dz, C.dx, vac + 3);
        ecoef(C, 1.0f, 0.0f, C.dx, C.dy, vac + 6);
        hcoef(C, 1.0f, 0.0f, C.dz, C.dy, vac + 9);
        hcoef(C, 1.0f, 0.0f, C.dx, C.dz, vac + 12);
        hcoef(C, 1.0f, 0.0f, C.dy, C.dx, vac + 15);
    }
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t x = (size_t)(i + nx * (j + (i64)ny * k));
                std::array<std::uint32_t, 13> key;
                key[0] = lab[x];
                key[1] = i > 0 ? lab[x - 1] : NONE;
                key[2] = j > 0 ? lab[x - nx] : NONE;
                key[3] = k > 0 ? lab[x - nxy] : NONE;
                key[4] = (i > 0 && j > 0) ? lab[x - 1 - nx] : NONE;
                key[5] = (j > 0 && k > 0) ? lab[x - nx - nxy] : NONE;
                key[6] = (i > 0 && k > 0) ? lab[x - 1 - nxy] : NONE;
                key[7] = (std::uint32_t)((i == 0) | (i == nx - 1) << 1 | (j == 0) << 2 | (j == ny - 1) << 3 |
                                         (k == 0) << 4 | (k == nz - 1) << 5);
                for (int c = 0; c < 3; ++c) {
                    key[8 + c] = NONE;
                    if (!sig_over[c].empty()) {
                        auto it = sig_over[c].find((i64)x);
                        if (it != sig_over[c].end()) std::memcpy(&key[8 + c], &it->second, 4);
                    }
                }
                key[11] = 0;
                if (!res_of.empty()) {
                    auto it = res_of.find((i64)x);
                    if (it != res_of.end()) key[11] = (std::uint32_t)it->second + 1u;
                }
                key[12] = 0;
                if (!xs_of.empty()) {
                    auto it = xs_of.find((i64)x);
                    if (it != xs_of.end()) key[12] = (std::uint32_t)it->second + 1u;
                }
                if (key == prev) { cellcls[x] = (std::uint32_t)prev_cls; continue; }
                auto it = key2cls.find(key);
                int cid;
                if (it != key2cls.end()) {
                    cid = it->second;
                } else {
                    float t[18];
                    std::memcpy(t, vac, sizeof t);
                    const Mat &m0 = mats[key[0]];
                    auto M = [&](std::uint32_t l) -> const Mat & { return mats[l]; };
                    // Ex: i<nx-1, 0<j<ny-1, 0<k<nz-1; labels (i,j,k),(j-1),(k-1),(j-1,k-1)
                    if (i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
                        float e = avg4(m0.v[0], M(key[2]).v[0], M(key[3]).v[0], M(key[5]).v[0]);
                        float s = edge_sigma(0, x, avg4(m0.v[3], M(key[2]).v[3], M(key[3]).v[3], M(key[5]).v[3]));
                        if (key[12]) xcoef(in.sources[key[12] - 1], e, s, t + 0);
                        else ecoef(C, e, s, C.dy, C.dz, t + 0);
                    }
                    // Ey: 0<i<nx-1, j<ny-1, 0<k<nz-1; labels (i,j,k),(i-1),(k-1),(i-1,k-1)
                    if (i > 0 && i < nx - 1 && j < ny - 1 && k > 0 && k < nz - 1) {
                        float e = avg4(m0.v[1], M(key[1]).v[1], M(key[3]).v[1], M(key[6]).v[1]);
                        float s = edge_sigma(1, x, avg4(m0.v[4], M(key[1]).v[4], M(key[3]).v[4], M(key[6]).v[4]));
                        ecoef(C, e, s, C.dz, C.dx, t + 3);
                    }
                    // Ez: 0<i<nx-1, 0<j<ny-1, k<nz-1; labels (i,j,k),(i-1),(j-1),(i-1,j-1)
                    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k < nz - 1) {
                        float e = avg4(m0.v[2], M(key[1]).v[2], M(key[2]).v[2], M(key[4]).v[2]);
                        float s = edge_sigma(2, x, avg4(m0.v[5], M(key[1]).v[5], M(key[2]).v[5], M(key[4]).v[5]));
                        if (key[11]) rcoef(in.resistors[key[11] - 1], e, s, t + 6);
                        else ecoef(C, e, s, C.dx, C.dy, t + 6);
                    }
                    // Hx: j<ny-1, k<nz-1; mu_x harmonic over (i, i-1) when i > 0
                    if (j < ny - 1 && k < nz - 1) {
                        float mu = i > 0 ? harm(m0.v[6], M(key[1]).v[6]) : 1.0f;
                        float sg = i > 0 ? harm(m0.v[9], M(key[1]).v[9]) : 0.0f;
                        hcoef(C, mu, sg, C.dz, C.dy, t + 9);
                    }
                    // Hy: i<nx-1, k<nz-1; over (j, j-1) when j > 0
                    if (i < nx - 1 && k < nz - 1) {
                        float mu = j > 0 ? harm(m0.v[7], M(key[2]).v[7]) : 1.0f;
                        float sg = j > 0 ? harm(m0.v[10], M(key[2]).v[10]) : 0.0f;
                        hcoef(C, mu, sg, C.dx, C.dz, t + 12);
                    }
                    // Hz: i<nx-1, j<ny-1; over (k, k-1) when k > 0
                    if (i < nx - 1 && j < ny - 1) {
                        float mu = k > 0 ? harm(m0.v[8], M(key[3]).v[8]) : 1.0f;
                        float sg = k > 0 ? harm(m0.v[11], M(key[3]).v[11]) : 0.0f;
                        hcoef(C, mu, sg, C.dy, C.dx, t + 15);
                    }
                    std::array<std::uint32_t, 18> tb;
                    std::memcpy(tb.data(), t, sizeof t);
                    auto jt = tup2cls.find(tb);
                    if (jt != tup2cls.end()) {
                        cid = jt->second;
                    } else {
                        cid = (int)tuples.size();
                        std::array<float, 18> ta;
                        std::memcpy(ta.data(), t, sizeof t);
                        tuples.push_back(ta);
                        tup2cls.emplace(tb, cid);
                    }
                    key2cls.emplace(key, cid);
                }
                prev = key;
                prev_cls = cid;
                cellcls[x] = (std::uint32_t)cid;
            }
    const int ncls = (int)tuples.size();
    bool e_uni = true, h_uni = true;
    for (int c = 1; c < ncls; ++c) {
        if (std::memcmp(tuples[c].data(), tuples[0].data(), 9 * sizeof(float)) != 0) e_uni = false;
        if (std::memcmp(tuples[c].data() + 9, tuples[0].data() + 9, 9 * sizeof(float)) != 0) h_uni = false;
    }
    out.e_uniform = e_uni;
    out.h_uniform = h_uni;
    if (e_uni && h_uni) {
        out.mode = CLS_UNIFORM;
        out.ncls = 1;
        out.cls.clear();
    } else {
        out.mode = ncls <= 65536 ? CLS_U16 : CLS_U32;
        out.ncls = ncls;
        out.cls.assign((size_t)px * ny * nz, 0u);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                std::memcpy(&out.cls[(size_t)px * (j + (i64)ny * k)], &cellcls[(size_t)nx * (j + (i64)ny * k)],
                            (size_t)nx * sizeof(std::uint32_t));
    }

    // ---- storage scales: stored = SI * scale.  FP32 stores SI values.
    // FP16/BF16 store E/eta0 and H with one common gain chosen so that the
    // source-gap field V/d maps to 2^11, leaving 2^5 of headroom below the
    // FP16 maximum and the normal range for 14 decades below it.
    float sc[6] = {1, 1, 1, 1, 1, 1};
    if (in.storage != 0) {
        double vmax = 0.0;
        for (const Source &s : in.sources) vmax = std::max(vmax, std::fabs(s.amp));
        if (vmax == 0.0) vmax = 1.0;
        const double dmin = std::min(in.dx, std::min(in.dy, in.dz));
        const double eref = vmax / dmin;
        const double eta0 = std::sqrt(in.mu0 / in.eps0);
        const double gain = 2048.0 / (eref / eta0);
        for (int c = 0; c < 3; ++c) sc[c] = (float)(gain / eta0);
        for (int c = 3; c < 6; ++c) sc[c] = (float)gain;
    }
    for (int c = 0; c < 6; ++c) {
        out.scale[c] = sc[c];
        out.probe_unscale[c] = 1.0f / sc[c];
    }
    const bool scaled = in.storage != 0;
    // E coefficient (comp c, slot s) multiplies a difference of H component
    // hcomp[c][s]; H coefficient multiplies a difference of ecomp[c][s].
    const int hcomp[3][2] = {{5, 4}, {3, 5}, {4, 3}};   // Ex: Hz, Hy; Ey: Hx, Hz; Ez: Hy, Hx
    const int ecomp[3][2] = {{1, 2}, {2, 0}, {0, 1}};   // Hx: Ey, Ez; Hy: Ez, Ex; Hz: Ex, Ey
    const int ntab = out.ncls;
    out.etab.assign((size_t)ntab * 12, 0.0f);
    out.htab.assign((size_t)ntab * 12, 0.0f);
    out.etab_si.assign((size_t)ntab * 12, 0.0f);
    out.htab_si.assign((size_t)ntab * 12, 0.0f);
    for (int c = 0; c < ntab; ++c) {
        const float *t = tuples[(size_t)(ntab == 1 ? 0 : c)].data();
        for (int comp = 0; comp < 3; ++comp) {
            float *e = &out.etab[(size_t)c * 12 + comp * 4];
            float *h = &out.htab[(size_t)c * 12 + comp * 4];
            e[0] = t[comp * 3 + 0];
            e[1] = t[comp * 3 + 1];
            e[2] = t[comp * 3 + 2];
            h[0] = t[9 + comp * 3 + 0];
            h[1] = t[9 + comp * 3 + 1];
            h[2] = t[9 + comp * 3 + 2];
            std::memcpy(&out.etab_si[(size_t)c * 12 + comp * 4], e, 4 * sizeof(float));
            std::memcpy(&out.htab_si[(size_t)c * 12 + comp * 4], h, 4 * sizeof(float));
            if (scaled) {
                e[1] = (float)((double)e[1] * sc[comp] / sc[hcomp[comp][0]]);
                e[2] = (float)((double)e[2] * sc[comp] / sc[hcomp[comp][1]]);
                h[1] = (float)((double)h[1] * sc[3 + comp] / sc[ecomp[comp][0]]);
                h[2] = (float)((double)h[2] * sc[3 + comp] / sc[ecomp[comp][1]]);
            }
        }
    }

    // ---- CPML 1-D profiles (setABCPML, fdtdrav.h:1863-2155), FP64 -> FP32
    const CpmlParams &P = in.cpml;
    const int dims[3] = {nx, ny, nz};
    const double dd[3] = {in.dx, in.dy, in.dz};
    for (int a = 0; a < 3; ++a) {
        const int n = dims[a], Ln = P.thick[2 * a], Lp = P.thick[2 * a + 1];
        double smax = P.sigma_max[a];
        if (!(smax > 0.0)) smax = P.sigma_factor * (P.order + 1.0) / (150.0 * M_PI * dd[a]);
        out.sigma_max_used[a] = smax;
        std::vector<float> &pr = out.prof[a];
        pr.assign((size_t)2 * n * 4, 0.0f);
        for (int i = 0; i < n; ++i) {
            pr[(size_t)i * 4 + 2] = 1.0f;
            pr[(size_t)(n + i) * 4 + 2] = 1.0f;
            double re, rm;
            if (i < Ln) {
                re = (Ln - i - 0.75) / Ln;
                rm = (Ln - i - 0.25) / Ln;
            } else if (i >= n - Lp) {
                re = (i - (n - Lp - 1) - 0.75) / Lp;
                rm = (i - (n - Lp - 1) - 0.25) / Lp;
            } else {
                continue;
            }
            const double m = P.order, e0 = in.eps0, u0 = in.mu0;
            double sge = smax * std::pow(re, m);
            double sgm = smax * std::pow(rm, m) * (u0 / e0);
            double kpe = 1.0 + (P.kappa_max - 1.0) * std::pow(re, m);
            double kpm = 1.0 + (P.kappa_max - 1.0) * std::pow(rm, m);
            double ale = P.alpha_min + (P.alpha_max - P.alpha_min) * (1.0 - re);
            double alm = (P.alpha_min + (P.alpha_max - P.alpha_min) * (1.0 - rm)) * (u0 / e0);
            double be = std::exp((-in.dt / e0) * ((sge / kpe) + ale));
            double bm = std::exp((-in.dt / u0) * ((sgm / kpm) + alm));
            double dne = kpe * (sge + kpe * ale), dnm = kpm * (sgm + kpm * alm);
            pr[(size_t)i * 4 + 0] = (float)be;
            pr[(size_t)i * 4 + 1] = (float)(dne != 0.0 ? (1.0 / dd[a]) * (be - 1.0) * sge / dne : 0.0);
            pr[(size_t)i * 4 + 2] = (float)kpe;
            pr[(size_t)(n + i) * 4 + 0] = (float)bm;
            pr[(size_t)(n + i) * 4 + 1] = (float)(dnm != 0.0 ? (1.0 / dd[a]) * (bm - 1.0) * sgm / dnm : 0.0);
            pr[(size_t)(n + i) * 4 + 2] = (float)kpm;
        }
    }

    // ---- sources: one record per driven Ez edge, coefficients from the
    // averaged eps_z / sigma_z of that edge (defineSourceZ,
    // fdtdrav.h:1391-1421).  An x-directed source leaves the Ez coefficients
    // as setCoefficients wrote them and keeps setZero's Cezj = 1.2e-38
    // (fdtdrav.h:1016): calc_e injects on the box's Ez edges whatever the
    // source direction (fdtdrav.h:726-729).
    out.events.clear();
    out.srcs.clear();
    out.srcs_si.clear();
    out.src_j0 = 1;
    out.src_j1 = 0;
    std::vector<Event> ev_e, ev_h;
    std::map<i64, int> owner;
    out.vpn.assign(in.sources.size() * (size_t)in.steps_cap, 0.0f);
    for (size_t q = 0; q < in.sources.size(); ++q) {
        const Source &s = in.sources[q];
        const int *b = s.b;
        const float _is = (float)b[0], _js = (float)b[1], _ks = (float)b[2];
        const float _ie = (float)b[3], _je = (float)b[4], _ke = (float)b[5];
        float r_comp = ((1.0f + _ie) - _is) * ((1.0f + _je) - _js) / (_ke - _ks);
        r_comp = r_comp * (float)s.R;
        const float rfactor = (C.dt * C.dz) / (r_comp * C.dx * C.dy);
        const float raux = r_comp * C.dx * C.dy;
        for (int k = b[2]; k < b[5]; ++k)
            for (int j = b[1]; j <= b[4]; ++j)
                for (int i = b[0]; i <= b[3]; ++i) {
                    const i64 cell = i + nx * (j + (i64)ny * k);
                    if (owner.count(cell)) { err = "lumped sources overlap"; return -1; }
                    owner[cell] = (int)q;
                    const size_t x = (size_t)cell;
                    float e = avg4(mats[lab[x]].v[2], mats[lab[x - 1]].v[2], mats[lab[x - nx]].v[2],
                                   mats[lab[x - 1 - nx]].v[2]);
                    float sg = edge_sigma(2, x, avg4(mats[lab[x]].v[5], mats[lab[x - 1]].v[5], mats[lab[x - nx]].v[5],
                                                     mats[lab[x - 1 - nx]].v[5]));
                    const float two = 2.0f;
                    SrcEdge se{};
                    if (s.dir == 1) {
                        float o[3];
                        ecoef(C, e, sg, C.dx, C.dy, o);
                        se.ceze = o[0];
                        se.cezhy = o[1];
                        se.cezhx = o[2];
                        se.cezj = 1.2e-38f;
                    } else {
                        float D = two * e * C.eps0 + C.dt * sg + rfactor;
                        se.ceze = (two * e * C.eps0 - C.dt * sg - rfactor) / D;
                        se.cezhy = (two * C.dt / C.dx) / D;
                        se.cezhx = (-2.0f * C.dt / C.dy) / D;
                        se.cezj = (-2.0f * C.dt / raux) / D;
                    }
                    out.srcs_si.push_back(se);
                    // the adjoint H gather meets this edge from rows j and j-1
                    out.src_j0 = std::min(out.src_j0 > out.src_j1 ? j : out.src_j0, std::max(0, j - 1));
                    out.src_j1 = std::max(out.src_j1, j);
                    if (scaled) {
                        se.cezhy = (float)((double)se.cezhy * sc[2] / sc[4]);
                        se.cezhx = (float)((double)se.cezhx * sc[2] / sc[3]);
                        se.cezj = (float)((double)se.cezj * sc[2]);
                    }
                    se.wave = (int)q;
                    ev_e.push_back(Event{k, EV_SOURCE, (int)out.srcs.size(), j * nx + i});
                    out.srcs.push_back(se);
                }
        const int vmag = b[5] - b[2];
        if (vmag == 0) continue;   // an x-directed source one plane thick drives no Ez edge
        for (int t = 0; t < in.steps_cap; ++t) {
            double v = waveform(s.wtype, s.amp, s.t0, s.tau, s.freq, (double)t * in.dt);
            out.vpn[q * (size_t)in.steps_cap + t] = (float)v / (float)vmag;
        }
    }
    // ---- probes
    for (size_t p = 0; p < in.probes.size(); ++p) {
        const Probe &pr = in.probes[p];
        Event e{pr.k, EV_PROBE + pr.comp, (int)p, pr.j * nx + pr.i};
        if (pr.comp < 3) ev_e.push_back(e);
        else ev_h.push_back(e);
    }
    auto lessev = [](const Event &a, const Event &b) {
        if (a.col != b.col) return a.col < b.col;
        if (a.k != b.k) return a.k < b.k;
        return a.kind < b.kind;   // a source edge before a probe of the same cell
    };
    std::sort(ev_e.begin(), ev_e.end(), lessev);
    std::sort(ev_h.begin(), ev_h.end(), lessev);
    out.e_first.assign((size_t)nx * ny, -1);
    out.h_first.assign((size_t)nx * ny, -1);
    for (size_t q = 0; q < ev_e.size(); ++q) {
        if (out.e_first[ev_e[q].col] < 0) out.e_first[ev_e[q].col] = (int)out.events.size();
        out.events.push_back(ev_e[q]);
    }
    // a column terminator keeps the device loop from reading past its column
    out.events.push_back(Event{1 << 30, -1, -1, -1});
    for (size_t q = 0; q < ev_h.size(); ++q) {
        if (out.h_first[ev_h[q].col] < 0) out.h_first[ev_h[q].col] = (int)out.events.size();
        out.events.push_back(ev_h[q]);
    }
    out.events.push_back(Event{1 << 30, -1, -1, -1});

    // ---- design edges: derivative of each E edge's update w.r.t. its
    // averaged eps (fdtdrav.h:1135-1177, 1271-1305):  with den = 2 eps eps0 +
    // dt sigma,  dCe/deps = 4 eps0 dt sigma / den^2 and dCb/deps = -Cb 2 eps0 /
    // den, so dE'/deps = dCe E - (2 eps0/den)(E' - Ce E) on an edge that is
    // neither a source edge nor in the CPML.  Stored over the design box
    // dilated by one edge on each high side; zero where no design cell
    // enters the edge's average.
    out.dcoef.clear();
    if (!in.design_eps.empty()) {
        const int *d = in.design;
        const int bx = d[3] - d[0] + 1, by = d[4] - d[1] + 1, bz = d[5] - d[2] + 1;
        const i64 nb = (i64)bx * by * bz;
        out.dcoef.assign((size_t)3 * nb * 4, 0.0f);
        auto in_design = [&](int i, int j, int k) {
            return i >= d[0] && i < d[3] && j >= d[1] && j < d[4] && k >= d[2] && k < d[5];
        };
        for (int kk = 0; kk < bz; ++kk)
            for (int jj = 0; jj < by; ++jj)
                for (int ii = 0; ii < bx; ++ii) {
                    const int i = d[0] + ii, j = d[1] + jj, k = d[2] + kk;
                    if (i < 1 || j < 1 || k < 1 || i >= nx - 1 || j >= ny - 1 || k >= nz - 1) continue;
                    const size_t x = (size_t)(i + nx * (j + (i64)ny * k));
                    const i64 q = ii + (i64)bx * (jj + (i64)by * kk);
                    // the four cells of each component's average
                    const size_t cx[4] = {x, x - nx, x - nxy, x - nx - nxy};
                    const size_t cy[4] = {x, x - 1, x - nxy, x - 1 - nxy};
                    const size_t cz[4] = {x, x - 1, x - nx, x - 1 - nx};
                    const bool dx_[3] = {in_design(i, j, k) || in_design(i, j - 1, k) || in_design(i, j, k - 1) ||
                                             in_design(i, j - 1, k - 1),
                                         in_design(i, j, k) || in_design(i - 1, j, k) || in_design(i, j, k - 1) ||
                                             in_design(i - 1, j, k - 1),
                                         in_design(i, j, k) || in_design(i - 1, j, k) || in_design(i, j - 1, k) ||
                                             in_design(i - 1, j - 1, k)};
                    const size_t *cs[3] = {cx, cy, cz};
                    for (int comp = 0; comp < 3; ++comp) {
                        if (!dx_[comp]) continue;
                        const float e = avg4(mats[lab[cs[comp][0]]].v[comp], mats[lab[cs[comp][1]]].v[comp],
                                             mats[lab[cs[comp][2]]].v[comp], mats[lab[cs[comp][3]]].v[comp]);
                        const float sg = avg4(mats[lab[cs[comp][0]]].v[3 + comp], mats[lab[cs[comp][1]]].v[3 + comp],
                                              mats[lab[cs[comp][2]]].v[3 + comp], mats[lab[cs[comp][3]]].v[3 + comp]);
                        const float ce = tuples[cellcls[x]][comp * 3];
                        const double den = 2.0 * (double)e * in.eps0 + in.dt * (double)sg;
                        float *o = &out.dcoef[((size_t)comp * nb + q) * 4];
                        o[0] = (float)(4.0 * in.eps0 * in.dt * (double)sg / (den * den));
                        o[1] = (float)(2.0 * in.eps0 / den);
                        o[2] = ce;
                    }
                }
    }

    // ---- DFT phasors exp(-j 2 pi f t_m) dt, t_m = m dt
    const int nf = (int)in.freqs.size();
    out.phasor.assign((size_t)in.steps_cap * nf * 2, 0.0);
    for (int t = 0; t < in.steps_cap; ++t)
        for (int f = 0; f < nf; ++f) {
            // one sincos call (glibc): a compiler may or may not merge
            // separate cos and sin calls into it, and the two forms can
            // differ in the last bit
            double sn, cs;
            const double th = 2.0 * M_PI * in.freqs[f] * ((double)t * in.dt);
            ::sincos(th, &sn, &cs);
            out.phasor[((size_t)t * nf + f) * 2 + 0] = cs * in.dt;
            out.phasor[((size_t)t * nf + f) * 2 + 1] = -sn * in.dt;
        }
    return 0;
}

}  // namespace fdtd
