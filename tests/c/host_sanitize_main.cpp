/* Host code of the engine under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_host.py builds it host-only and runs it; no GPU is touched):
 * the scene builder (class keys, coefficient tables, CPML profiles, events,
 * waveform and phasor tables, design-edge derivatives) for FP32, FP16 and
 * block-scaled storage on a scene with every kind of object, then the port
 * reductions and the results-file writer on synthetic samples.  Exits 0 when
 * every call succeeds; a sanitizer report fails the run. */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "internal.h"

using namespace fdtd;

static SceneIn scene(int storage)
{
    SceneIn in{};
    in.nx = 26; in.ny = 22; in.nz = 18;
    in.px = 64;
    in.dx = 1e-3; in.dy = 1.2e-3; in.dz = 0.9e-3;
    in.eps0 = 8.854187817e-12; in.mu0 = 4e-7 * M_PI;
    const double c = 1.0 / std::sqrt(in.eps0 * in.mu0);
    in.dt = 0.99 / (c * std::sqrt(1 / (in.dx * in.dx) + 1 / (in.dy * in.dy) + 1 / (in.dz * in.dz)));
    in.storage = storage;
    in.cpml = CpmlParams{{3, 4, 3, 3, 2, 3}, 3.0, {0, 0, 0}, 1.5, 5.0, 0.05, 0.0};
    in.boxes.push_back(Box{{5, 5, 4, 14, 12, 10}, {4, 4, 4, 0.01f, 0.01f, 0.01f, 1, 1, 1, 0, 0, 0}});
    in.boxes.push_back(Box{{10, 8, 6, 20, 16, 13}, {2, 3, 5, 0, 0.02f, 0, 2, 1.5f, 3, 10, 0, 5}});
    in.sources.push_back(Source{{8, 9, 7, 9, 10, 9}, 50.0, 2, 1.0, 40 * in.dt, 12 * in.dt, 0.0, 3});
    in.sources.push_back(Source{{16, 6, 12, 18, 7, 13}, 75.0, 3, 0.5, 0.0, 0.0, 3e9, 1});
    in.plates.push_back(Plate{3, {6, 6, 5, 12, 9, 5}, {1e10f, 1e10f, 1e10f}});
    in.plates.push_back(Plate{1, {13, 5, 5, 13, 9, 8}, {5.0f, 5.0f, 5.0f}});
    in.resistors.push_back(Resistor{{18, 14, 6, 18, 14, 8}, 100.0});
    in.sectors.push_back(Sector{{8, 12, 12, 12, 12, 16}, 11, 20, 18, 4.5e-3f, {1e8f, 1e8f}});
    for (int comp = 0; comp < 6; ++comp) in.probes.push_back(Probe{comp, 11 + comp, 10, 9});
    in.freqs = {1e9, 5e9, 12e9};
    const int d[6] = {17, 2, 4, 20, 5, 7};
    for (int q = 0; q < 6; ++q) in.design[q] = d[q];
    for (int q = 0; q < 3 * 3 * 3; ++q) in.design_eps.push_back(1.0f + 0.05f * (float)q);
    in.steps_cap = 40;
    return in;
}

int main()
{
    for (int storage : {0, 1, 2, 3}) {
        SceneIn in = scene(storage);
        Built b;
        std::string err;
        if (build_scene(in, b, err) != 0) {
            std::fprintf(stderr, "build_scene (storage %d): %s\n", storage, err.c_str());
            return 1;
        }
        if (b.etab.empty() || b.htab.empty() || b.etab.size() % 12 || b.htab.size() % 12 || b.vpn.empty() ||
            b.phasor.size() != (size_t)in.steps_cap * in.freqs.size() * 2 || b.dcoef.empty()) {
            std::fprintf(stderr, "build_scene (storage %d): unexpected table sizes\n", storage);
            return 1;
        }
        for (float v : b.etab)
            if (!std::isfinite(v)) {
                std::fprintf(stderr, "non-finite E coefficient\n");
                return 1;
            }
        std::printf("storage %d: %d classes, %zu events\n", storage, b.ncls, b.events.size());
    }
    // an invalid scene is refused with a message, not a crash
    {
        SceneIn in = scene(0);
        in.resistors.push_back(Resistor{{1, 1, 1, 1, 1, 3}, 50.0});   // inside the CPML slab
        Built b;
        std::string err;
        if (build_scene(in, b, err) == 0 || err.empty()) {
            std::fprintf(stderr, "a resistor in the CPML slab was accepted\n");
            return 1;
        }
    }
    // port voltage / current reductions on synthetic samples
    const PortDef pd{3, {8, 9, 4, 10, 10, 7}, 2, {8, 9, 7, 10, 9, 7}};
    PortPlan p;
    std::string err;
    if (port_plan(pd, 26, 22, 18, 1e-3f, 1.2e-3f, 0.9e-3f, p, err) != 0) {
        std::fprintf(stderr, "port_plan: %s\n", err.c_str());
        return 1;
    }
    const size_t ns = 64;
    std::vector<float> hist(p.samples.size() * ns);
    for (size_t q = 0; q < hist.size(); ++q) hist[q] = std::sin(0.37f * (float)q);
    std::vector<float> v(ns), cur(ns);
    port_series(p, hist.data(), ns, v.data(), cur.data());
    const double f[3] = {1e9, 5e9, 12e9};
    std::vector<double> out(3 * 8);
    port_spectrum(v.data(), cur.data(), ns, 1e-12, 50.0, 3, f, out.data());
    const char *path = std::getenv("SANITIZE_OUT");
    if (path && port_save(v.data(), ns, path) != 0) {
        std::fprintf(stderr, "port_save failed\n");
        return 1;
    }
    std::printf("port: %zu samples, V[ns-1] %g\n", p.samples.size(), (double)v[ns - 1]);
    return 0;
}
