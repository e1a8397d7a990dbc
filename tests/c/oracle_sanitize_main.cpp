/* The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_host.py builds this file, which includes the oracle's source,
 * with g++ and runs it): FP32 and FP64 runs of a scene with every kind of
 * object, the storage-rounding models, and the FP64 adjoint. */
#include "../../oracle/fdtd_oracle.cpp"

#include <cstdio>

int main()
{
    const int box[12] = {5, 5, 4, 14, 12, 10, 10, 8, 6, 20, 16, 13};
    const float mat[24] = {4, 4, 4, 0.01f, 0.01f, 0.01f, 1, 1, 1, 0, 0, 0,
                           2, 3, 5, 0, 0.02f, 0, 2, 1.5f, 3, 10, 0, 5};
    const int src[12] = {8, 9, 7, 9, 10, 9, 16, 6, 12, 18, 7, 13};
    const double R[2] = {50, 75}, wpar[8] = {1, 40e-13, 12e-13, 0, 0.5, 0, 0, 3e9};
    const int wtype[2] = {2, 3}, sdir[2] = {3, 1};
    const int probe[24] = {0, 11, 10, 9, 1, 12, 10, 9, 2, 13, 10, 9, 3, 14, 10, 9, 4, 15, 10, 9, 5, 16, 10, 9};
    const double freqs[3] = {1e9, 5e9, 12e9};
    const int plate[14] = {3, 6, 6, 5, 12, 9, 5, 1, 13, 5, 5, 13, 9, 8};
    const float psig[6] = {1e10f, 1e10f, 1e10f, 5, 5, 5};
    const int res[6] = {18, 14, 6, 18, 14, 8};
    const double resR[1] = {100};
    const int sector[9] = {8, 12, 12, 12, 12, 16, 11, 20, 18};
    const float sectf[3] = {4.5e-3f, 1e8f, 1e8f};
    float deps[3 * 3 * 3];
    for (int q = 0; q < 27; ++q) deps[q] = 1.0f + 0.05f * (float)q;
    oracle_scene sc{};
    sc.nx = 26; sc.ny = 22; sc.nz = 18;
    sc.dx = 1e-3; sc.dy = 1.2e-3; sc.dz = 0.9e-3;
    sc.eps0 = 8.854187817e-12; sc.mu0 = 4e-7 * M_PI;
    sc.dt = 1.5e-12;
    const int pml[6] = {3, 4, 3, 3, 2, 3};
    for (int q = 0; q < 6; ++q) sc.pml[q] = pml[q];
    sc.order = 3; sc.sigma_max[0] = sc.sigma_max[1] = sc.sigma_max[2] = 10;
    sc.kappa_max = 5; sc.alpha_max = 0.05; sc.alpha_min = 0;
    sc.nbox = 2; sc.box = box; sc.box_mat = mat;
    sc.nsrc = 2; sc.src = src; sc.src_R = R; sc.src_wtype = wtype; sc.src_wpar = wpar; sc.src_dir = sdir;
    sc.nprobe = 6; sc.probe = probe;
    sc.nfreq = 3; sc.freqs = freqs;
    const int d[6] = {17, 2, 4, 20, 5, 7};
    for (int q = 0; q < 6; ++q) sc.design[q] = d[q];
    sc.design_eps = deps;
    sc.nsteps = 30;
    sc.nplate = 2; sc.plate = plate; sc.plate_sigma = psig;
    sc.nres = 1; sc.res = res; sc.res_R = resR;
    sc.nsector = 1; sc.sector = sector; sc.sector_f = sectf;
    std::vector<double> pr(6 * 30), dft(6 * 3 * 2), grad(27);
    for (int prec : {32, 64}) {
        void *h = oracle_create(&sc, prec);
        if (!h || oracle_steps(h, 30) != 0 || oracle_get_probes(h, pr.data()) != 0 ||
            oracle_get_dft(h, dft.data()) != 0) {
            std::fprintf(stderr, "oracle run failed (FP%d)\n", prec);
            return 1;
        }
        oracle_destroy(h);
    }
    for (int mode : {16, 17, 48}) {
        void *h = oracle_create(&sc, 32);
        if (!h || oracle_set_quant(h, mode, 1.0, 376.73, -1) != 0 || oracle_steps(h, 30) != 0) {
            std::fprintf(stderr, "storage model %d failed\n", mode);
            return 1;
        }
        oracle_destroy(h);
    }
    void *h = oracle_create(&sc, 64);
    const int tp[2] = {2, 2}, tf[2] = {0, 2};
    const double tw[2] = {1.0, 0.5};
    double J = 0;
    if (!h || oracle_adjoint(h, 2, tp, tf, tw, &J, grad.data()) != 0 || !(J > 0)) {
        std::fprintf(stderr, "oracle adjoint failed\n");
        return 1;
    }
    oracle_destroy(h);
    std::printf("oracle ok: J %g, grad[0] %g\n", J, grad[0]);
    return 0;
}
