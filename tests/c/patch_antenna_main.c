/* A C main() for the reference's own patch-antenna scene (source.cu:17-325),
 * calling marchingLoop_mi355x (include/fdtd_marching_loop.h, the library's
 * counterpart of marchingLoop) with the same argument values, evaluated
 * with C's float/double rules.  tests/test_gpu_c_driver.py builds it against
 * libfdtd_amd.so and compares its output with the Python-API run.
 *
 * Differences from source.cu, each on purpose:
 *  - n_t_steps and NUMDEV come from argv[1] and a fixed 1 (one device);
 *  - X, Y, Z are the maxima of the objects' extents (60 dx, 100 dy, 3 dz):
 *    source.cu:256-273 compares every maximum with `<` and never updates y,
 *    which would size the grid from the smallest object;
 *  - literals are double, converted to float where they are stored or passed,
 *    as in source.cu;
 *  - V[m] and I[m] are also written raw (float32) to argv[2] for a bitwise
 *    comparison; results1003.dat / results1004.dat go to the working
 *    directory as the reference writes them.
 * The file is valid C and C++.  Built as C, sqrt of a float promotes to
 * double; built as C++ (as nvcc builds source.cu) it is the float overload,
 * so c and dt differ in the last float bits between the two builds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "fdtd_marching_loop.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: %s n_t_steps out.raw\n", argv[0]);
        return 2;
    }
    const int n = atoi(argv[1]);

    /* constants and cell sizes (source.cu:29-46) */
    float eps_0 = 8.854187817e-12;
    float pi = 3.1415;
    float mu_0 = (4 * pi) * 1e-7;
    float c = 1 / sqrt(mu_0 * eps_0);
    float dx = 0.389e-3, dy = 0.4e-3, dz = 0.265e-3;
    float dt = 1 / (c * sqrt(1 / pow(dx, 2) + 1 / pow(dy, 2) + 1 / pow(dz, 2)));   /* Courant factor 1 */

    /* substrate (source.cu:72-103) */
    float b_min[3] = {0, 0, 0}, b_max[3] = {60 * dx, 100 * dy, 3 * dz};
    float b_se[3] = {0.0004, 0.0004, 0.0004}, b_eps[3] = {2.2, 2.2, 2.2};
    float b_sm[3] = {1.2e-38, 1.2e-38, 1.2e-38}, b_mu[3] = {1, 1, 1};

    /* ground plane, feed line, patch (source.cu:105-151) */
    float p_min_x[3] = {0, 7.535e-3, 5.445e-3}, p_min_y[3] = {0, 0, 50 * dy}, p_min_z[3] = {0, 3 * dz, 3 * dz};
    float p_max_x[3] = {60 * dx, 9.869e-3, 17.895e-3};
    float p_max_y[3] = {100 * dy, 50 * dy, (float)((50 * dy) + (16e-3))};
    float p_max_z[3] = {0, 3 * dz, 3 * dz};
    float p_sig[3] = {1e10, 1e10, 1e10};

    /* Gaussian source, 50 ohm, z-directed (source.cu:154-178) */
    float nc = 20;
    float tau = (nc * dy) / (2 * c);
    float t_0 = 3 * tau;

    /* bounding box of the objects: the maxima (see the header comment) */
    int X = (int)round(b_max[0] / dx), Y = (int)round(b_max[1] / dy), Z = (int)round(b_max[2] / dz);

    float *vi = (float *)malloc(sizeof(float) * 2 * (size_t)(n > 0 ? n : 1));
    if (!vi) return 1;
    int rc = marchingLoop_mi355x(1, eps_0, pi, mu_0, c, dx, dy, dz, dt, X, Y, Z, n,
        10, 10, 10, 10, 10, 10, 15, 15, 15, 15, 15, 15,
        1, 1e8, 1, 7.535e-3, 0, 0, 9.869e-3, 0, 3 * dz, 3, 50, nc, tau, t_0,
        &b_min[0], &b_min[1], &b_min[2], &b_max[0], &b_max[1], &b_max[2],
        &b_se[0], &b_se[1], &b_se[2], &b_eps[0], &b_eps[1], &b_eps[2],
        &b_sm[0], &b_sm[1], &b_sm[2], &b_mu[0], &b_mu[1], &b_mu[2], 1, 1,
        p_min_x, p_min_y, p_min_z, p_max_x, p_max_y, p_max_z, p_sig, p_sig, p_sig, 1, 3,
        7e-3, 0, 0, 8e-3, 2e-3, 4e-3, 50, 1, 0,                      /* resistor_opt = 0 (source.cu:190) */
        7.535e-3, 10 * dy, 0, 9.869e-3, 10 * dy, 3 * dz, 3,
        7.535e-3, 10 * dy, 3 * dz, 9.869e-3, 10 * dy, 3 * dz, 2,
        vi, vi + n);
    if (rc != FDTD_OK) {
        fprintf(stderr, "marchingLoop_mi355x: %d\n", rc);
        return 1;
    }
    FILE *f = fopen(argv[2], "wb");
    if (!f || fwrite(vi, sizeof(float), 2 * (size_t)n, f) != 2 * (size_t)n) {
        fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    fclose(f);
    free(vi);
    printf("grid %d %d %d dt %.9e steps %d\n", X + 50, Y + 50, Z + 50, dt, n);
    return 0;
}
