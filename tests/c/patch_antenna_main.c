/* A C main() for the reference's own patch-antenna scene (source.cu:17-325),
 * calling the marchingLoop-compatible shim of INTEGRATION.md with the same
 * argument values, evaluated with C's float/double rules.  Test driver:
 * tests/test_gpu_c_driver.py builds it against the shim and
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
 *    directory as the reference writes them. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "fdtd.h"

int marchingLoop_mi355x(const int NUMDEV, float eps_0, float pi, float mu_0, float c, float dx, float dy, float dz,
    float dt, int X, int Y, int Z, int n_t_steps,
    int pml_x_n, int pml_y_n, int pml_z_n, int pml_x_p, int pml_y_p, int pml_z_p,
    int air_buff_x_n, int air_buff_y_n, int air_buff_z_n, int air_buff_x_p, int air_buff_y_p, int air_buff_z_p,
    int source_tp, float source_freq, float source_amp,
    float source_min_x, float source_min_y, float source_min_z,
    float source_max_x, float source_max_y, float source_max_z,
    int source_direction, float rs, float nc, float tau, float t_0,
    float *brick_min_x, float *brick_min_y, float *brick_min_z,
    float *brick_max_x, float *brick_max_y, float *brick_max_z,
    float *brick_sigma_e_x, float *brick_sigma_e_y, float *brick_sigma_e_z,
    float *brick_eps_r_x, float *brick_eps_r_y, float *brick_eps_r_z,
    float *brick_sigma_m_x, float *brick_sigma_m_y, float *brick_sigma_m_z,
    float *brick_mu_r_x, float *brick_mu_r_y, float *brick_mu_r_z,
    int brick_opt, int brick_num,
    float *pec_min_x, float *pec_min_y, float *pec_min_z, float *pec_max_x, float *pec_max_y, float *pec_max_z,
    float *pec_sigma_e_x, float *pec_sigma_e_y, float *pec_sigma_e_z, int pec_opt, int pec_num,
    float resistor_min_x, float resistor_min_y, float resistor_min_z,
    float resistor_max_x, float resistor_max_y, float resistor_max_z,
    float resistor_resist, int resistor_direction, int resistor_opt,
    float v_min_x, float v_min_y, float v_min_z, float v_max_x, float v_max_y, float v_max_z, int voltage_direction,
    float i_min_x, float i_min_y, float i_min_z, float i_max_x, float i_max_y, float i_max_z, int current_direction,
    float **probes_out, int *n_probes_out);

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
    float p_max_y[3] = {100 * dy, 50 * dy, (50 * dy) + (16e-3)};
    float p_max_z[3] = {0, 3 * dz, 3 * dz};
    float p_sig[3] = {1e10, 1e10, 1e10};

    /* Gaussian source, 50 ohm, z-directed (source.cu:154-178) */
    float nc = 20;
    float tau = (nc * dy) / (2 * c);
    float t_0 = 3 * tau;

    /* bounding box of the objects: the maxima (see the header comment) */
    int X = (int)round(b_max[0] / dx), Y = (int)round(b_max[1] / dy), Z = (int)round(b_max[2] / dz);

    float *vi = NULL;
    int nvi = 0;
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
        &vi, &nvi);
    if (rc != FDTD_OK) {
        fprintf(stderr, "marchingLoop_mi355x: %d\n", rc);
        return 1;
    }
    FILE *f = fopen(argv[2], "wb");
    if (!f || fwrite(vi, sizeof(float), (size_t)nvi * n, f) != (size_t)nvi * n) {
        fprintf(stderr, "cannot write %s\n", argv[2]);
        return 1;
    }
    fclose(f);
    free(vi);
    printf("grid %d %d %d dt %.9e steps %d\n", X + 50, Y + 50, Z + 50, dt, n);
    return 0;
}
