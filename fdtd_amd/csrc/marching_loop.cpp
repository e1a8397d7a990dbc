// marching_loop.cpp -- marchingLoop's parameter list on top of the fdtd_* ABI
// (fdtdrav.h:3425-4387 for one device).  Physical coordinates become cell
// indices as marchingLoop computes them: air buffer + CPML + calc_index
// (fdtdrav.h:198-206, 3529-3537, 3596-3601); the logical grid is
// NXX x NYY x NZ (fdtdrav.h:3487-3491) without the NUMDEV split or the
// 16-cell padding, which the library does not need.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/fdtd_marching_loop.h"

namespace {

int cidx(float coord, float d) { return (int)std::roundf(coord / d); }   // calc_index, fdtdrav.h:198-206

// On failure: report the library's message, release the context, return the code.
int failed(fdtd_ctx *ctx, const char *what, int rc)
{
    std::fprintf(stderr, "marchingLoop_mi355x: %s: %s\n", what, ctx ? fdtd_last_error(ctx) : "");
    if (ctx) fdtd_destroy(ctx);
    return rc;
}

}  // namespace

#define ML_CHECK(call)                                  \
    do {                                                \
        const int rc_ = (call);                         \
        if (rc_ != FDTD_OK) return failed(ctx, #call, rc_); \
    } while (0)

extern "C" int marchingLoop_mi355x(const int NUMDEV, float eps_0, float pi, float mu_0, float c, float dx, float dy,
    float dz, float dt, int X, int Y, int Z, int n_t_steps,
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
    float *v_out, float *i_out)
{
    (void)NUMDEV;
    (void)c;
    (void)nc;
    (void)resistor_direction;
    fdtd_ctx *ctx = nullptr;
    if (source_direction != 1 && source_direction != 3) return FDTD_E_UNSUPPORTED;
    if (source_tp != 1 && source_tp != 3) return FDTD_E_UNSUPPORTED;   // case 2 fills nothing, fdtdrav.h:4077-4079
    const int ox = air_buff_x_n + pml_x_n, oy = air_buff_y_n + pml_y_n, oz = air_buff_z_n + pml_z_n;
    fdtd_grid g;
    std::memset(&g, 0, sizeof g);
    g.nx = X + air_buff_x_n + air_buff_x_p + pml_x_n + pml_x_p;   // NXX, fdtdrav.h:3487
    g.ny = Y + air_buff_y_n + air_buff_y_p + pml_y_n + pml_y_p;   // NYY
    g.nz = Z + air_buff_z_n + air_buff_z_p + pml_z_n + pml_z_p;   // NZ before the NUMDEV padding
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.dt = dt;
    g.eps0 = eps_0;
    g.mu0 = mu_0;
    int rc = fdtd_create(&g, FDTD_STORE_F32, &ctx);
    if (rc != FDTD_OK) return failed(nullptr, "fdtd_create", rc);

    // setABCPML's constants (fdtdrav.h:1774-1812): m = 3, kappa 5, alpha 0.05 -> 0, and
    // sigma_max = 1.5 sigma_opt with sigma_opt = (m + 1) / (150 pi d), float
    // arithmetic with the caller's pi (fdtdrav.h:1793-1808)
    const float npml = 3, sigma_factor = 1.5f;
    const float smax[3] = {sigma_factor * ((npml + 1) / (150 * pi * 1 * dx)),
                           sigma_factor * ((npml + 1) / (150 * pi * 1 * dy)),
                           sigma_factor * ((npml + 1) / (150 * pi * 1 * dz))};
    const fdtd_cpml cp = {{pml_x_n, pml_x_p, pml_y_n, pml_y_p, pml_z_n, pml_z_p}, 3.0,
                          {smax[0], smax[1], smax[2]}, 1.5, 5.0, 0.05, 0.0};
    ML_CHECK(fdtd_set_cpml(ctx, &cp));

    for (int b = 0; brick_opt && b < brick_num; ++b) {   // setBrick loop, fdtdrav.h:4104-4128
        const float eps[3] = {brick_eps_r_x[b], brick_eps_r_y[b], brick_eps_r_z[b]};
        const float se[3] = {brick_sigma_e_x[b], brick_sigma_e_y[b], brick_sigma_e_z[b]};
        const float mu[3] = {brick_mu_r_x[b], brick_mu_r_y[b], brick_mu_r_z[b]};
        const float sm[3] = {brick_sigma_m_x[b], brick_sigma_m_y[b], brick_sigma_m_z[b]};
        ML_CHECK(fdtd_add_box(ctx, ox + cidx(brick_min_x[b], dx), oy + cidx(brick_min_y[b], dy),
                              oz + cidx(brick_min_z[b], dz), ox + cidx(brick_max_x[b], dx),
                              oy + cidx(brick_max_y[b], dy), oz + cidx(brick_max_z[b], dz), eps, se, mu, sm));
    }
    for (int p = 0; pec_opt && p < pec_num; ++p) {   // definePlateZ on every plate, fdtdrav.h:4183-4210
        const float sg[3] = {pec_sigma_e_x[p], pec_sigma_e_y[p], pec_sigma_e_z[p]};
        ML_CHECK(fdtd_add_plate(ctx, 3, ox + cidx(pec_min_x[p], dx), oy + cidx(pec_min_y[p], dy),
                                oz + cidx(pec_min_z[p], dz), ox + cidx(pec_max_x[p], dx),
                                oy + cidx(pec_max_y[p], dy), oz + cidx(pec_max_z[p], dz), sg));
    }
    if (resistor_opt)   // defineResistor, fdtdrav.h:4296-4309
        ML_CHECK(fdtd_add_resistor(ctx, 3, ox + cidx(resistor_min_x, dx), oy + cidx(resistor_min_y, dy),
                                   oz + cidx(resistor_min_z, dz), ox + cidx(resistor_max_x, dx),
                                   oy + cidx(resistor_max_y, dy), oz + cidx(resistor_max_z, dz), resistor_resist));

    // setGauss / setSine; the library's sine is sin(2 pi_lib f t), so the
    // frequency carries the caller's pi: sin(2 pi f dt m), fdtdrav.h:1515
    const double f_eff = (double)source_freq * ((double)pi / M_PI);
    const fdtd_waveform w = {source_tp, source_amp, t_0, tau, f_eff};
    ML_CHECK(fdtd_add_lumped_source(ctx, source_direction, ox + cidx(source_min_x, dx),
                                    oy + cidx(source_min_y, dy), oz + cidx(source_min_z, dz),
                                    ox + cidx(source_max_x, dx), oy + cidx(source_max_y, dy),
                                    oz + cidx(source_max_z, dz), rs, &w));

    // sampled-voltage and sampled-current boxes, fdtdrav.h:3614-3650
    ML_CHECK(fdtd_set_port(ctx, voltage_direction, ox + cidx(v_min_x, dx), oy + cidx(v_min_y, dy),
                           oz + cidx(v_min_z, dz), ox + cidx(v_max_x, dx), oy + cidx(v_max_y, dy),
                           oz + cidx(v_max_z, dz), current_direction, ox + cidx(i_min_x, dx),
                           oy + cidx(i_min_y, dy), oz + cidx(i_min_z, dz), ox + cidx(i_max_x, dx),
                           oy + cidx(i_max_y, dy), oz + cidx(i_max_z, dz)));

    ML_CHECK(fdtd_finalize(ctx, n_t_steps));
    ML_CHECK(fdtd_run(ctx, n_t_steps));                                     // fdtdrav.h:4320-4355
    ML_CHECK(fdtd_save_port(ctx, "results1003.dat", "results1004.dat"));    // fdtdrav.h:4357-4367
    if (v_out || i_out) {
        // the port series of every step; the one the caller did not ask for
        // goes to a temporary buffer
        float *tmp = (v_out && i_out) ? nullptr : new float[(size_t)n_t_steps];
        rc = fdtd_get_port(ctx, v_out ? v_out : tmp, i_out ? i_out : tmp, (size_t)n_t_steps);
        delete[] tmp;
        if (rc != FDTD_OK) return failed(ctx, "fdtd_get_port", rc);
    }
    return fdtd_destroy(ctx);
}
