/* fdtd_marching_loop.h -- the reference's own entry point on the C ABI.
 *
 * marchingLoop_mi355x replaces marchingLoop (fdtdrav.h:3425-3453), called
 * from source.cu:297-325, for one MI355X: the same parameter list and
 * meaning, mapped onto the fdtd_* calls of fdtd.h in marchingLoop's set-up
 * order (fdtdrav.h:4104-4309), the time loop (fdtdrav.h:4320-4355) and
 * results1003.dat / results1004.dat in the working directory
 * (fdtdrav.h:4357-4367).
 *
 * Differences a caller sees:
 *  - it returns FDTD_OK or a negative FDTD_E_* code (the reference calls
 *    exit() on errors and returns nothing);
 *  - NUMDEV is ignored (one device); c is implied by eps_0 and mu_0, which
 *    are used as given; pi enters where the reference uses it: the CPML's
 *    sigma_max = 1.5 (npml + 1) / (150 pi d) per axis, evaluated in float
 *    (fdtdrav.h:1793-1808), and the sine source sin(2 pi f dt m)
 *    (fdtdrav.h:1505-1520);
 *  - source_direction 1 (defineSourceX) and 3 (defineSourceZ) are supported,
 *    other values return FDTD_E_UNSUPPORTED; voltage_direction must be 3 (the
 *    reference samples only Ez, fdtdrav.h:740);
 *  - source_tp 1 (setGauss) and 3 (setSine) are supported; source_tp 2
 *    returns FDTD_E_UNSUPPORTED: the reference's switch leaves its signal
 *    buffer unwritten for it (fdtdrav.h:4070-4085), so there is no waveform
 *    to reproduce (the Gaussian derivative of fdtd_add_lumped_source's
 *    waveform type 2 is this library's own);
 *  - the resistor is added only when resistor_opt != 0: the reference runs
 *    defineResistor whenever resistor_direction != 0 but computes its indices
 *    only when resistor_opt != 0 (fdtdrav.h:3575-3587, 4296);
 *  - v_out and i_out, when not NULL, receive the port voltage and the
 *    averaged port current of every step (n_t_steps floats each, caller
 *    allocated), the series written to the results files;
 *  - all device memory is freed before it returns.
 */
#ifndef FDTD_MARCHING_LOOP_H
#define FDTD_MARCHING_LOOP_H

#include "fdtd.h"

#ifdef __cplusplus
extern "C" {
#endif

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
    float *v_out, float *i_out);

#ifdef __cplusplus
}
#endif

#endif /* FDTD_MARCHING_LOOP_H */
