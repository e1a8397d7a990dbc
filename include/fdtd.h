/*
 * fdtd.h -- C ABI of the MI355X (gfx950) 3-D Yee/CPML FDTD engine.
 *
 * This is the drop-in replacement for the host orchestration boundary of
 * ravelec/FDTD, `marchingLoop(...)` (fdtdrav.h:3425-3453), whose ~110
 * positional arguments are split here into a handle plus one call per
 * concept: grid, CPML, material boxes, lumped sources, probes, the time-step
 * loop and the outputs.  Every entry point below names the reference code it
 * replaces.
 *
 * Conventions
 *   - Every call returns int: FDTD_OK (0) on success, a negative FDTD_E_* code
 *     otherwise.  Nothing calls exit(); the reference's HANDLE_ERROR did
 *     (fdtdrav.h:24-33).  fdtd_last_error() gives the text of the last error.
 *   - Indices are cell indices (i, j, k) on the logical grid
 *     [0,nx) x [0,ny) x [0,nz), which includes the CPML cells -- the
 *     reference's NXX x NYY x NZ (fdtdrav.h:3487-3492).  Flat arrays handed in
 *     or out are x-fastest: idx = i + nx*(j + ny*k).
 *   - Ownership: the library owns all device memory.  Input arrays are copied
 *     during the call; output arrays are caller-allocated.
 *   - Threading: a context is not thread-safe.  Use one host thread per
 *     context.  A context runs on the HIP device that was current when
 *     fdtd_create() was called, on one internal stream (a checkpointed
 *     fdtd_adjoint_gradient adds a second one for the call).
 *   - Life cycle: create -> set_cpml / add_box / add_lumped_source /
 *     add_probe / set_dft / set_design_region -> finalize -> run ... ->
 *     get_* -> destroy.  fdtd_reset() rewinds to step 0 with zero fields.
 */
#ifndef FDTD_H
#define FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---------------------------------------------------- */
#define FDTD_OK             0
#define FDTD_E_INVALID     -1  /* invalid argument or configuration        */
#define FDTD_E_STATE       -2  /* call not allowed in the current state    */
#define FDTD_E_RANGE       -3  /* index or box outside the grid            */
#define FDTD_E_NOMEM       -4  /* host or device allocation failed         */
#define FDTD_E_DEVICE      -5  /* HIP runtime error                        */
#define FDTD_E_UNSUPPORTED -6  /* feature not provided by the library       */
#define FDTD_E_NONFINITE   -7  /* a probe recorded a non-finite value      */
#define FDTD_E_CHECKSUM    -8  /* checkpoint blob is corrupt or mismatched */

/* Opaque context: one simulation on one device. */
typedef struct fdtd_ctx fdtd_ctx;

/* ---- enumerations --------------------------------------------------- */
/* Field storage precision.  Arithmetic is FP32 in every form. */
#define FDTD_STORE_F32   0
#define FDTD_STORE_F16   1  /* FP16 storage of E, H and psi, scaled       */
#define FDTD_STORE_BF16  2  /* BF16 storage of E, H and psi, scaled       */
#define FDTD_STORE_BQ16  3  /* E and H as int16 mantissas with one power-of-two
                               exponent byte per 8 x-cells (2.125 B/value),
                               psi FP16, scaled; no adjoint               */

/* Field components (probe and field-access selectors). */
#define FDTD_EX 0
#define FDTD_EY 1
#define FDTD_EZ 2
#define FDTD_HX 3
#define FDTD_HY 4
#define FDTD_HZ 5

/* Waveform types; the numbering follows the reference's source_tp
 * (source.cu:160, fdtdrav.h:4071-4085). */
#define FDTD_WAVE_GAUSSIAN       1  /* amp*exp(-((t-t0)/tau)^2)                     */
#define FDTD_WAVE_GAUSS_DERIV    2  /* amp*sqrt(2e)*((t-t0)/tau)*exp(-((t-t0)/tau)^2) */
#define FDTD_WAVE_SINE           3  /* amp*sin(2*pi*freq*t)                          */

/* ---- parameter blocks ----------------------------------------------- */

/* Grid and constants.  Replaces the constants and dt of source.cu:30-46 and
 * the sizing of fdtdrav.h:3487-3498 (NUMDEV = 1, so no z padding). */
typedef struct {
    int    nx, ny, nz;      /* cells per axis, CPML included                 */
    double dx, dy, dz;      /* cell sizes [m]                                */
    double dt;              /* time step [s]; <= 0 selects courant * 3-D CFL;
                               may exceed the CFL limit by float resolution
                               (4 FLT_EPSILON), no more                       */
    double courant;         /* Courant factor used when dt <= 0 (source.cu:41-46) */
    double eps0, mu0;       /* <= 0 selects 8.854187817e-12 and 4*pi*1e-7     */
} fdtd_grid;

/* CPML parameters.  The reference hard-codes them in setABCPML
 * (fdtdrav.h:1774-1812); widths come from pml_{x,y,z}_{n,p}
 * (source.cu:53-59).  A zero width disables that face (PEC backing only). */
typedef struct {
    int    thick[6];        /* x-, x+, y-, y+, z-, z+ widths in cells         */
    double order;           /* polynomial grading m (reference: 3)            */
    double sigma_max[3];    /* per axis; <= 0 selects sigma_factor*sigma_opt,
                               sigma_opt = (m+1)/(150*pi*d) (fdtdrav.h:1783-1798) */
    double sigma_factor;    /* reference: 1.5                                 */
    double kappa_max;       /* reference: 5                                   */
    double alpha_max;       /* reference: 0.05                                */
    double alpha_min;       /* reference: 0                                   */
} fdtd_cpml;

/* Source waveform.  Replaces setGauss / setSine (fdtdrav.h:1486-1520) and the
 * source_tp / source_amp / tau / t_0 / source_freq arguments (fdtdrav.h:3430-3433). */
typedef struct {
    int    type;            /* FDTD_WAVE_*                                    */
    double amp;             /* [V]                                            */
    double t0;              /* [s]                                            */
    double tau;             /* [s]                                            */
    double freq;            /* [Hz], FDTD_WAVE_SINE only                      */
} fdtd_waveform;

/* Adjoint objective: J = sum_t weight[t] * |DFT(probe[t], freq[t])|^2. */
typedef struct {
    int           n_terms;
    const int    *probe;    /* probe index (an E-component probe)            */
    const int    *freq;     /* DFT frequency index                            */
    const double *weight;   /* NULL selects 1 for every term                  */
} fdtd_objective;

/* Run statistics. */
typedef struct {
    int      nx, ny, nz;
    int      steps_done;
    int      steps_capacity;
    int      n_probes, n_freqs, n_sources;
    int      n_classes;          /* distinct E/H coefficient classes           */
    int      storage;            /* FDTD_STORE_*                               */
    double   dt;
    double   last_run_ms;        /* hipEvent time of the last fdtd_run() loop  */
    double   e_kernel_ms;        /* mean E-kernel launch time (fdtd_profile)   */
    double   h_kernel_ms;        /* mean H-kernel launch time (fdtd_profile)   */
    double   step_kernel_ms;     /* mean time of one step's launches            */
    int      interior_split;     /* 1: the sweeps run interior / x-edge /
                                    general region variants; 0: general only */
    int      pad1;
    double   adjoint_forward_ms; /* last fdtd_adjoint_gradient: forward pass   */
    double   adjoint_total_ms;   /* ... forward + recompute + reverse sweep     */
    int      adjoint_segments;   /* ... checkpoint segments used                */
    int      pad0;
    uint64_t device_bytes;       /* device memory held by the context          */
    uint64_t bytes_per_cell_e;   /* algorithmic bytes per cell, E sweep        */
    uint64_t bytes_per_cell_h;   /* algorithmic bytes per cell, H sweep        */
    uint64_t adjoint_device_bytes; /* last fdtd_adjoint_gradient: peak device
                                    memory, context plus adjoint buffers      */
} fdtd_info;

/* ---- life cycle ----------------------------------------------------- */

/* Create a context on the current HIP device.  Replaces marchingLoop's
 * sizing, P2P set-up and allocation (fdtdrav.h:3455-3498, 3874-4045) for one
 * device. storage is FDTD_STORE_*. */
int fdtd_create(const fdtd_grid *grid, int storage, fdtd_ctx **out);

/* Release every resource.  The reference frees nothing (fdtdrav.h:4384). */
int fdtd_destroy(fdtd_ctx *ctx);

/* Text of the last error on this context; with ctx NULL, of the last
 * fdtd_create on the calling thread that failed (thread-local storage). */
const char *fdtd_last_error(const fdtd_ctx *ctx);

/* ---- scene ---------------------------------------------------------- */

/* CPML parameters; replaces setABCPML's constants (fdtdrav.h:1747-1812). */
int fdtd_set_cpml(fdtd_ctx *ctx, const fdtd_cpml *cpml);

/* Paint a material box on cells is<=i<ie, js<=j<je, ks<=k<ke; later boxes
 * win.  Replaces setBrick (fdtdrav.h:1524-1571) and its launch loop
 * (fdtdrav.h:4104-4128).  Each array holds the x, y, z values. */
int fdtd_add_box(fdtd_ctx *ctx, int is, int js, int ks, int ie, int je, int ke,
                 const float eps_r[3], const float sigma_e[3],
                 const float mu_r[3], const float sigma_m[3]);

/* Resistive lumped voltage source.  dir 3: the Ez edges is<=i<=ie,
 * js<=j<=je, ks<=k<ke; replaces defineSourceZ (fdtdrav.h:1368-1426) with the
 * injection of calc_e (fdtdrav.h:726-729).  Any number of non-overlapping
 * z-sources is allowed; the reference had one.
 * dir 1 reproduces defineSourceX (fdtdrav.h:1311-1366) as marchingLoop runs
 * it: the Ex edges is<=i<ie, js<=j<=ie (the reference bounds j by ie),
 * ks<=k<=ke get the source-loaded coefficients; calc_e's Ex injection is
 * commented out (fdtdrav.h:681), and its Ez injection still adds
 * 1.2e-38 * V(t)/(ke-ks) (setZero's Cezj, fdtdrav.h:1016) on the box's Ez
 * edges is<=i<=ie, js<=j<=je, ks<=k<ke.  Needs ie > is; refused at
 * fdtd_finalize inside a CPML slab. */
int fdtd_add_lumped_source(fdtd_ctx *ctx, int dir, int is, int js, int ks,
                           int ie, int je, int ke, double resistance,
                           const fdtd_waveform *wave);

/* Zero-thickness conducting plate: overrides the averaged sigma_e of the
 * edges lying in the plate after the edge averaging and before the
 * coefficients, in call order (later plates win), as definePlateX / Y / Z
 * do (fdtdrav.h:1574-1678; marchingLoop applies definePlateZ,
 * fdtdrav.h:4183-4210).  axis 3: the plane k = ks, Ex edges is<=i<ie,
 * js<=j<=je get sigma[0] and Ey edges is<=i<=ie, js<=j<je get sigma[1];
 * axis 1 (plane i = is: Ey edges j<je, k<=ke get sigma[1], Ez edges j<=je,
 * k<ke get sigma[2]) and axis 2 (plane j = js: Ez edges i<=ie, k<ke get
 * sigma[2], Ex edges i<ie, k<=ke get sigma[0]) likewise.  sigma >= 0 (S/m;
 * the reference's PEC plates use 1e10). */
/* Zero-thickness conducting sector in the plane k = kz: a node (i, j) lies
 * in it when it is on the side of the line (i0,j0)-(i1,j1) where (i2,j2) is,
 * on the side of the line (i1,j1)-(i2,j2) where (i0,j0) is, and within
 * `radius` metres of the apex (i1, j1) -- tricheck (fdtdrav.h:56-128) on the
 * node coordinates i*dx, j*dy.  The Ex edges of such nodes with 0<=i<max_x,
 * 0<=j<=max_y get sigma_x, the Ey edges with 0<=i<=max_x, 0<=j<max_y get
 * sigma_y: defineSec (fdtdrav.h:1681-1744), applied after the edge averaging
 * and before the plates (its marchingLoop slot, fdtdrav.h:4152-4180, where
 * the call is disabled).  Sectors apply in call order.
 * 0<kz<nz-1, 0<=max_x<nx, 0<=max_y<ny, radius > 0. */
int fdtd_add_sector(fdtd_ctx *ctx, int i0, int j0, int i1, int j1, int i2, int j2, int kz, float radius,
                    int max_x, int max_y, float sigma_x, float sigma_y);

int fdtd_add_plate(fdtd_ctx *ctx, int axis, int is, int js, int ks, int ie, int je, int ke,
                   const float sigma[3]);

/* Lumped resistor on the Ez edges is<=i<=ie, js<=j<=je, ks<=k<ke: the Ez
 * coefficients of those edges are replaced with the resistor-loaded ones
 * of defineResistor (fdtdrav.h:1429-1484; the reference applies it after
 * the source and the CPML, fdtdrav.h:4296-4309).  dir must be 3 (the
 * reference's kernel loads Ez whatever its direction argument).  R > 0.
 * Resistors may not overlap a lumped source or a CPML slab. */
int fdtd_add_resistor(fdtd_ctx *ctx, int dir, int is, int js, int ks, int ie, int je, int ke,
                      double resistance);

/* Point probe of one component, sampled once per step after that
 * component's main update and before its CPML correction, as sampleField is
 * called in calc_e / calc_h (fdtdrav.h:234-246, 740, 347-391).  Returns the
 * probe index through *index (may be NULL). */
int fdtd_add_probe(fdtd_ctx *ctx, int comp, int i, int j, int k, int *index);

/* On-the-fly DFT of every probe at nfreq frequencies [Hz]:
 * X(f) = sum_m s[m] * exp(-j*2*pi*f*m*dt) * dt, accumulated in FP64.  No
 * counterpart runs in the reference; its cuFFT path calc_otm
 * (fdtdrav.h:2988-3051) is disabled at fdtdrav.h:4376. */
int fdtd_set_dft(fdtd_ctx *ctx, int nfreq, const double *freqs);

/* Design region is<=i<ie, js<=j<je, ks<=k<ke whose per-cell relative
 * permittivity is set by fdtd_set_design_eps and differentiated by
 * fdtd_adjoint_gradient.  No counterpart in the reference. */
int fdtd_set_design_region(fdtd_ctx *ctx, int is, int js, int ks,
                           int ie, int je, int ke);

/* Per-cell isotropic eps_r of the design region (x-fastest, n = its cell
 * count), painted after every box.  Allowed after finalize: coefficients are
 * rebuilt and the context is reset to step 0. */
int fdtd_set_design_eps(fdtd_ctx *ctx, const float *eps_r, size_t n);

/* Build coefficients, CPML profiles and source loading in the reference
 * order -- averaging, coefficients, CPML, then sources (fdtdrav.h:4134-4290)
 * -- upload them and allocate the step buffers for n_steps steps
 * (the reference's n_t_steps, fdtdrav.h:3588). */
int fdtd_finalize(fdtd_ctx *ctx, int n_steps);

/* ---- time loop ------------------------------------------------------ */

/* Advance nsteps steps (E then H each step, fdtdrav.h:3188-3422 and the
 * loop of fdtdrav.h:4320-4355).  Blocking; returns FDTD_E_NONFINITE if a
 * probe went non-finite. */
int fdtd_run(fdtd_ctx *ctx, int nsteps);

/* Advance one step: the E sweep then the H sweep of step steps_done.
 * Replaces one call of solver (fdtdrav.h:3188-3422), which the reference's
 * loop makes per step (fdtdrav.h:4320-4355).  Blocking; same results as
 * fdtd_run(ctx, 1). */
int fdtd_step(fdtd_ctx *ctx);

/* Rewind to step 0: zero fields, psi, probe history and DFT accumulators. */
int fdtd_reset(fdtd_ctx *ctx);

/* Run nsteps steps with a HIP event pair around every E and every H launch
 * and store the mean launch times in fdtd_info.  Same results as fdtd_run. */
int fdtd_profile(fdtd_ctx *ctx, int nsteps);

int fdtd_get_info(const fdtd_ctx *ctx, fdtd_info *info);

/* ---- outputs -------------------------------------------------------- */

/* Probe history [probe][step], steps_done columns per row (n >= probes *
 * steps_done).  Replaces the sampled-buffer read-back (fdtdrav.h:4364-4365). */
int fdtd_get_probes(fdtd_ctx *ctx, float *out, size_t n);

/* DFT accumulators [probe][freq][re, im] (n >= 2 * probes * freqs). */
int fdtd_get_dft(fdtd_ctx *ctx, double *out, size_t n);

/* One field component as FP32 SI values on the logical grid (n >= nx*ny*nz). */
int fdtd_get_field(fdtd_ctx *ctx, int comp, float *out, size_t n);

/* ---- port voltage / current (the reference's results1003/1004.dat) ---- */

/* Port of the run: the sampled-voltage box and the sampled-current box, each
 * given as is, js, ks, ie, je, ke with the reference's inclusive indices
 * (calc_index of the v_* / i_* extents, fdtdrav.h:3614-3650) and a direction
 * (1 x, 2 y, 3 z).
 *   V[m] = -dz / ((ie-is+1)(je-js+1)) * sum Ez over is<=i<=ie, js<=j<=je,
 *          ks<=k<ke   (voltMgpu case 3, fdtdrav.h:2252-2281; the reference
 *          samples only Ez, so vdir must be 3);
 *   I[m] = loop integral of H around the current box at the box's last
 *          plane along idir (currMgpu, fdtdrav.h:2301-2487), averaged with
 *          the next sample (correcI, fdtdrav.h:2490-2507).
 * The field values the two reductions read are sampled as probes (appended
 * after the caller's; they count in fdtd_get_probes).  One port per context,
 * set before fdtd_finalize.  Replaces voltMgpu / currMgpu / sampledAdj /
 * correcI of calcVoltCurrent (fdtdrav.h:3125-3185). */
int fdtd_set_port(fdtd_ctx *ctx, int vdir, int vis, int vjs, int vks, int vie, int vje, int vke,
                  int idir, int iis, int ijs, int iks, int iie, int ije, int ike);

/* V[m] and the averaged I[m] of the steps run so far, in volts and amperes
 * (FP32, the reference's summation order).  n >= steps_done. */
int fdtd_get_port(fdtd_ctx *ctx, float *v, float *i, size_t n);

/* Port spectrum at nfreq frequencies (Hz): out[8q .. 8q+7] = Re, Im of V(f),
 * I(f), S11 and Zin, with X(f) = sum_m x[m] exp(-2 pi i f m dt) (cuFFT's R2C
 * bin k at f = k / (n dt)) and c_zinS11's algebra (fdtdrav.h:2510-2546):
 * a, b = (V +- z0 I) / (2 sqrt(z0)), S11 = b / a, Zin = z0 (1 + S11) / (1 - S11).
 * Replaces calc_otm (fdtdrav.h:2988-3051, disabled in the reference).
 * z0 > 0; n >= 8 * nfreq. */
int fdtd_port_spectrum(fdtd_ctx *ctx, double z0, int nfreq, const double *freqs, double *out, size_t n);

/* Writes V and the averaged I as saveTDomain does (fdtdrav.h:2703-2717): one
 * "%e \n" line per step, to vpath and ipath (results1003.dat and
 * results1004.dat in the reference's working directory). */
int fdtd_save_port(fdtd_ctx *ctx, const char *vpath, const char *ipath);

/* ---- checkpoint / resume (no counterpart in the reference) ---------- */

int fdtd_checkpoint_size(const fdtd_ctx *ctx, size_t *bytes);
/* Full state: step index, E, H, psi, probe history, DFT accumulators, and a
 * checksum.  Resume is bitwise equal to an uninterrupted run. */
int fdtd_checkpoint_save(fdtd_ctx *ctx, void *buf, size_t bytes);
int fdtd_checkpoint_restore(fdtd_ctx *ctx, const void *buf, size_t bytes);

/* ---- adjoint (no counterpart in the reference) ---------------------- */

/* Run n_steps forward steps from step 0 (checkpointing every
 * checkpoint_every steps), then the discrete adjoint backward in segments,
 * recomputing each segment's forward fields from its checkpoint.
 * checkpoint_every = 0 stores the design region's E of every step instead
 * (no recompute); < 0 chooses: that history when it fits in half the free
 * device memory, else checkpoints every ceil(sqrt(n_steps)) steps.
 * J_out receives J; grad_out receives dJ/d(eps_r) of every design cell
 * (x-fastest over the design region, FP64). */
int fdtd_adjoint_gradient(fdtd_ctx *ctx, const fdtd_objective *obj,
                          int n_steps, int checkpoint_every,
                          double *J_out, double *grad_out, size_t n_grad);

/* Device scratch of fdtd_adjoint_gradient_noalloc, in bytes: a static array
 * the library holds on every GPU it runs on, until the process ends. */
#define FDTD_NOALLOC_SCRATCH_BYTES ((size_t)320 << 20)

/* fdtd_adjoint_gradient with checkpoint_every = 0 (the design-region E
 * history of the whole run is stored), without allocating: no hipMalloc or
 * hipFree, no stream or event creation.  Its device buffers (the FP32 adjoint
 * fields and psi, the cells' transposed quantities, the history, the
 * gradient accumulators) are carved from
 * the static scratch of FDTD_NOALLOC_SCRATCH_BYTES per GPU; a run that needs
 * more returns FDTD_E_NOMEM with the figure (adjoint128, 1000 steps, needs
 * 297 MiB).  Same results as fdtd_adjoint_gradient(..., 0, ...), bit for
 * bit.  Every call on one GPU shares that GPU's scratch: calls that overlap
 * on one GPU (other threads or contexts) race.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it). */
int fdtd_adjoint_gradient_noalloc(fdtd_ctx *ctx, const fdtd_objective *obj, int n_steps,
                                  double *J_out, double *grad_out, size_t n_grad);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_H */
