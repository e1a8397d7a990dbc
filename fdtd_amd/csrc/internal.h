// internal.h -- structures shared by the host scene builder (scene.cpp), the
// gfx950 kernels (kernels.hip) and the C ABI (api.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fdtd {

typedef std::int64_t i64;

// Block-scaled int16 storage: cells per exponent along x, exponent bias, and
// exponent bytes per block in the E (or H) exponent array (x, y, z, unused).
constexpr int BQ_BLOCK = 8;
constexpr int BQ_BIAS = 128;
constexpr int BQ_STRIDE = 4;

// Coefficient-class map index width.  Class 0..n-1 index the E and H tables.
enum ClassMode { CLS_UNIFORM = 0, CLS_U16 = 1, CLS_U32 = 2 };

// One event on a (i, j) column: a source edge or a probe cell at plane k.
// Events are sorted by column, then k; `col` = j * nx + i.
struct Event {
    int k;
    int kind;   // EV_SOURCE, or EV_PROBE + component (0..5)
    int id;     // source-edge record or probe index
    int col;
};
enum { EV_SOURCE = 0, EV_PROBE = 8 };

// Per source edge: Ez coefficients with the resistive load folded in
// (defineSourceZ, fdtdrav.h:1402-1421) and the waveform row.
struct SrcEdge {
    float ceze, cezhy, cezhx, cezj;   // cezj already carries the E storage scale
    int wave;                         // row of the per-node waveform table
    int pad[3];
};

// Sweep-kernel region variants: bit V_X / V_Y / V_Z set when the tiles of
// the variant can meet an x / y / z CPML slab or outer face.
enum SweepRegion { V_IN = 0, V_X = 1, V_Y = 2, V_Z = 4, V_GEN = 7 };

// Kernel argument block.  Field pointers point at (i,j,k) = (0,0,0) of an
// allocation that has one zero guard plane below k = 0 and one above
// k = nz-1, so every i+-1, j+-1, k+-1 neighbour address is inside it.
struct KArgs {
    void *ex, *ey, *ez, *hx, *hy, *hz;
    // block-scaled storage: the fields' exponent bytes, one per 8 x-cells,
    // [nz+2][ny][px/8] pointed at plane k = 0 (null for other storage types)
    std::uint8_t *qex, *qey, *qez, *qhx, *qhy, *qhz;
    // psi slabs: x [nz][ny][Lx], y [nz][Ly][px], z [Lz][ny][px]
    void *psi_eyx, *psi_ezx, *psi_hyx, *psi_hzx;
    void *psi_exy, *psi_ezy, *psi_hxy, *psi_hzy;
    void *psi_exz, *psi_eyz, *psi_hxz, *psi_hyz;
    const void *cls;           // class map [nz][ny][px] (u16/u32) or null
    const float4 *etab;        // [class][3] (ce, cb_a, cb_b, 0) for Ex, Ey, Ez
    const float4 *htab;        // [class][3] (ch, cb_a, cb_b, 0) for Hx, Hy, Hz
    // 1-D CPML profiles, indexed by the global cell index of the axis:
    // float4 (b, a, kappa, 0), E at [0,n), H at [n,2n)
    const float4 *px_prof, *py_prof, *pz_prof;
    const int *e_first, *h_first;   // per column: first event or -1
    const Event *events;
    const SrcEdge *srcs;
    const float *vpn;          // [wave][steps_cap]
    float *probe_hist;         // [probe][steps_cap]
    const double2 *phasor;     // [step][nfreq]
    double2 *dft;              // [probe][nfreq]
    float probe_unscale[6];    // stored -> SI per component
    float d[3];                // dx, dy, dz (cpsi = Cb * d)
    int nx, ny, nz, px;
    i64 plane;                 // px * ny
    int L[6];                  // x-, x+, y-, y+, z-, z+
    int Lx, Ly, Lz;            // slab widths summed per axis
    int steps_cap, nfreq;
    int kchunk;                // planes per z-march block
    int tiles_x, tiles_y;      // tile grid (for the XCD-aware remap)
    // Interior tile box [ib0, ib3) x [ib1, ib4) x [ib2, ib5) of (tile x,
    // tile y, z-block): every cell of it lies off the CPML slabs and the
    // outer faces.  With a box, three region variants share the sweep: V_IN
    // over the box, V_X over the tiles beside it in x (same rows and
    // z-blocks), V_GEN over the rows and z-blocks outside it.  Without one
    // (all zero), V_GEN covers every tile.
    int ib[6];
    // zsplit: V_Z takes the box's tile columns and rows in the z-blocks
    // outside it (cells that can meet only a z slab or face), V_GEN the rest.
    int zsplit;
    // Light-cone recompute: only the cell box [lc0, lc3) x [lc1, lc4) x
    // [lc2, lc5) is swept (lc[3] == 0: the whole grid), as sub = (tile x,
    // tile row, z-block of its first tile; tiles, rows, z-blocks), set for
    // each sweep's tiling when the step is launched (no interior box then).
    int lc[6];
    int sub[6];
    int rec;                   // record probes / DFT (0 during adjoint recompute)
    int e_uni, h_uni;          // E / H coefficients are one class (no class-map read)
    // design-region snapshot: E^{m+1} of the E edges in the box
    // [sb[0], sb[0]+sb[3]) x [sb[1], sb[1]+sb[4]) x [sb[2], sb[2]+sb[5]) is
    // written in SI units to snap[((m+1-snap_base)*3 + comp)*nb + q]
    float *snap;
    int sb[6];
    int snap_base;
};

// Tile counts of the region variants of a sweep (tile_of_region).
inline void region_tiles(const KArgs &a, int &n_in, int &n_x, int &n_z, int &n_gen)
{
    const int *ib = a.ib;
    const int tzs = (a.nz + a.kchunk - 1) / a.kchunk;
    if (ib[5] == ib[2]) {   // no box: every tile, or the light-cone tiles
        n_in = n_x = n_z = 0;
        n_gen = a.sub[3] > 0 ? a.sub[3] * a.sub[4] * a.sub[5] : a.tiles_x * a.tiles_y * tzs;
        return;
    }
    const int bx = ib[3] - ib[0], by = ib[4] - ib[1], bz = ib[5] - ib[2];
    n_in = bx * by * bz;
    n_x = (a.tiles_x - bx) * by * bz;
    n_z = a.zsplit ? bx * by * (tzs - bz) : 0;
    n_gen = a.tiles_x * a.tiles_y * (tzs - bz) - n_z + a.tiles_x * (a.tiles_y - by) * bz;
}

// Adjoint argument block (FP32 SI adjoint state; same layouts as the forward).
struct AArgs {
    float *eb[3], *hb[3];      // dJ/dE, dJ/dH (guard-plane layout of the fields)
    // where a reverse step writes its dJ/dE and dJ/dH: eb / hb themselves
    // (in place), or, for the fused interior sweep, the other buffer of a
    // pair (the host swaps the two after each step).  The H sweep reads
    // dJ/dE from ebn (the values the E sweep just wrote).
    float *ebn[3], *hbn[3];
    // interior tiles (each its index in the box, x fastest): ntl > 0
    // restricts the interior E / H sweeps to tl (the tiles the fused sweep
    // does not take), fl the fused sweep's
    const int *tl;
    int ntl;
    const int *fl;
    int nfl;
    // transposed quantities of every cell, same layout: the factors its
    // adjoint contributes to each field difference it reads, written by the
    // sweep that updates it, gathered by the other one
    float *hq[6];              // H cells: sy sz tz tx rx ry
    float *eq[6];              // E cells: ax bx ay by az bz
    float *pb[12];             // dJ/dpsi: eyx ezx hyx hzx exy ezy hxy hzy exz eyz hxz hyz
    const float *g;            // [probe][gsteps] dJ/d(probe sample m)
    int gsteps;
    const float *hist;         // [slot][3][nb] E snapshots (SI), slot = step - hist_base
    int hist_base;
    const float4 *dcoef;       // [3][nb] (dCe/deps, 2 eps0/den, Ce, 0); zero off the design
    double *gedge;             // [3][nb] dJ/d(averaged eps) per E edge
    int sb[6];                 // dilated design box (i0, j0, k0, bx, by, bz)
    int src_j0, src_j1;        // rows j whose row j or j+1 holds a source edge (empty: j0 > j1)
};

// Host-side scene, filled by the fdtd_* scene calls.
struct Box {
    int b[6];
    float mat[12];  // eps_r xyz, sigma_e xyz, mu_r xyz, sigma_m xyz
};
struct Source {
    int b[6];
    double R;
    int wtype;
    double amp, t0, tau, freq;
    int dir = 3;   // 3: z-directed (loaded, driven Ez edges); 1: x-directed (loaded Ex edges)
};
struct Probe {
    int comp, i, j, k;
};
struct Plate {
    int axis, b[6];
    float sigma[3];
};
struct Sector {
    int p[6];             // i0 j0 i1 j1 i2 j2 (apex i1, j1)
    int kz, max_x, max_y;
    float radius, sigma[2];
};
struct Resistor {
    int b[6];
    double R;
};

struct CpmlParams {
    int thick[6];
    double order, sigma_max[3], sigma_factor, kappa_max, alpha_max, alpha_min;
};

// Result of the host build: everything the device needs.
struct Built {
    int mode;                       // ClassMode
    bool e_uniform, h_uniform;
    int ncls;
    std::vector<std::uint32_t> cls; // per cell (padded pitch), empty if uniform
    std::vector<float> etab, htab;  // [class][3][4], in stored (scaled) units
    std::vector<float> etab_si, htab_si;   // the same in SI units (adjoint)
    std::vector<float> prof[3];     // [2n][4] per axis
    std::vector<int> e_first, h_first;
    std::vector<Event> events;
    std::vector<SrcEdge> srcs;
    std::vector<SrcEdge> srcs_si;   // SI coefficients (adjoint)
    int src_j0 = 1, src_j1 = 0;     // rows j whose row j or j+1 holds a source edge: [src_j0, src_j1]
    std::vector<float> vpn;         // [wave][steps_cap]
    std::vector<double> phasor;     // [step][nfreq][2]
    std::vector<float> dcoef;       // [3][nb][4] design-edge eps derivatives (adjoint)
    float probe_unscale[6];
    float scale[6];                 // stored = SI * scale, per component
    double sigma_max_used[3];
};

struct SceneIn {
    int nx, ny, nz, px;
    double dx, dy, dz, dt, eps0, mu0;
    int storage;
    CpmlParams cpml;
    std::vector<Box> boxes;
    std::vector<Source> sources;
    std::vector<Sector> sectors;
    std::vector<Plate> plates;
    std::vector<Resistor> resistors;
    std::vector<Probe> probes;
    std::vector<double> freqs;
    int design[6];
    std::vector<float> design_eps;
    int steps_cap;
};

// scene.cpp
int build_scene(const SceneIn &in, Built &out, std::string &err);

// port.cpp: port voltage / current boxes (fdtd_set_port) and the probes the
// reductions read.  v: voltage samples in summation order; ia..id: the four
// line sums of the current loop, I = ca*sum(ia) + cb*sum(ib) - cc*sum(ic) -
// cd*sum(id).  Sample indices are into `samples`.
struct PortDef {
    int vdir, vbox[6];   // is, js, ks, ie, je, ke
    int idir, ibox[6];
};
struct PortPlan {
    PortDef def;
    std::vector<Probe> samples;
    std::vector<int> v, ia, ib, ic, id;
    float vscale, ca, cb, cc, cd;
};
int port_plan(const PortDef &d, int nx, int ny, int nz, float dx, float dy, float dz, PortPlan &p,
              std::string &err);
// hist: [sample][ns] probe histories of plan.samples; v, cur: [ns]
void port_series(const PortPlan &p, const float *hist, size_t ns, float *v, float *cur);
// out: [nfreq][8] = Re/Im of V(f), I(f), S11, Zin
void port_spectrum(const float *v, const float *cur, size_t ns, double dt, double z0, int nfreq,
                   const double *freqs, double *out);
int port_save(const float *x, size_t ns, const char *path);

}  // namespace fdtd
