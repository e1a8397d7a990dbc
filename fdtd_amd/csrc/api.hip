// api.hip -- the C ABI declared in include/fdtd.h.
//
// Replaces the reference's host orchestration, marchingLoop
// (fdtdrav.h:3425-4387), and its per-step driver, solver (fdtdrav.h:3188-3422),
// for one device: the z-slab split and ghost-plane copies of the reference
// collapse into a single domain whose neighbour planes are plain HBM reads.
// All work goes to one non-blocking stream; a step is the E sweep then the
// H sweep (one launch per non-empty region variant each), with no host
// synchronisation between them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fdtd.h"
#include "internal.h"

namespace fdtd {
hipError_t launch_half_step(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s);
hipError_t launch_adj(int which, int mode, int euni, int huni, const KArgs &a, const AArgs &b, int m, int fstep,
                      int sten, hipStream_t s);
hipError_t launch_snap(int storage, const KArgs &a, float *hist, int slot, hipStream_t s);
bool adj_stores_quantities(int nx);
void adj_fuse_tiles(const KArgs &a, const std::vector<int> &e_first, const std::vector<int> &h_first,
                    const int sb[6], std::vector<int> &quiet, std::vector<int> &noisy);
hipError_t launch_pair_half_step(bool efield, int storage, int mode, const KArgs &a, int m, hipStream_t s,
                                 bool gen_cells);
int tile_rows();
int tile_cols();
int pair_tile_cols();
}  // namespace fdtd

using namespace fdtd;

struct fdtd_ctx {
    SceneIn in;
    bool cpml_set = false;
    bool finalized = false;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Built built;
    KArgs ka{};
    int e_mode = 0, h_mode = 0;
    int steps_done = 0;
    bool has_port = false;          // fdtd_set_port: its samples are probes port_base ..
    int port_base = 0;
    PortPlan port;
    double last_run_ms = 0.0, e_ms = 0.0, h_ms = 0.0;
    size_t esz = 4;                 // bytes per stored value
    i64 field_elems = 0;            // per field allocation, guard planes included
    std::vector<void *> allocs;     // every device allocation
    long n_allocs = 0;              // device allocations attempted (failure injection)
    // Field and psi allocation bases (the sweeps update them in place).
    void *fields[6] = {};
    void *psi[12] = {};
    // block-scaled storage: the E and H exponent arrays (BQ_STRIDE bytes per
    // 8 x-cells: the x, y, z components' exponents), guard planes included,
    // and the component pointers into them (qexp[f] = qarr[f / 3] + f % 3)
    std::uint8_t *qarr[2] = {};
    std::uint8_t *qexp[6] = {};
    size_t qarr_bytes = 0;          // per array
    bool pair = false;              // 16-bit storage: two cells per thread in the half-step kernels
    bool pair_gen_cells = false;    // ... except the general variant, which runs one cell per thread
    int pair_tiles_x = 0, pair_kchunk = 0;
    int kchunk_h = 0;               // planes per block of the one-cell H sweep (E: ka.kchunk)
    int box[2][6] = {};             // interior tile box per sweep (0 = E, 1 = H)
    double step_ms = 0.0;
    size_t psi_elems[12] = {};
    uint64_t dev_bytes = 0;
    // SI coefficient tables for the adjoint
    const float4 *etab_si = nullptr, *htab_si = nullptr;
    const SrcEdge *srcs_si = nullptr;
    const float4 *dcoef = nullptr;
    double adj_fwd_ms = 0.0, adj_total_ms = 0.0;
    int adj_segments = 0;
    uint64_t adj_bytes = 0;         // peak device memory of the last adjoint run
    hipEvent_t adj_ev[3] = {};      // timing events of the adjoint (made at create)
    unsigned char *scratch = nullptr;   // this device's copy of the static scratch
};

// Device scratch of fdtd_adjoint_gradient_noalloc: a file-scope device array,
// one copy per GPU, held from the library's first use on that GPU until the
// process ends.  Its name is a host-side stand-in, never a device address:
// host code uses only the per-device address hipGetSymbolAddress returns
// (scratch_address), for kernel arguments, memsets and copies.  Every call
// on one GPU shares that GPU's copy, and it keeps the last call's contents.
static __device__ __align__(256) unsigned char scratch[FDTD_NOALLOC_SCRATCH_BYTES];

namespace {

// Text of the last refused fdtd_create on this thread (no context to hold it).
thread_local char g_create_err[256] = "no error";

int create_fail(int rc, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_create_err, sizeof g_create_err, fmt, ap);
    va_end(ap);
    return rc;
}

// The address of this device's copy of scratch, looked up once per device
// with that device current (fdtd_create's device); nullptr on failure.
unsigned char *scratch_address(int device)
{
    static std::mutex mu;
    static std::vector<unsigned char *> addr;
    static std::vector<char> looked;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return nullptr;
    if ((size_t)device >= addr.size()) {
        addr.resize((size_t)device + 1, nullptr);
        looked.resize((size_t)device + 1, 0);
    }
    if (!looked[(size_t)device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) {
            addr[(size_t)device] = (unsigned char *)p;
            looked[(size_t)device] = 1;
        } else {
            (void)hipGetLastError();   // not cached: a later context may retry
        }
    }
    return addr[(size_t)device];
}

int fail(fdtd_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(c, e_ == hipErrorOutOfMemory ? FDTD_E_NOMEM : FDTD_E_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                       \
    } while (0)

int dev_alloc(fdtd_ctx *c, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    // failure injection: FDTD_FAIL_ALLOC=n makes the context's n-th device
    // allocation fail as out of memory (tests of the error paths)
    ++c->n_allocs;
    if (const char *fa = std::getenv("FDTD_FAIL_ALLOC"))
        if (std::atol(fa) == c->n_allocs)
            return fail(c, FDTD_E_NOMEM, "hipMalloc(%zu) failed: injected (FDTD_FAIL_ALLOC=%s)", bytes, fa);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, FDTD_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs.push_back(*p);
    c->dev_bytes += bytes;
    return FDTD_OK;
}

int upload(fdtd_ctx *c, void **p, const void *src, size_t bytes)
{
    int rc = dev_alloc(c, p, bytes);
    if (rc) return rc;
    if (bytes && src) HIPCHK(c, hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, c->stream));
    return FDTD_OK;
}

void free_device(fdtd_ctx *c)
{
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    c->dev_bytes = 0;
    for (auto &f : c->fields) f = nullptr;
    for (auto &p : c->psi) p = nullptr;
    for (auto &q : c->qexp) q = nullptr;
    for (auto &q : c->qarr) q = nullptr;
}

// The buffers holding the current state.
void *fld(const fdtd_ctx *c, int f) { return c->fields[f]; }
void *psv(const fdtd_ctx *c, int q) { return c->psi[q]; }

bool box_in_grid(const fdtd_ctx *c, int i, int j, int k)
{
    return i >= 0 && j >= 0 && k >= 0 && i < c->in.nx && j < c->in.ny && k < c->in.nz;
}

// Zero every piece of time-dependent state.
int zero_state(fdtd_ctx *c)
{
    for (int f = 0; f < 6; ++f) {
        if (c->fields[f]) HIPCHK(c, hipMemsetAsync(c->fields[f], 0, (size_t)c->field_elems * c->esz, c->stream));
    }
    for (int q = 0; q < 2; ++q)
        if (c->qarr[q]) HIPCHK(c, hipMemsetAsync(c->qarr[q], 0, c->qarr_bytes, c->stream));
    for (int p = 0; p < 12; ++p)
        if (c->psi_elems[p] && c->psi[p])
            HIPCHK(c, hipMemsetAsync(c->psi[p], 0, c->psi_elems[p] * c->esz, c->stream));
    const size_t np = c->in.probes.size(), nf = c->in.freqs.size();
    if (np) HIPCHK(c, hipMemsetAsync(c->ka.probe_hist, 0, np * (size_t)c->in.steps_cap * sizeof(float), c->stream));
    if (np && nf) HIPCHK(c, hipMemsetAsync(c->ka.dft, 0, np * nf * sizeof(double2), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->steps_done = 0;
    return FDTD_OK;
}

// Upload the material-dependent tables (class map, E/H tables, sources).
int upload_tables(fdtd_ctx *c);

// Planes per z-march block.  Many short marches keep every CU's wave slots
// full through the tail of a launch; each block re-reads one plane of the
// carried z-neighbour, so a block of `planes` costs 1/planes extra on that
// stream.  Measured on slab512: the FP32 E sweep is fastest at 4 planes, the
// FP32 H sweep at 2, the 16-bit pair sweeps at 8 (FDTD_CHUNKS overrides all
// three with a number of blocks along z; FDTD_E_PLANES, FDTD_H_PLANES and
// FDTD_PAIR_PLANES set one sweep's planes per block before the small-grid
// halving).
int choose_chunks(const fdtd_ctx *c, int tiles, int planes, const char *planes_env)
{
    const char *ov = std::getenv("FDTD_CHUNKS");
    if (const char *pe = std::getenv(planes_env)) planes = std::max(1, std::atoi(pe));
    // Small grids: halve the planes per block until a launch has 1024
    // workgroups (4 per CU), down to one plane (FDTD_MIN_WG overrides 1024).
    // Measured: 64^3 FP32 18.4 -> 22.7 G, FP16 10.3 -> 22.3 G; 128x128x64
    // FP32 keeps its 4-plane E blocks (2-plane ones measured 3.5 % slower).
    const char *mw = std::getenv("FDTD_MIN_WG");
    const long min_wg = mw ? std::atol(mw) : 1024;
    while (planes > 1 && (long)tiles * (c->in.nz / planes) < min_wg) planes /= 2;
    int nchunks = ov ? std::atoi(ov)
                     : std::min((32 * 1024 * 8 / planes + tiles - 1) / tiles, c->in.nz / planes);
    nchunks = std::max(1, std::min(nchunks, c->in.nz));
    return (c->in.nz + nchunks - 1) / nchunks;
}

// Interior tile box of one sweep (tiles tile_cols wide and tile_rows() high,
// z-blocks of kchunk planes): along each axis, the run of tiles whose cell
// range reaches neither a CPML slab nor an outer face (nor the grid's end).
// The x-edge variant takes the tiles beside the box in x, the general
// variant the rows and z-blocks outside it: z-edge variants of their own
// measured slower on 512x512x256 (their launch tails cost more than their
// leaner code saves).  FDTD_INTERIOR=0, or an axis without interior tiles,
// leaves the box empty (z range [0, 0)) and every tile to the general
// variant.  (On 128x128x64 an x-edge launch over the interior rows and
// z-blocks beside a general one measured slower than the general variant
// alone: 32.4 against 26.9 ms for 1000 steps; the two half-size launches
// of a small grid leave CUs idle.)
void sweep_split(fdtd_ctx *c, int sweep, int tile_cols, int kchunk)
{
    const SceneIn &in = c->in;
    const int *L = in.cpml.thick;
    int *ib = c->box[sweep];
    for (int q = 0; q < 6; ++q) ib[q] = 0;
    const char *ev = std::getenv("FDTD_INTERIOR");
    if (ev && std::strcmp(ev, "0") == 0) return;
    const int by = tile_rows();
    const int nts[3] = {(in.nx + tile_cols - 1) / tile_cols, (in.ny + by - 1) / by, (in.nz + kchunk - 1) / kchunk};
    const int w[3] = {tile_cols, by, kchunk}, nn[3] = {in.nx, in.ny, in.nz};
    int box[6];
    for (int ax = 0; ax < 3; ++ax) {
        const int lo_cells = std::max(L[2 * ax], 1), hi_cells = nn[ax] - std::max(L[2 * ax + 1], 1);
        int lo = nts[ax], hi = 0;
        for (int t = 0; t < nts[ax]; ++t)
            if (t * w[ax] >= lo_cells && (t + 1) * w[ax] <= hi_cells) {
                lo = std::min(lo, t);
                hi = t + 1;
            }
        if (hi <= lo) return;
        box[ax] = lo;
        box[ax + 3] = hi;
    }
    for (int q = 0; q < 6; ++q) ib[q] = box[q];
}

int do_finalize(fdtd_ctx *c)
{
    const SceneIn &in = c->in;
    Built &b = c->built;
    std::string err;
    {
        // The sweep kernels address every array with 32-bit byte offsets:
        // refuse a field array (with its two guard planes) of 4 GiB or more
        // before building anything.
        const size_t fb = (size_t)in.px * in.ny * (in.nz + 2) * (in.storage == FDTD_STORE_F32 ? 4 : 2);
        if (fb >= ((size_t)1 << 32))
            return fail(c, FDTD_E_UNSUPPORTED,
                        "grid too large: each field array (with its two guard planes) must be under 4 GiB "
                        "(%zu bytes per field)", fb);
    }
    if (build_scene(in, b, err) != 0) return fail(c, FDTD_E_INVALID, "%s", err.c_str());
    c->e_mode = b.e_uniform ? CLS_UNIFORM : b.mode;
    c->h_mode = b.h_uniform ? CLS_UNIFORM : b.mode;
    c->esz = in.storage == FDTD_STORE_F32 ? 4 : 2;

    KArgs &a = c->ka;
    a = KArgs{};
    a.nx = in.nx; a.ny = in.ny; a.nz = in.nz; a.px = in.px;
    a.plane = (i64)in.px * in.ny;
    for (int q = 0; q < 6; ++q) a.L[q] = in.cpml.thick[q];
    a.Lx = a.L[0] + a.L[1];
    a.Ly = a.L[2] + a.L[3];
    a.Lz = a.L[4] + a.L[5];
    a.d[0] = (float)in.dx; a.d[1] = (float)in.dy; a.d[2] = (float)in.dz;
    a.steps_cap = in.steps_cap;
    a.nfreq = (int)in.freqs.size();
    for (int q = 0; q < 6; ++q) a.probe_unscale[q] = b.probe_unscale[q];
    a.tiles_x = (in.nx + tile_cols() - 1) / tile_cols();
    a.tiles_y = (in.ny + tile_rows() - 1) / tile_rows();
    a.kchunk = choose_chunks(c, a.tiles_x * a.tiles_y, 4, "FDTD_E_PLANES");
    c->kchunk_h = choose_chunks(c, a.tiles_x * a.tiles_y, 2, "FDTD_H_PLANES");
    // 16-bit storage (FP16, BF16, block-scaled int16) sweeps a cell pair per
    // thread (128-cell tile rows); FDTD_PAIR=0 selects the one-cell kernels
    // instead (same stored bits)
    const char *pv = std::getenv("FDTD_PAIR");
    c->pair = in.storage != FDTD_STORE_F32 && !(pv && std::strcmp(pv, "0") == 0);
    // The pair sweeps' general variant (slab rows and z-blocks) runs as the
    // one-cell kernel (same stored bits): on 512x512x256 FP16 the step took
    // 0.772 ms against 0.788 with the pair kernel's own general variant,
    // which holds two cells of slab state per thread.  FDTD_GEN16=pair
    // selects the latter.
    const char *gv = std::getenv("FDTD_GEN16");
    c->pair_gen_cells = !(gv && std::strcmp(gv, "pair") == 0);
    c->pair_tiles_x = (in.nx + pair_tile_cols() - 1) / pair_tile_cols();
    c->pair_kchunk = choose_chunks(c, c->pair_tiles_x * a.tiles_y, 8, "FDTD_PAIR_PLANES");
    a.rec = 1;
    a.snap = nullptr;

    // fields: one zero guard plane below and above
    c->field_elems = a.plane * (in.nz + 2);
    // The sweep kernels address every array with 32-bit byte offsets.
    const size_t cls_bytes = (size_t)a.plane * in.nz * (b.mode == CLS_U32 ? 4 : 2);
    if ((size_t)c->field_elems * c->esz >= ((size_t)1 << 32) || cls_bytes >= ((size_t)1 << 32))
        return fail(c, FDTD_E_UNSUPPORTED,
                    "grid too large: each field array (with its two guard planes) and the class map "
                    "must be under 4 GiB (%zu bytes per field)", (size_t)c->field_elems * c->esz);
    for (int f = 0; f < 6; ++f) {
        int rc = dev_alloc(c, &c->fields[f], (size_t)c->field_elems * c->esz);
        if (rc) return rc;
    }
    if (in.storage == FDTD_STORE_BQ16) {
        // BQ_STRIDE exponent bytes per 8 x-cells of E and of H (px is a
        // multiple of 64)
        c->qarr_bytes = (size_t)c->field_elems / BQ_BLOCK * BQ_STRIDE;
        for (int q = 0; q < 2; ++q) {
            void *d = nullptr;
            int rc = dev_alloc(c, &d, c->qarr_bytes);
            if (rc) return rc;
            c->qarr[q] = (std::uint8_t *)d;
        }
        for (int f = 0; f < 6; ++f) c->qexp[f] = c->qarr[f / 3] + f % 3;
    }
    // interior / general split of the two sweeps
    {
        const int tcols = c->pair ? pair_tile_cols() : tile_cols();
        sweep_split(c, 0, tcols, c->pair ? c->pair_kchunk : a.kchunk);
        sweep_split(c, 1, tcols, c->pair ? c->pair_kchunk : c->kchunk_h);
    }
    // psi slabs
    const size_t px_elems = (size_t)in.nz * in.ny * a.Lx;
    const size_t py_elems = (size_t)in.nz * a.Ly * in.px;
    const size_t pz_elems = (size_t)a.Lz * in.ny * in.px;
    for (int p = 0; p < 12; ++p) {
        c->psi_elems[p] = p < 4 ? px_elems : p < 8 ? py_elems : pz_elems;
        int rc = dev_alloc(c, &c->psi[p], std::max<size_t>(c->psi_elems[p], 1) * c->esz);
        if (rc) return rc;
    }
    a.e_uni = b.e_uniform ? 1 : 0;
    a.h_uni = b.h_uniform ? 1 : 0;
    int rc;
    void *tmp;
    for (int ax = 0; ax < 3; ++ax) {
        if ((rc = upload(c, &tmp, b.prof[ax].data(), b.prof[ax].size() * sizeof(float)))) return rc;
        (ax == 0 ? a.px_prof : ax == 1 ? a.py_prof : a.pz_prof) = (const float4 *)tmp;
    }
    if ((rc = upload(c, &tmp, b.e_first.data(), b.e_first.size() * sizeof(int)))) return rc;
    a.e_first = (const int *)tmp;
    if ((rc = upload(c, &tmp, b.h_first.data(), b.h_first.size() * sizeof(int)))) return rc;
    a.h_first = (const int *)tmp;
    if ((rc = upload(c, &tmp, b.events.data(), b.events.size() * sizeof(Event)))) return rc;
    a.events = (const Event *)tmp;
    if ((rc = upload(c, &tmp, b.vpn.data(), b.vpn.size() * sizeof(float)))) return rc;
    a.vpn = (const float *)tmp;
    if ((rc = upload(c, &tmp, b.phasor.data(), b.phasor.size() * sizeof(double)))) return rc;
    a.phasor = (const double2 *)tmp;
    const size_t np = in.probes.size(), nf = in.freqs.size();
    if ((rc = dev_alloc(c, &tmp, np * (size_t)in.steps_cap * sizeof(float)))) return rc;
    a.probe_hist = (float *)tmp;
    if ((rc = dev_alloc(c, &tmp, np * nf * sizeof(double2)))) return rc;
    a.dft = (double2 *)tmp;
    if ((rc = upload_tables(c))) return rc;
    if ((rc = zero_state(c))) return rc;
    c->finalized = true;
    return FDTD_OK;
}

int upload_tables(fdtd_ctx *c)
{
    Built &b = c->built;
    KArgs &a = c->ka;
    int rc;
    void *tmp;
    if ((rc = upload(c, &tmp, b.etab.data(), b.etab.size() * sizeof(float)))) return rc;
    a.etab = (const float4 *)tmp;
    if ((rc = upload(c, &tmp, b.htab.data(), b.htab.size() * sizeof(float)))) return rc;
    a.htab = (const float4 *)tmp;
    if ((rc = upload(c, &tmp, b.srcs.data(), b.srcs.size() * sizeof(SrcEdge)))) return rc;
    a.srcs = (const SrcEdge *)tmp;
    if ((rc = upload(c, &tmp, b.etab_si.data(), b.etab_si.size() * sizeof(float)))) return rc;
    c->etab_si = (const float4 *)tmp;
    if ((rc = upload(c, &tmp, b.htab_si.data(), b.htab_si.size() * sizeof(float)))) return rc;
    c->htab_si = (const float4 *)tmp;
    if ((rc = upload(c, &tmp, b.srcs_si.data(), b.srcs_si.size() * sizeof(SrcEdge)))) return rc;
    c->srcs_si = (const SrcEdge *)tmp;
    c->dcoef = nullptr;
    if (!b.dcoef.empty()) {
        if ((rc = upload(c, &tmp, b.dcoef.data(), b.dcoef.size() * sizeof(float)))) return rc;
        c->dcoef = (const float4 *)tmp;
    }
    a.cls = nullptr;
    if (b.mode == CLS_U16) {
        std::vector<uint16_t> m16(b.cls.size());
        for (size_t q = 0; q < b.cls.size(); ++q) m16[q] = (uint16_t)b.cls[q];
        if ((rc = upload(c, &tmp, m16.data(), m16.size() * sizeof(uint16_t)))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        a.cls = tmp;
    } else if (b.mode == CLS_U32) {
        if ((rc = upload(c, &tmp, b.cls.data(), b.cls.size() * sizeof(uint32_t)))) return rc;
        a.cls = tmp;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<uint32_t>().swap(b.cls);   // host copy no longer needed
    return FDTD_OK;
}

// Point a kernel argument block at the field and psi buffers; the sweeps
// read and write them in place.
void bind_sets(const fdtd_ctx *c, KArgs &a)
{
    const size_t g = (size_t)c->ka.plane * c->esz;   // guard plane
    void **rd[6] = {&a.ex, &a.ey, &a.ez, &a.hx, &a.hy, &a.hz};
    for (int f = 0; f < 6; ++f) *rd[f] = (char *)c->fields[f] + g;
    std::uint8_t **qd[6] = {&a.qex, &a.qey, &a.qez, &a.qhx, &a.qhy, &a.qhz};
    for (int f = 0; f < 6; ++f)
        *qd[f] = c->qexp[f] ? c->qexp[f] + (size_t)c->ka.plane / BQ_BLOCK * BQ_STRIDE : nullptr;
    void **pr[12] = {&a.psi_eyx, &a.psi_ezx, &a.psi_hyx, &a.psi_hzx, &a.psi_exy, &a.psi_ezy,
                     &a.psi_hxy, &a.psi_hzy, &a.psi_exz, &a.psi_eyz, &a.psi_hxz, &a.psi_hyz};
    for (int p = 0; p < 12; ++p) *pr[p] = c->psi[p];
}

// One time step m.  base carries the per-call flags (rec, snapshot); the
// tile geometry and buffer pointers are filled in here.
hipError_t do_step(fdtd_ctx *c, const KArgs &base, int m, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                   hipEvent_t ev2 = nullptr, hipEvent_t ev3 = nullptr)
{
    hipError_t e;
    KArgs a = base;
    a.tiles_x = c->pair ? c->pair_tiles_x : c->ka.tiles_x;
    a.tiles_y = c->ka.tiles_y;
    a.kchunk = c->pair ? c->pair_kchunk : c->ka.kchunk;
    bind_sets(c, a);
    auto half = [c](bool efield, int storage, int mode, const KArgs &ka, int mm, hipStream_t st) {
        return c->pair ? launch_pair_half_step(efield, storage, mode, ka, mm, st, c->pair_gen_cells)
                       : launch_half_step(efield, storage, mode, ka, mm, st);
    };
    auto split = [&](int sweep) {
        for (int q = 0; q < 6; ++q) a.ib[q] = c->box[sweep][q];
        a.sub[3] = 0;
        if (a.lc[3] > 0) {   // light-cone box: its tiles of this sweep's tiling, general variant
            for (int q = 0; q < 6; ++q) a.ib[q] = 0;
            const int tc = c->pair ? pair_tile_cols() : tile_cols(), by = tile_rows();
            a.sub[0] = a.lc[0] / tc;
            a.sub[3] = (a.lc[3] + tc - 1) / tc - a.sub[0];
            a.sub[1] = a.lc[1] / by;
            a.sub[4] = (a.lc[4] + by - 1) / by - a.sub[1];
            a.sub[2] = a.lc[2] / a.kchunk;
            a.sub[5] = (a.lc[5] + a.kchunk - 1) / a.kchunk - a.sub[2];
        }
    };
    split(0);
    if (ev0) (void)hipEventRecord(ev0, c->stream);
    if ((e = half(true, c->in.storage, c->e_mode, a, m, c->stream)) != hipSuccess) return e;
    if (ev1) (void)hipEventRecord(ev1, c->stream);
    if (ev2) (void)hipEventRecord(ev2, c->stream);
    if (!c->pair) a.kchunk = c->kchunk_h;
    split(1);
    e = half(false, c->in.storage, c->h_mode, a, m, c->stream);
    if (ev3) (void)hipEventRecord(ev3, c->stream);
    return e;
}

// A stream of one call, drained and destroyed on every return path.
struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard()
    {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};

// The events of one call, destroyed on every return path.
struct EventSet {
    std::vector<hipEvent_t> v;
    ~EventSet()
    {
        for (hipEvent_t e : v)
            if (e) (void)hipEventDestroy(e);
    }
};

int run_steps(fdtd_ctx *c, int nsteps, bool profile)
{
    if (!c->finalized) return fail(c, FDTD_E_STATE, "fdtd_finalize has not been called");
    if (nsteps < 0) return fail(c, FDTD_E_INVALID, "nsteps < 0");
    if (c->steps_done + nsteps > c->in.steps_cap)
        return fail(c, FDTD_E_RANGE, "run past the finalized horizon (%d + %d > %d)", c->steps_done, nsteps,
                    c->in.steps_cap);
    EventSet all;
    all.v.assign(2 + (profile ? (size_t)4 * nsteps : 0), nullptr);
    for (auto &e : all.v) HIPCHK(c, hipEventCreate(&e));
    hipEvent_t t0 = all.v[0], t1 = all.v[1];
    hipEvent_t *evs = all.v.data() + 2;
    const int m0 = c->steps_done;
    HIPCHK(c, hipEventRecord(t0, c->stream));
    for (int s = 0; s < nsteps; ++s) {
        const int m = m0 + s;
        if (profile)
            HIPCHK(c, do_step(c, c->ka, m, evs[4 * s + 0], evs[4 * s + 1], evs[4 * s + 2], evs[4 * s + 3]));
        else
            HIPCHK(c, do_step(c, c->ka, m));
    }
    HIPCHK(c, hipEventRecord(t1, c->stream));
    HIPCHK(c, hipEventSynchronize(t1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, t0, t1));
    c->last_run_ms = ms;
    if (profile && nsteps > 0) {
        double se = 0, sh = 0;
        for (int s = 0; s < nsteps; ++s) {
            float a = 0, b = 0;
            HIPCHK(c, hipEventElapsedTime(&a, evs[4 * s + 0], evs[4 * s + 1]));
            HIPCHK(c, hipEventElapsedTime(&b, evs[4 * s + 2], evs[4 * s + 3]));
            se += a;
            sh += b;
        }
        c->e_ms = se / nsteps;
        c->h_ms = sh / nsteps;
        c->step_ms = c->e_ms + c->h_ms;
    }
    c->steps_done = m0 + nsteps;
    // blow-up detection on the probe samples of this call
    const size_t np = c->in.probes.size();
    if (np && nsteps > 0) {
        std::vector<float> h(np * (size_t)nsteps);
        HIPCHK(c, hipMemcpy2D(h.data(), (size_t)nsteps * sizeof(float), c->ka.probe_hist + m0,
                              (size_t)c->in.steps_cap * sizeof(float), (size_t)nsteps * sizeof(float), np,
                              hipMemcpyDeviceToHost));
        for (size_t q = 0; q < h.size(); ++q)
            if (!std::isfinite(h[q]))
                return fail(c, FDTD_E_NONFINITE, "probe %zu is non-finite at step %zu", q / nsteps,
                            m0 + q % nsteps);
    }
    return FDTD_OK;
}

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h = 1469598103934665603ull)
{
    for (size_t q = 0; q < n; ++q) { h ^= p[q]; h *= 1099511628211ull; }
    return h;
}

struct CkHeader {
    char magic[8];
    int32_t version, nx, ny, nz, storage, step, nprobe, nfreq, steps_cap, pad;
    uint64_t payload, checksum;
};

}  // namespace


// ---- adjoint orchestration --------------------------------------------------
namespace {

// Device memory the forward state occupies (fields + psi), per checkpoint.
size_t state_bytes(const fdtd_ctx *c)
{
    size_t b = 6 * (size_t)c->field_elems * c->esz;
    for (int q = 0; q < 12; ++q) b += c->psi_elems[q] * c->esz;
    return b;
}

int copy_state(fdtd_ctx *c, char *dst, bool save)
{
    size_t off = 0;
    for (int f = 0; f < 6; ++f) {
        const size_t n = (size_t)c->field_elems * c->esz;
        if (save) HIPCHK(c, hipMemcpyAsync(dst + off, fld(c, f), n, hipMemcpyDeviceToDevice, c->stream));
        else HIPCHK(c, hipMemcpyAsync(fld(c, f), dst + off, n, hipMemcpyDeviceToDevice, c->stream));
        off += n;
    }
    for (int q = 0; q < 12; ++q) {
        const size_t n = c->psi_elems[q] * c->esz;
        if (!n) continue;
        if (save) HIPCHK(c, hipMemcpyAsync(dst + off, psv(c, q), n, hipMemcpyDeviceToDevice, c->stream));
        else HIPCHK(c, hipMemcpyAsync(psv(c, q), dst + off, n, hipMemcpyDeviceToDevice, c->stream));
        off += n;
    }
    return FDTD_OK;
}

// The device buffers of one adjoint call: hipMalloc'd and freed on return,
// or (arena set) carved in 256-byte-aligned pieces from the static scratch,
// with no allocation at all.
struct DevBuf {
    std::vector<void *> ptrs;
    uint64_t bytes = 0;
    unsigned char *arena = nullptr;
    size_t arena_cap = 0;
    hipStream_t stream = nullptr;   // the arena's work is queued here
    static size_t rounded(size_t b) { return (std::max<size_t>(b, 256) + 255) / 256 * 256; }
    ~DevBuf()
    {
        // an early return may leave work queued on these buffers
        if (!ptrs.empty()) (void)hipDeviceSynchronize();
        for (void *p : ptrs) (void)hipFree(p);
        if (arena && bytes && stream) (void)hipStreamSynchronize(stream);
    }
    void *get(size_t bytes_)
    {
        const size_t b = rounded(bytes_);
        if (arena) {
            if (bytes + b > arena_cap) return nullptr;
            void *p = arena + bytes;
            bytes += b;
            return p;
        }
        void *p = nullptr;
        if (hipMalloc(&p, b) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptrs.push_back(p);
        bytes += b;
        return p;
    }
};

}  // namespace

// arena: nullptr allocates the call's buffers with hipMalloc; otherwise they
// are carved from arena (the static scratch, arena_cap bytes), which allows
// one segment only (no checkpoint pool, no second stream).
int run_adjoint(fdtd_ctx *c, const fdtd_objective *obj, int n_steps, int K, double *J_out, double *grad_out,
                size_t n_grad, unsigned char *arena = nullptr, size_t arena_cap = 0)
{
    if (!c->finalized) return fail(c, FDTD_E_STATE, "fdtd_finalize has not been called");
    const SceneIn &in = c->in;
    if (in.design_eps.empty() || !c->dcoef) return fail(c, FDTD_E_STATE, "no design region");
    if (in.storage == FDTD_STORE_BQ16)
        return fail(c, FDTD_E_UNSUPPORTED, "the adjoint runs with FP32, FP16 or BF16 storage, not block-scaled int16");
    if (n_steps < 1 || n_steps > in.steps_cap) return fail(c, FDTD_E_RANGE, "n_steps outside [1, %d]", in.steps_cap);
    // the adjoint sweeps address their FP32 arrays (two guard planes
    // included) and the class map with 32-bit byte offsets
    if ((size_t)c->field_elems * 4 >= ((size_t)1 << 32))
        return fail(c, FDTD_E_UNSUPPORTED, "grid too large for the adjoint: each FP32 adjoint array (with its two "
                    "guard planes) must be under 4 GiB (%zu bytes)", (size_t)c->field_elems * 4);
    const int *d = in.design;
    const int dnx = d[3] - d[0], dny = d[4] - d[1], dnz = d[5] - d[2];
    if (n_grad < (size_t)dnx * dny * dnz) return fail(c, FDTD_E_INVALID, "gradient buffer too small");
    const int *L = in.cpml.thick;
    const int nx = in.nx, ny = in.ny, nz = in.nz;
    // the dilated design box must be updated, CPML-free E edges
    if (d[0] < std::max(1, L[0]) || d[3] >= nx - std::max(1, L[1]) || d[1] < std::max(1, L[2]) ||
        d[4] >= ny - std::max(1, L[3]) || d[2] < std::max(1, L[4]) || d[5] >= nz - std::max(1, L[5]))
        return fail(c, FDTD_E_RANGE, "design region (dilated by one edge) must lie outside the CPML");
    for (const Source &s : in.sources)
        for (int k = s.b[2]; k < s.b[5]; ++k)
            for (int j = s.b[1]; j <= s.b[4]; ++j)
                for (int i = s.b[0]; i <= s.b[3]; ++i)
                    if (i >= d[0] && i <= d[3] && j >= d[1] && j <= d[4] && k >= d[2] && k < d[5])
                        return fail(c, FDTD_E_INVALID, "a source edge touches the design region");
    const int np = (int)in.probes.size(), nf = (int)in.freqs.size();
    if (obj->n_terms < 1) return fail(c, FDTD_E_INVALID, "empty objective");
    for (int t = 0; t < obj->n_terms; ++t) {
        const int p = obj->probe[t], f = obj->freq[t];
        if (p < 0 || p >= np || f < 0 || f >= nf) return fail(c, FDTD_E_RANGE, "objective term %d out of range", t);
        const Probe &pr = in.probes[p];
        if (pr.comp > 2) return fail(c, FDTD_E_INVALID, "objective probes must sample E");
        if (pr.i < std::max(1, L[0]) || pr.i >= nx - std::max(1, L[1]) || pr.j < std::max(1, L[2]) ||
            pr.j >= ny - std::max(1, L[3]) || pr.k < std::max(1, L[4]) || pr.k >= nz - std::max(1, L[5]))
            return fail(c, FDTD_E_INVALID, "objective probe %d lies in the CPML or on a face", p);
    }
    const int n = n_steps;
    const int bx = dnx + 1, by = dny + 1, bz = dnz + 1;
    const i64 nb = (i64)bx * by * bz;
    // adjoint fields (6) and, on narrow grids, every cell's transposed
    // quantities (12 arrays), guard planes included
    const int nq = adj_stores_quantities(nx) ? 12 : 0;
    const size_t adj_bytes = (6 + nq) * (size_t)c->field_elems * 4 + [&] {
        size_t t = 0;
        for (int q = 0; q < 12; ++q) t += c->psi_elems[q] * 4;
        return t;
    }();
    size_t freeb = 0, totb = 0;
    if (K < 0) {
        // automatic: the design history of the whole run when it fits in half
        // the free device memory (no recompute), else checkpoints every
        // ceil(sqrt(n)) steps
        HIPCHK(c, hipMemGetInfo(&freeb, &totb));
        const size_t all = (size_t)(n + 1) * 3 * nb * sizeof(float) + adj_bytes + (size_t)std::max(np, 1) * n * 4;
        K = all < freeb / 2 ? 0 : (int)std::ceil(std::sqrt((double)n));
    }
    int Kseg = K > 0 ? std::min(K, n) : n;
    const int nseg = (n + Kseg - 1) / Kseg;
    const size_t hist_bytes = (size_t)(Kseg + 1) * 3 * nb * sizeof(float);
    const size_t ck_bytes = state_bytes(c);
    // segmented runs double-buffer the history: segment s-1 is recomputed on
    // a second stream while the reverse sweep of segment s runs
    const int nhist = nseg > 1 ? 2 : 1;
    if (arena) {
        if (nseg > 1)
            return fail(c, FDTD_E_INVALID, "the no-allocation adjoint stores the design history of the whole run: "
                        "checkpoint_every must be 0 or >= n_steps");
        // exactly the pieces carved below, each rounded to 256 bytes
        size_t carve = (6 + nq) * DevBuf::rounded((size_t)c->field_elems * 4) + DevBuf::rounded(hist_bytes) +
                       DevBuf::rounded(3 * nb * sizeof(double)) +
                       DevBuf::rounded((size_t)std::max(np, 1) * n * sizeof(float));
        for (int q = 0; q < 12; ++q) carve += DevBuf::rounded(std::max<size_t>(c->psi_elems[q], 1) * 4);
        if (carve > arena_cap)
            return fail(c, FDTD_E_NOMEM, "the no-allocation adjoint needs %zu bytes of device scratch for this "
                        "run; the static scratch holds %zu (FDTD_NOALLOC_SCRATCH_BYTES)", carve, arena_cap);
    } else {
        HIPCHK(c, hipMemGetInfo(&freeb, &totb));
        const size_t need = nhist * hist_bytes + (size_t)(nseg - 1) * ck_bytes + adj_bytes +
                            (size_t)np * n * 4 + 3 * nb * 8;
        if ((double)need > 0.9 * (double)freeb)
            return fail(c, FDTD_E_NOMEM, "adjoint needs %.2f GB of device memory, %.2f GB free; use a larger "
                        "checkpoint interval", need / 1e9, freeb / 1e9);
    }

    DevBuf mem;
    mem.arena = arena;
    mem.arena_cap = arena_cap;
    mem.stream = c->stream;
    AArgs b{};
    for (int q = 0; q < 3; ++q) {
        char *e = (char *)mem.get((size_t)c->field_elems * 4);
        char *h = (char *)mem.get((size_t)c->field_elems * 4);
        if (!e || !h) return fail(c, FDTD_E_NOMEM, "adjoint field allocation failed");
        HIPCHK(c, hipMemsetAsync(e, 0, (size_t)c->field_elems * 4, c->stream));
        HIPCHK(c, hipMemsetAsync(h, 0, (size_t)c->field_elems * 4, c->stream));
        b.eb[q] = (float *)(e + (size_t)c->ka.plane * 4);
        b.hb[q] = (float *)(h + (size_t)c->ka.plane * 4);
        b.ebn[q] = b.eb[q];   // in place unless the fused interior sweep is set up below
        b.hbn[q] = b.hb[q];
    }
    for (int q = 0; q < nq; ++q) {
        char *t = (char *)mem.get((size_t)c->field_elems * 4);
        if (!t) return fail(c, FDTD_E_NOMEM, "adjoint transposed-quantity allocation failed");
        HIPCHK(c, hipMemsetAsync(t, 0, (size_t)c->field_elems * 4, c->stream));
        (q < 6 ? b.hq[q] : b.eq[q - 6]) = (float *)(t + (size_t)c->ka.plane * 4);
    }
    for (int q = 0; q < 12; ++q) {
        b.pb[q] = (float *)mem.get(std::max<size_t>(c->psi_elems[q], 1) * 4);
        if (!b.pb[q]) return fail(c, FDTD_E_NOMEM, "adjoint psi allocation failed");
        HIPCHK(c, hipMemsetAsync(b.pb[q], 0, std::max<size_t>(c->psi_elems[q], 1) * 4, c->stream));
    }
    float *hbuf[2] = {(float *)mem.get(hist_bytes), nhist > 1 ? (float *)mem.get(hist_bytes) : nullptr};
    float *hist = hbuf[0];
    double *gedge = (double *)mem.get(3 * nb * sizeof(double));
    float *gdev = (float *)mem.get((size_t)std::max(np, 1) * n * sizeof(float));
    char *pool = nseg > 1 ? (char *)mem.get((size_t)(nseg - 1) * ck_bytes) : nullptr;
    if (!hist || (nhist > 1 && !hbuf[1]) || !gedge || !gdev || (nseg > 1 && !pool))
        return fail(c, FDTD_E_NOMEM, "adjoint buffers");
    for (int q = 0; q < nhist; ++q) HIPCHK(c, hipMemsetAsync(hbuf[q], 0, hist_bytes, c->stream));
    HIPCHK(c, hipMemsetAsync(gedge, 0, 3 * nb * sizeof(double), c->stream));
    b.gsteps = n;
    b.g = gdev;
    b.hist = hist;
    b.dcoef = c->dcoef;
    b.gedge = gedge;
    b.sb[0] = d[0]; b.sb[1] = d[1]; b.sb[2] = d[2]; b.sb[3] = bx; b.sb[4] = by; b.sb[5] = bz;
    b.src_j0 = c->built.src_j0;
    b.src_j1 = c->built.src_j1;

    int rc = zero_state(c);
    if (rc) return rc;
    KArgs fa = c->ka;               // forward arguments
    fa.sb[0] = d[0]; fa.sb[1] = d[1]; fa.sb[2] = d[2]; fa.sb[3] = bx; fa.sb[4] = by; fa.sb[5] = bz;
    KArgs aa = c->ka;               // adjoint arguments: SI tables
    if (const char *ov = std::getenv("FDTD_ADJ_CHUNKS")) {
        const int nch = std::max(1, std::min(std::atoi(ov), nz));
        aa.kchunk = (nz + nch - 1) / nch;
    } else {
        // z-blocks of the adjoint sweeps: those of the forward E sweep, halved
        // until a launch has 2048 workgroups (8 per CU: the reverse sweeps of a
        // small grid are latency-bound).  Measured on 128x128x64, 1000 steps,
        // stored history: 4-plane blocks 83.0 ms, 2-plane 78.3, 1-plane 78.2.
        const int tiles = ((nx + tile_cols() - 1) / tile_cols()) * ((ny + tile_rows() - 1) / tile_rows());
        while (aa.kchunk > 1 && (long)tiles * ((nz + aa.kchunk - 1) / aa.kchunk) < 2048) aa.kchunk /= 2;
    }
    aa.etab = c->etab_si;
    aa.htab = c->htab_si;
    aa.srcs = c->srcs_si;
    // interior / general split of the adjoint sweeps (FDTD_INTERIOR=0: general only)
    {
        const char *ev = std::getenv("FDTD_INTERIOR");
        aa.ib[0] = (ev && std::strcmp(ev, "0") == 0) ? -1 : 0;
    }
    const int mode = c->built.mode;
    const int euni = c->built.e_uniform ? 1 : 0, huni = c->built.h_uniform ? 1 : 0;
    // fused interior sweep (LDS gather path, one H class, FDTD_ADJ_FUSE=0
    // disables): its tile columns and the others, and the second dJ/dE,
    // dJ/dH buffers it writes (the pair is swapped after every step).  Not
    // with the no-allocation scratch, nor when the pair would take more than
    // half the free memory.
    bool pp = false;
    {
        const char *fv = std::getenv("FDTD_ADJ_FUSE");
        const bool want = !(fv && std::strcmp(fv, "0") == 0);
        std::vector<int> quiet, noisy;
        if (want && !arena && huni && !adj_stores_quantities(nx) && aa.ib[0] >= 0)
            adj_fuse_tiles(aa, c->built.e_first, c->built.h_first, b.sb, quiet, noisy);
        const size_t pair_bytes = 6 * (size_t)c->field_elems * 4 + (quiet.size() + noisy.size() + 2) * 4;
        if (!quiet.empty()) HIPCHK(c, hipMemGetInfo(&freeb, &totb));
        if (!quiet.empty() && pair_bytes < freeb / 2) {
            for (int q = 0; q < 6; ++q) {
                char *f = (char *)mem.get((size_t)c->field_elems * 4);
                if (!f) return fail(c, FDTD_E_NOMEM, "adjoint field allocation failed");
                HIPCHK(c, hipMemsetAsync(f, 0, (size_t)c->field_elems * 4, c->stream));
                (q < 3 ? b.ebn[q] : b.hbn[q - 3]) = (float *)(f + (size_t)c->ka.plane * 4);
            }
            int *fl = (int *)mem.get(quiet.size() * 4);
            int *tl = (int *)mem.get(std::max<size_t>(noisy.size(), 1) * 4);
            if (!fl || !tl) return fail(c, FDTD_E_NOMEM, "adjoint tile lists");
            HIPCHK(c, hipMemcpyAsync(fl, quiet.data(), quiet.size() * 4, hipMemcpyHostToDevice, c->stream));
            if (!noisy.empty())
                HIPCHK(c, hipMemcpyAsync(tl, noisy.data(), noisy.size() * 4, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));   // the host vectors go out of scope
            b.fl = fl;
            b.nfl = (int)quiet.size();
            b.tl = tl;
            b.ntl = (int)noisy.size();
            pp = true;
        }
    }
    // after a step the buffers just written hold the current state
    auto swap_e = [&]() { for (int q = 0; q < 3; ++q) std::swap(b.eb[q], b.ebn[q]); };
    auto swap_h = [&]() { for (int q = 0; q < 3; ++q) std::swap(b.hb[q], b.hbn[q]); };

    hipEvent_t t0 = c->adj_ev[0], t1 = c->adj_ev[1], t2 = c->adj_ev[2];
    // second stream (default priority) for the reverse sweeps of a segmented
    // run, and the events that order it against the recompute stream:
    // ready[q] = history buffer q filled, freed[q] = its reverse sweep done
    StreamGuard rsg;
    EventSet sync;
    hipStream_t rs = c->stream;
    if (nseg > 1) {
        HIPCHK(c, hipStreamCreateWithFlags(&rsg.s, hipStreamNonBlocking));
        rs = rsg.s;
        sync.v.assign(4, nullptr);
        for (auto &e : sync.v) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // the fused interior pass runs on a stream of its own (default priority)
    // beside the E sweep's other launches, which touch other cells; the H
    // sweep waits for it (FDTD_ADJ_STREAMS=1: all on one stream).  fev[0]:
    // the previous step's H sweep done, fev[1]: the fused pass done.
    StreamGuard fsg;
    EventSet fev;
    const char *sv = std::getenv("FDTD_ADJ_STREAMS");
    const bool fz_async = pp && !(sv && std::strcmp(sv, "1") == 0);
    if (fz_async) {
        HIPCHK(c, hipStreamCreateWithFlags(&fsg.s, hipStreamNonBlocking));
        fev.v.assign(2, nullptr);
        for (auto &e : fev.v) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    HIPCHK(c, hipEventRecord(t0, c->stream));
    // ---- forward pass: DFT at the probes; checkpoints at segment starts, or
    // the whole design history when there is a single segment
    fa.rec = 1;
    fa.snap = nseg == 1 ? hist : nullptr;
    fa.snap_base = 0;
    for (int m = 0; m < n; ++m) {
        if (nseg > 1 && m % Kseg == 0 && m > 0)
            if ((rc = copy_state(c, pool + (size_t)(m / Kseg - 1) * ck_bytes, true))) return rc;
        HIPCHK(c, do_step(c, fa, m));
    }
    HIPCHK(c, hipEventRecord(t1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->steps_done = n;
    // ---- objective J = sum w |X|^2 and dJ/ds[m] = sum w 2 Re(conj(X) phasor[m])
    std::vector<double> X((size_t)np * nf * 2);
    if (np && nf) HIPCHK(c, hipMemcpy(X.data(), c->ka.dft, X.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<float> g((size_t)np * n, 0.0f);
    std::vector<double> gd((size_t)np * n, 0.0);
    double J = 0.0;
    for (int t = 0; t < obj->n_terms; ++t) {
        const int p = obj->probe[t], f = obj->freq[t];
        const double w = obj->weight ? obj->weight[t] : 1.0;
        const double xr = X[((size_t)p * nf + f) * 2], xi = X[((size_t)p * nf + f) * 2 + 1];
        J += w * (xr * xr + xi * xi);
        for (int m = 0; m < n; ++m) {
            const double pr = c->built.phasor[((size_t)m * nf + f) * 2], pi = c->built.phasor[((size_t)m * nf + f) * 2 + 1];
            gd[(size_t)p * n + m] += w * 2.0 * (xr * pr + xi * pi);
        }
    }
    for (size_t q = 0; q < g.size(); ++q) g[q] = (float)gd[q];
    HIPCHK(c, hipMemcpyAsync(gdev, g.data(), g.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    // ---- reverse sweep, segment by segment.  Segmented: the forward state of
    // segment sgi is recomputed from its checkpoint on c->stream into history
    // buffer sgi % 2 while the reverse sweep of segment sgi + 1 runs on rs.
    // A step moves information one cell per axis (E reads H at i-1, H reads E
    // at i+1), so over a segment of at most Kseg steps the design history
    // depends only on the checkpoint within Kseg + 1 cells of the snapshot
    // box: the recompute sweeps that box alone, dilated by Kseg + 2 at a
    // segment's first step and a cell less at each later one, whenever the
    // step's box holds under 60 % of the grid's cells (the light-cone tiles
    // run the general variant; a larger box runs the whole grid).
    // FDTD_LIGHTCONE=0: the whole grid always; FDTD_LIGHTCONE=fixed: the
    // first box for every step.  Cells outside it keep stale values, and every
    // segment restarts from a full checkpoint.
    const char *lce = std::getenv("FDTD_LIGHTCONE");
    const bool lc_on = nseg > 1 && !(lce && std::strcmp(lce, "0") == 0);
    const bool lc_fixed = lce && std::strcmp(lce, "fixed") == 0;
    // the box dilated by R, or all zero (the whole grid)
    auto cone = [&](int R, int lc[6]) {
        const int lo[3] = {std::max(0, d[0] - R), std::max(0, d[1] - R), std::max(0, d[2] - R)};
        const int hi[3] = {std::min(nx, d[0] + bx + R), std::min(ny, d[1] + by + R), std::min(nz, d[2] + bz + R)};
        const double frac = (double)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]) / ((double)nx * ny * nz);
        const bool use = lc_on && frac < 0.6;
        for (int v = 0; v < 3; ++v) {
            lc[v] = use ? lo[v] : 0;
            lc[v + 3] = use ? hi[v] : 0;
        }
    };
    auto recompute = [&](int sgi) -> int {
        const int s0 = sgi * Kseg, s1 = std::min(n, s0 + Kseg), q = sgi % 2;
        if (sgi + 2 < nseg) HIPCHK(c, hipStreamWaitEvent(c->stream, sync.v[2 + q], 0));   // buffer q free
        if (sgi == 0) {
            for (int f = 0; f < 6; ++f)
                HIPCHK(c, hipMemsetAsync(fld(c, f), 0, (size_t)c->field_elems * c->esz, c->stream));
            for (int p = 0; p < 12; ++p)
                if (c->psi_elems[p]) HIPCHK(c, hipMemsetAsync(psv(c, p), 0, c->psi_elems[p] * c->esz, c->stream));
        } else {
            const int r = copy_state(c, pool + (size_t)(sgi - 1) * ck_bytes, false);
            if (r) return r;
        }
        KArgs ra = fa;
        ra.rec = 0;
        ra.snap = hbuf[q];
        ra.snap_base = s0;
        cone(Kseg + 2, ra.lc);
        bind_sets(c, ra);
        HIPCHK(c, launch_snap(in.storage, ra, hbuf[q], 0, c->stream));
        for (int m = s0; m < s1; ++m) {
            // the cone narrows a cell per step: step m feeds only the
            // snapshots of steps m .. s1-1, so the design box dilated by
            // s1 - m + 2 (Kseg + 2 at most) suffices
            if (!lc_fixed) cone(s1 - m + 2, ra.lc);
            HIPCHK(c, do_step(c, ra, m));
        }
        HIPCHK(c, hipEventRecord(sync.v[q], c->stream));   // buffer q ready
        return FDTD_OK;
    };
    if (nseg > 1 && (rc = recompute(nseg - 1))) return rc;
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
        const int s0 = sgi * Kseg, s1 = std::min(n, s0 + Kseg), q = sgi % 2;
        if (nseg > 1) HIPCHK(c, hipStreamWaitEvent(rs, sync.v[q], 0));
        b.hist = hbuf[q];
        b.hist_base = s0;
        for (int m = s1 - 1; m >= s0; --m) {
            const int fst = m < s1 - 1 ? m + 1 : -1;
            if (pp) {
                const hipStream_t fs = fz_async ? fsg.s : rs;
                if (fz_async) {
                    HIPCHK(c, hipEventRecord(fev.v[0], rs));
                    HIPCHK(c, hipStreamWaitEvent(fs, fev.v[0], 0));
                }
                HIPCHK(c, launch_adj(2, mode, euni, huni, aa, b, m, fst, 1, fs));
                if (fz_async) HIPCHK(c, hipEventRecord(fev.v[1], fs));
                HIPCHK(c, launch_adj(0, mode, euni, huni, aa, b, m, fst, 1, rs));
                if (fz_async) HIPCHK(c, hipStreamWaitEvent(rs, fev.v[1], 0));
                HIPCHK(c, launch_adj(1, mode, euni, huni, aa, b, m, -1, 1, rs));
                swap_e();
                swap_h();
            } else {
                HIPCHK(c, launch_adj(0, mode, euni, huni, aa, b, m, fst, 1, rs));
                HIPCHK(c, launch_adj(1, mode, euni, huni, aa, b, m, -1, 1, rs));
            }
        }
        HIPCHK(c, launch_adj(0, mode, euni, huni, aa, b, s0, s0, 0, rs));
        if (pp) swap_e();
        if (nseg > 1) {
            HIPCHK(c, hipEventRecord(sync.v[2 + q], rs));   // buffer q free
            if (sgi > 0 && (rc = recompute(sgi - 1))) return rc;
        }
    }
    if (nseg > 1) {
        HIPCHK(c, hipEventRecord(sync.v[2], rs));
        HIPCHK(c, hipStreamWaitEvent(c->stream, sync.v[2], 0));
    }
    HIPCHK(c, hipEventRecord(t2, c->stream));
    HIPCHK(c, hipEventSynchronize(t2));
    float ms1 = 0.f, ms2 = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms1, t0, t1));
    HIPCHK(c, hipEventElapsedTime(&ms2, t0, t2));
    c->adj_fwd_ms = ms1;
    c->adj_total_ms = ms2;
    c->adj_segments = nseg;
    c->adj_bytes = c->dev_bytes + mem.bytes;
    // ---- chain rule through the 4-cell averages (fdtdrav.h:1135-1177)
    std::vector<double> ge((size_t)3 * nb);
    HIPCHK(c, hipMemcpy(ge.data(), gedge, ge.size() * sizeof(double), hipMemcpyDeviceToHost));
    auto q = [&](int a_, int b_, int c_) { return (size_t)(a_ + (i64)bx * (b_ + (i64)by * c_)); };
    for (int k = 0; k < dnz; ++k)
        for (int j = 0; j < dny; ++j)
            for (int i = 0; i < dnx; ++i) {
                double acc = 0.0;
                for (int u = 0; u < 2; ++u)
                    for (int v = 0; v < 2; ++v) {
                        acc += ge[q(i, j + u, k + v)];
                        acc += ge[(size_t)nb + q(i + u, j, k + v)];
                        acc += ge[(size_t)2 * nb + q(i + u, j + v, k)];
                    }
                grad_out[i + (size_t)dnx * (j + (size_t)dny * k)] = 0.25 * acc;
            }
    *J_out = J;
    // leave the context as after an uninterrupted forward run of n steps
    // when there was a single segment; otherwise rewind it
    if (nseg > 1) return zero_state(c);
    return FDTD_OK;
}

// ============================================================================
extern "C" {

const char *fdtd_last_error(const fdtd_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err; }

int fdtd_create(const fdtd_grid *g, int storage, fdtd_ctx **out)
{
    if (!out) return create_fail(FDTD_E_INVALID, "null output pointer");
    *out = nullptr;
    if (!g) return create_fail(FDTD_E_INVALID, "null grid");
    if (g->nx < 4 || g->ny < 4 || g->nz < 4)
        return create_fail(FDTD_E_INVALID, "the grid needs at least 4 cells along each axis (got %d x %d x %d)", g->nx,
                           g->ny, g->nz);
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0))
        return create_fail(FDTD_E_INVALID, "cell sizes must be positive");
    if (storage < FDTD_STORE_F32 || storage > FDTD_STORE_BQ16)
        return create_fail(FDTD_E_INVALID, "unknown storage form %d", storage);
    double eps0 = g->eps0 > 0 ? g->eps0 : 8.854187817e-12;
    double mu0 = g->mu0 > 0 ? g->mu0 : 4.0 * M_PI * 1e-7;
    double dt = g->dt;
    const double cfl = 1.0 / ((1.0 / std::sqrt(mu0 * eps0)) *
                              std::sqrt(1.0 / (g->dx * g->dx) + 1.0 / (g->dy * g->dy) + 1.0 / (g->dz * g->dz)));
    if (!(dt > 0)) {
        if (!(g->courant > 0) || g->courant > 1.0)
            return create_fail(FDTD_E_INVALID, "courant must lie in (0, 1] when dt is not given");
        dt = g->courant * cfl;
    } else if (dt > cfl * (1.0 + 4.0 * FLT_EPSILON)) {
        // above the 3-D Courant limit: unstable.  A dt computed in float at
        // Courant factor 1 (source.cu:45-46 built as C++, where sqrt of a
        // float is a float) lands a few float ulps either side of the double
        // limit, so the test allows float resolution.
        return create_fail(FDTD_E_INVALID, "dt %.9g s is above the 3-D Courant limit %.9g s", dt, cfl);
    }
    fdtd_ctx *c = new (std::nothrow) fdtd_ctx();
    if (!c) return create_fail(FDTD_E_NOMEM, "out of host memory");
    SceneIn &in = c->in;
    in.nx = g->nx; in.ny = g->ny; in.nz = g->nz;
    in.px = (g->nx + 63) / 64 * 64;
    in.dx = g->dx; in.dy = g->dy; in.dz = g->dz;
    in.dt = dt; in.eps0 = eps0; in.mu0 = mu0;
    in.storage = storage;
    in.cpml = CpmlParams{{0, 0, 0, 0, 0, 0}, 3.0, {0, 0, 0}, 1.5, 5.0, 0.05, 0.0};
    for (int q = 0; q < 6; ++q) in.design[q] = 0;
    in.steps_cap = 0;
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return create_fail(FDTD_E_DEVICE, "no HIP device");
    }
    for (hipEvent_t &e : c->adj_ev)
        if (hipEventCreate(&e) != hipSuccess) {
            (void)hipGetLastError();
            fdtd_destroy(c);
            return create_fail(FDTD_E_DEVICE, "hipEventCreate failed");
        }
    // the no-allocation adjoint's scratch on this (current) device
    c->scratch = scratch_address(c->device);
    *out = c;
    return FDTD_OK;
}

int fdtd_destroy(fdtd_ctx *c)
{
    if (!c) return FDTD_E_INVALID;
    free_device(c);
    for (hipEvent_t e : c->adj_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FDTD_OK;
}

int fdtd_set_cpml(fdtd_ctx *c, const fdtd_cpml *p)
{
    if (!c || !p) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    const int n[3] = {c->in.nx, c->in.ny, c->in.nz};
    for (int a = 0; a < 3; ++a) {
        const int lo = p->thick[2 * a], hi = p->thick[2 * a + 1];
        if (lo < 0 || hi < 0 || lo + hi > n[a] - 2)
            return fail(c, FDTD_E_INVALID, "CPML widths %d+%d do not fit axis %d of %d cells", lo, hi, a, n[a]);
    }
    if (!(p->order >= 0) || !(p->kappa_max >= 1) || !(p->alpha_max >= 0) || !(p->alpha_min >= 0))
        return fail(c, FDTD_E_INVALID, "invalid CPML grading parameters");
    CpmlParams &q = c->in.cpml;
    for (int a = 0; a < 6; ++a) q.thick[a] = p->thick[a];
    q.order = p->order;
    for (int a = 0; a < 3; ++a) q.sigma_max[a] = p->sigma_max[a];
    q.sigma_factor = p->sigma_factor > 0 ? p->sigma_factor : 1.5;
    q.kappa_max = p->kappa_max;
    q.alpha_max = p->alpha_max;
    q.alpha_min = p->alpha_min;
    c->cpml_set = true;
    return FDTD_OK;
}

int fdtd_add_box(fdtd_ctx *c, int is, int js, int ks, int ie, int je, int ke, const float eps_r[3],
                 const float sigma_e[3], const float mu_r[3], const float sigma_m[3])
{
    if (!c || !eps_r || !sigma_e || !mu_r || !sigma_m) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    for (int q = 0; q < 3; ++q) {
        if (!(eps_r[q] > 0) || !(mu_r[q] > 0) || !(sigma_e[q] >= 0) || !(sigma_m[q] >= 0) ||
            !std::isfinite(eps_r[q]) || !std::isfinite(mu_r[q]) || !std::isfinite(sigma_e[q]) ||
            !std::isfinite(sigma_m[q]))
            return fail(c, FDTD_E_INVALID, "invalid material values");
    }
    Box b;
    b.b[0] = is; b.b[1] = js; b.b[2] = ks; b.b[3] = ie; b.b[4] = je; b.b[5] = ke;
    for (int q = 0; q < 3; ++q) {
        b.mat[q] = eps_r[q];
        b.mat[3 + q] = sigma_e[q];
        b.mat[6 + q] = mu_r[q];
        b.mat[9 + q] = sigma_m[q];
    }
    c->in.boxes.push_back(b);
    return FDTD_OK;
}

int fdtd_add_sector(fdtd_ctx *c, int i0, int j0, int i1, int j1, int i2, int j2, int kz, float radius, int max_x,
                    int max_y, float sigma_x, float sigma_y)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (!(radius > 0) || !std::isfinite(radius)) return fail(c, FDTD_E_INVALID, "sector radius must be > 0");
    if (!(sigma_x >= 0) || !(sigma_y >= 0) || !std::isfinite(sigma_x) || !std::isfinite(sigma_y))
        return fail(c, FDTD_E_INVALID, "invalid sector sigma");
    if (kz < 1 || kz > c->in.nz - 2 || max_x < 0 || max_x >= c->in.nx || max_y < 0 || max_y >= c->in.ny)
        return fail(c, FDTD_E_RANGE, "sector plane %d or extent (%d, %d) outside the grid", kz, max_x, max_y);
    const int lim = 1 << 24;   // node coordinates stay exact in float
    for (int v : {i0, j0, i1, j1, i2, j2})
        if (v <= -lim || v >= lim) return fail(c, FDTD_E_RANGE, "sector point index out of range");
    c->in.sectors.push_back(Sector{{i0, j0, i1, j1, i2, j2}, kz, max_x, max_y, radius, {sigma_x, sigma_y}});
    return FDTD_OK;
}

int fdtd_add_plate(fdtd_ctx *c, int axis, int is, int js, int ks, int ie, int je, int ke, const float sigma[3])
{
    if (!c || !sigma) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (axis < 1 || axis > 3) return fail(c, FDTD_E_INVALID, "plate axis must be 1, 2 or 3");
    for (int q = 0; q < 3; ++q)
        if (!(sigma[q] >= 0) || !std::isfinite(sigma[q])) return fail(c, FDTD_E_INVALID, "invalid plate sigma");
    if (is < 0 || js < 0 || ks < 0 || ie >= c->in.nx || je >= c->in.ny || ke >= c->in.nz || is > ie || js > je ||
        ks > ke)
        return fail(c, FDTD_E_RANGE, "plate (%d,%d,%d)-(%d,%d,%d) outside the grid", is, js, ks, ie, je, ke);
    Plate p{axis, {is, js, ks, ie, je, ke}, {sigma[0], sigma[1], sigma[2]}};
    c->in.plates.push_back(p);
    return FDTD_OK;
}

int fdtd_add_resistor(fdtd_ctx *c, int dir, int is, int js, int ks, int ie, int je, int ke, double R)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (dir != 3) return fail(c, FDTD_E_UNSUPPORTED, "only z-directed resistors (dir = 3) are built");
    if (!(R > 0) || !std::isfinite(R)) return fail(c, FDTD_E_INVALID, "resistance must be > 0");
    const SceneIn &in = c->in;
    if (is < 1 || js < 1 || ks < 0 || ie > in.nx - 2 || je > in.ny - 2 || ke > in.nz - 1 || is > ie || js > je ||
        ks >= ke)
        return fail(c, FDTD_E_RANGE, "resistor edges (%d,%d,%d)-(%d,%d,%d) outside the updated Ez edges", is, js, ks,
                    ie, je, ke);
    Resistor r{{is, js, ks, ie, je, ke}, R};
    c->in.resistors.push_back(r);
    return FDTD_OK;
}

int fdtd_add_lumped_source(fdtd_ctx *c, int dir, int is, int js, int ks, int ie, int je, int ke, double R,
                           const fdtd_waveform *w)
{
    if (!c || !w) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (dir != 3 && dir != 1) return fail(c, FDTD_E_UNSUPPORTED, "lumped sources are x- (dir = 1) or z-directed (dir = 3)");
    if (!(R > 0) || !std::isfinite(R)) return fail(c, FDTD_E_INVALID, "source resistance must be > 0");
    if (w->type < FDTD_WAVE_GAUSSIAN || w->type > FDTD_WAVE_SINE)
        return fail(c, FDTD_E_INVALID, "unknown waveform type %d", w->type);
    if ((w->type != FDTD_WAVE_SINE && !(w->tau > 0)) || !std::isfinite(w->amp))
        return fail(c, FDTD_E_INVALID, "invalid waveform parameters");
    if (dir == 1) {
        // defineSourceX's loaded edges: is<=i<ie, js<=j<=ie, ks<=k<=ke, all updated Ex
        // edges (i<nx-1, 0<j<ny-1, 0<k<nz-1); calc_e's Ez injection still runs on the
        // box's Ez edges is<=i<=ie, js<=j<=je, ks<=k<ke (none when ks == ke)
        if (is >= ie || js > ie || js > je || ks > ke) return fail(c, FDTD_E_INVALID, "empty x-source box (needs ie > is)");
        if (is < 1 || ie > c->in.nx - 2 || js < 1 || std::max(ie, je) > c->in.ny - 2 || ks < 1 || ke > c->in.nz - 2)
            return fail(c, FDTD_E_RANGE, "x-source box outside the updated Ex / Ez region");
    } else {
        if (is > ie || js > je || ks >= ke) return fail(c, FDTD_E_INVALID, "empty source box");
        // every edge must be an updated Ez edge: 0<i<nx-1, 0<j<ny-1, k<nz-1
        if (is < 1 || ie > c->in.nx - 2 || js < 1 || je > c->in.ny - 2 || ks < 0 || ke > c->in.nz - 1)
            return fail(c, FDTD_E_RANGE, "source box outside the updated Ez region");
    }
    Source s;
    s.dir = dir;
    s.b[0] = is; s.b[1] = js; s.b[2] = ks; s.b[3] = ie; s.b[4] = je; s.b[5] = ke;
    s.R = R;
    s.wtype = w->type;
    s.amp = w->amp; s.t0 = w->t0; s.tau = w->tau; s.freq = w->freq;
    c->in.sources.push_back(s);
    return FDTD_OK;
}

int fdtd_add_probe(fdtd_ctx *c, int comp, int i, int j, int k, int *index)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (comp < FDTD_EX || comp > FDTD_HZ) return fail(c, FDTD_E_INVALID, "invalid component %d", comp);
    if (!box_in_grid(c, i, j, k)) return fail(c, FDTD_E_RANGE, "probe (%d,%d,%d) outside the grid", i, j, k);
    if (index) *index = (int)c->in.probes.size();
    c->in.probes.push_back(Probe{comp, i, j, k});
    return FDTD_OK;
}

int fdtd_set_dft(fdtd_ctx *c, int nfreq, const double *freqs)
{
    if (!c || nfreq < 0 || (nfreq > 0 && !freqs)) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    for (int f = 0; f < nfreq; ++f)
        if (!std::isfinite(freqs[f])) return fail(c, FDTD_E_INVALID, "non-finite DFT frequency");
    c->in.freqs.assign(freqs, freqs + nfreq);
    return FDTD_OK;
}

int fdtd_set_design_region(fdtd_ctx *c, int is, int js, int ks, int ie, int je, int ke)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (is < 0 || js < 0 || ks < 0 || ie > c->in.nx || je > c->in.ny || ke > c->in.nz || is >= ie || js >= je ||
        ks >= ke)
        return fail(c, FDTD_E_RANGE, "design region outside the grid or empty");
    int *d = c->in.design;
    d[0] = is; d[1] = js; d[2] = ks; d[3] = ie; d[4] = je; d[5] = ke;
    c->in.design_eps.assign((size_t)(ie - is) * (je - js) * (ke - ks), 1.0f);
    return FDTD_OK;
}

int fdtd_set_design_eps(fdtd_ctx *c, const float *eps, size_t n)
{
    if (!c || !eps) return FDTD_E_INVALID;
    if (c->in.design_eps.empty()) return fail(c, FDTD_E_STATE, "no design region");
    if (n != c->in.design_eps.size()) return fail(c, FDTD_E_INVALID, "design eps has %zu cells, expected %zu", n,
                                                  c->in.design_eps.size());
    for (size_t q = 0; q < n; ++q)
        if (!(eps[q] > 0) || !std::isfinite(eps[q])) return fail(c, FDTD_E_INVALID, "invalid design eps_r");
    std::memcpy(c->in.design_eps.data(), eps, n * sizeof(float));
    if (!c->finalized) return FDTD_OK;
    // rebuild every table on the new materials and restart from step 0
    free_device(c);
    c->finalized = false;
    const int rc = do_finalize(c);
    if (rc) free_device(c);
    return rc;
}

int fdtd_finalize(fdtd_ctx *c, int n_steps)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "already finalized");
    if (n_steps < 1) return fail(c, FDTD_E_INVALID, "n_steps must be >= 1");
    c->in.steps_cap = n_steps;
    int rc = do_finalize(c);
    if (rc) free_device(c);
    return rc;
}

int fdtd_run(fdtd_ctx *c, int nsteps)
{
    if (!c) return FDTD_E_INVALID;
    return run_steps(c, nsteps, false);
}

int fdtd_step(fdtd_ctx *c)
{
    if (!c) return FDTD_E_INVALID;
    return run_steps(c, 1, false);
}

int fdtd_profile(fdtd_ctx *c, int nsteps)
{
    if (!c) return FDTD_E_INVALID;
    return run_steps(c, nsteps, true);
}

int fdtd_reset(fdtd_ctx *c)
{
    if (!c) return FDTD_E_INVALID;
    if (!c->finalized) return fail(c, FDTD_E_STATE, "fdtd_finalize has not been called");
    return zero_state(c);
}

int fdtd_get_info(const fdtd_ctx *c, fdtd_info *o)
{
    if (!c || !o) return FDTD_E_INVALID;
    std::memset(o, 0, sizeof *o);
    o->nx = c->in.nx; o->ny = c->in.ny; o->nz = c->in.nz;
    o->steps_done = c->steps_done;
    o->steps_capacity = c->in.steps_cap;
    o->n_probes = (int)c->in.probes.size();
    o->n_freqs = (int)c->in.freqs.size();
    o->n_sources = (int)c->in.sources.size();
    o->n_classes = c->built.ncls;
    o->storage = c->in.storage;
    o->dt = c->in.dt;
    o->last_run_ms = c->last_run_ms;
    o->e_kernel_ms = c->e_ms;
    o->h_kernel_ms = c->h_ms;
    o->step_kernel_ms = c->step_ms;
    o->interior_split = c->box[0][5] > c->box[0][2] ? 1 : 0;
    o->device_bytes = c->dev_bytes;
    // algorithmic bytes per cell: E sweep reads E, H and writes E (9 values);
    // H sweep likewise; plus one class index per cell in a classed sweep and
    // psi read+write over the slabs (averaged over the grid).
    // block-scaled storage: 2 bytes per value plus BQ_STRIDE exponent bytes
    // per 8 cells of the three components; its psi is FP16
    const double s = c->in.storage == FDTD_STORE_BQ16 ? 2.0 + (double)BQ_STRIDE / (3 * BQ_BLOCK) : (double)c->esz;
    const double sp = (double)c->esz;
    const double fx = (double)(c->in.cpml.thick[0] + c->in.cpml.thick[1]) / c->in.nx;
    const double fy = (double)(c->in.cpml.thick[2] + c->in.cpml.thick[3]) / c->in.ny;
    const double fz = (double)(c->in.cpml.thick[4] + c->in.cpml.thick[5]) / c->in.nz;
    const double psi = 4.0 * sp * (fx + fy + fz);
    const double cls = c->built.mode == CLS_U16 ? 2.0 : c->built.mode == CLS_U32 ? 4.0 : 0.0;
    o->adjoint_forward_ms = c->adj_fwd_ms;
    o->adjoint_total_ms = c->adj_total_ms;
    o->adjoint_segments = c->adj_segments;
    o->adjoint_device_bytes = c->adj_bytes;
    o->bytes_per_cell_e = (uint64_t)std::llround(9 * s + (c->e_mode ? cls : 0.0) + psi);
    o->bytes_per_cell_h = (uint64_t)std::llround(9 * s + (c->h_mode ? cls : 0.0) + psi);
    return FDTD_OK;
}

int fdtd_get_probes(fdtd_ctx *c, float *out, size_t n)
{
    if (!c || (!out && n)) return FDTD_E_INVALID;
    if (!c->finalized) return fail(c, FDTD_E_STATE, "not finalized");
    const size_t np = c->in.probes.size(), ns = (size_t)c->steps_done;
    if (n < np * ns) return fail(c, FDTD_E_INVALID, "output holds %zu values, need %zu", n, np * ns);
    if (np && ns)
        HIPCHK(c, hipMemcpy2D(out, ns * sizeof(float), c->ka.probe_hist, (size_t)c->in.steps_cap * sizeof(float),
                              ns * sizeof(float), np, hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_set_port(fdtd_ctx *c, int vdir, int vis, int vjs, int vks, int vie, int vje, int vke, int idir, int iis,
                  int ijs, int iks, int iie, int ije, int ike)
{
    if (!c) return FDTD_E_INVALID;
    if (c->finalized) return fail(c, FDTD_E_STATE, "scene is finalized");
    if (c->has_port) return fail(c, FDTD_E_STATE, "the port is already set (one port per context)");
    PortDef d{vdir, {vis, vjs, vks, vie, vje, vke}, idir, {iis, ijs, iks, iie, ije, ike}};
    std::string err;
    const SceneIn &in = c->in;
    int rc = port_plan(d, in.nx, in.ny, in.nz, (float)in.dx, (float)in.dy, (float)in.dz, c->port, err);
    if (rc != FDTD_OK) return fail(c, rc, "%s", err.c_str());
    c->port_base = (int)c->in.probes.size();
    c->in.probes.insert(c->in.probes.end(), c->port.samples.begin(), c->port.samples.end());
    c->has_port = true;
    return FDTD_OK;
}

// V and averaged I of the steps run so far (host copies of the port's probe
// histories, reduced by port_series).
static int port_read(fdtd_ctx *c, std::vector<float> &v, std::vector<float> &cur)
{
    if (!c->finalized) return fail(c, FDTD_E_STATE, "not finalized");
    if (!c->has_port) return fail(c, FDTD_E_STATE, "no port was set");
    const size_t np = c->port.samples.size(), ns = (size_t)c->steps_done;
    v.assign(ns, 0.f);
    cur.assign(ns, 0.f);
    if (!ns) return FDTD_OK;
    std::vector<float> hist(np * ns);
    HIPCHK(c, hipMemcpy2D(hist.data(), ns * sizeof(float), c->ka.probe_hist + (size_t)c->port_base * c->in.steps_cap,
                          (size_t)c->in.steps_cap * sizeof(float), ns * sizeof(float), np, hipMemcpyDeviceToHost));
    port_series(c->port, hist.data(), ns, v.data(), cur.data());
    return FDTD_OK;
}

int fdtd_get_port(fdtd_ctx *c, float *v, float *i, size_t n)
{
    if (!c || ((!v || !i) && n)) return FDTD_E_INVALID;
    std::vector<float> vv, ii;
    if (int rc = port_read(c, vv, ii)) return rc;
    if (n < vv.size()) return fail(c, FDTD_E_INVALID, "output holds %zu values, need %zu", n, vv.size());
    std::copy(vv.begin(), vv.end(), v);
    std::copy(ii.begin(), ii.end(), i);
    return FDTD_OK;
}

int fdtd_port_spectrum(fdtd_ctx *c, double z0, int nfreq, const double *freqs, double *out, size_t n)
{
    if (!c || nfreq < 0 || (nfreq && (!freqs || !out))) return FDTD_E_INVALID;
    if (!(z0 > 0.0) || !std::isfinite(z0)) return fail(c, FDTD_E_INVALID, "z0 must be positive");
    if (n < 8 * (size_t)nfreq) return fail(c, FDTD_E_INVALID, "output holds %zu values, need %zu", n, 8 * (size_t)nfreq);
    std::vector<float> vv, ii;
    if (int rc = port_read(c, vv, ii)) return rc;
    port_spectrum(vv.data(), ii.data(), vv.size(), c->in.dt, z0, nfreq, freqs, out);
    return FDTD_OK;
}

int fdtd_save_port(fdtd_ctx *c, const char *vpath, const char *ipath)
{
    if (!c || !vpath || !ipath) return FDTD_E_INVALID;
    std::vector<float> vv, ii;
    if (int rc = port_read(c, vv, ii)) return rc;
    if (port_save(vv.data(), vv.size(), vpath) != FDTD_OK) return fail(c, FDTD_E_INVALID, "cannot write %s", vpath);
    if (port_save(ii.data(), ii.size(), ipath) != FDTD_OK) return fail(c, FDTD_E_INVALID, "cannot write %s", ipath);
    return FDTD_OK;
}

int fdtd_get_dft(fdtd_ctx *c, double *out, size_t n)
{
    if (!c || (!out && n)) return FDTD_E_INVALID;
    if (!c->finalized) return fail(c, FDTD_E_STATE, "not finalized");
    const size_t cnt = c->in.probes.size() * c->in.freqs.size() * 2;
    if (n < cnt) return fail(c, FDTD_E_INVALID, "output holds %zu values, need %zu", n, cnt);
    if (cnt) HIPCHK(c, hipMemcpy(out, c->ka.dft, cnt * sizeof(double), hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_get_field(fdtd_ctx *c, int comp, float *out, size_t n)
{
    if (!c || !out) return FDTD_E_INVALID;
    if (!c->finalized) return fail(c, FDTD_E_STATE, "not finalized");
    if (comp < 0 || comp > 5) return fail(c, FDTD_E_INVALID, "invalid component");
    const int nx = c->in.nx, ny = c->in.ny, nz = c->in.nz, px = c->in.px;
    const size_t cells = (size_t)nx * ny * nz;
    if (n < cells) return fail(c, FDTD_E_INVALID, "output too small");
    std::vector<unsigned char> raw((size_t)px * ny * nz * c->esz);
    const char *base = (const char *)c->fields[comp] + (size_t)c->ka.plane * c->esz;
    HIPCHK(c, hipMemcpy(raw.data(), base, raw.size(), hipMemcpyDeviceToHost));
    std::vector<std::uint8_t> qraw;   // the component's exponent array, k = 0 .. nz-1
    if (c->qarr[comp / 3]) {
        qraw.resize((size_t)px * ny * nz / BQ_BLOCK * BQ_STRIDE);
        HIPCHK(c, hipMemcpy(qraw.data(), c->qarr[comp / 3] + (size_t)c->ka.plane / BQ_BLOCK * BQ_STRIDE,
                            qraw.size(), hipMemcpyDeviceToHost));
    }
    const float un = c->built.probe_unscale[comp];
    for (size_t r = 0; r < (size_t)ny * nz; ++r)
        for (int i = 0; i < nx; ++i) {
            const size_t src = r * px + i;
            float v;
            if (c->esz == 4) {
                std::memcpy(&v, &raw[src * 4], 4);
            } else {
                uint16_t h;
                std::memcpy(&h, &raw[src * 2], 2);
                if (c->in.storage == FDTD_STORE_BF16) {
                    uint32_t u = (uint32_t)h << 16;
                    std::memcpy(&v, &u, 4);
                } else if (c->in.storage == FDTD_STORE_BQ16) {
                    v = std::ldexp((float)(int16_t)h, (int)qraw[src / BQ_BLOCK * BQ_STRIDE + comp % 3] - BQ_BIAS);
                } else {
                    _Float16 f;
                    std::memcpy(&f, &h, 2);
                    v = (float)f;
                }
            }
            out[r * nx + i] = v * un;
        }
    return FDTD_OK;
}

// ---- checkpoint -------------------------------------------------------------
int fdtd_checkpoint_size(const fdtd_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return FDTD_E_INVALID;
    if (!c->finalized) return FDTD_E_STATE;
    size_t p = 6 * (size_t)c->field_elems * c->esz + 2 * c->qarr_bytes;
    for (int q = 0; q < 12; ++q) p += c->psi_elems[q] * c->esz;
    p += c->in.probes.size() * (size_t)c->in.steps_cap * sizeof(float);
    p += c->in.probes.size() * c->in.freqs.size() * sizeof(double2);
    *bytes = sizeof(CkHeader) + p;
    return FDTD_OK;
}

int fdtd_checkpoint_save(fdtd_ctx *c, void *buf, size_t bytes)
{
    if (!c || !buf) return FDTD_E_INVALID;
    size_t need = 0;
    int rc = fdtd_checkpoint_size(c, &need);
    if (rc) return fail(c, rc, "not finalized");
    if (bytes < need) return fail(c, FDTD_E_INVALID, "checkpoint buffer holds %zu bytes, need %zu", bytes, need);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned char *p = (unsigned char *)buf + sizeof(CkHeader);
    auto get = [&](const void *dev, size_t n) -> int {
        if (n) HIPCHK(c, hipMemcpy(p, dev, n, hipMemcpyDeviceToHost));
        p += n;
        return FDTD_OK;
    };
    for (int f = 0; f < 6; ++f)
        if ((rc = get(fld(c, f), (size_t)c->field_elems * c->esz))) return rc;
    for (int q = 0; q < 2; ++q)
        if ((rc = get(c->qarr[q], c->qarr_bytes))) return rc;
    for (int q = 0; q < 12; ++q)
        if ((rc = get(psv(c, q), c->psi_elems[q] * c->esz))) return rc;
    if ((rc = get(c->ka.probe_hist, c->in.probes.size() * (size_t)c->in.steps_cap * sizeof(float)))) return rc;
    if ((rc = get(c->ka.dft, c->in.probes.size() * c->in.freqs.size() * sizeof(double2)))) return rc;
    CkHeader h{};
    std::memcpy(h.magic, "FDTDCKP1", 8);
    h.version = 1;
    h.nx = c->in.nx; h.ny = c->in.ny; h.nz = c->in.nz;
    h.storage = c->in.storage;
    h.step = c->steps_done;
    h.nprobe = (int)c->in.probes.size();
    h.nfreq = (int)c->in.freqs.size();
    h.steps_cap = c->in.steps_cap;
    h.payload = need - sizeof(CkHeader);
    h.checksum = fnv1a((const unsigned char *)buf + sizeof(CkHeader), h.payload);
    std::memcpy(buf, &h, sizeof h);
    return FDTD_OK;
}

int fdtd_checkpoint_restore(fdtd_ctx *c, const void *buf, size_t bytes)
{
    if (!c || !buf) return FDTD_E_INVALID;
    size_t need = 0;
    int rc = fdtd_checkpoint_size(c, &need);
    if (rc) return fail(c, rc, "not finalized");
    if (bytes < need) return fail(c, FDTD_E_CHECKSUM, "checkpoint truncated");
    CkHeader h;
    std::memcpy(&h, buf, sizeof h);
    if (std::memcmp(h.magic, "FDTDCKP1", 8) != 0 || h.version != 1 || h.nx != c->in.nx || h.ny != c->in.ny ||
        h.nz != c->in.nz || h.storage != c->in.storage || h.nprobe != (int)c->in.probes.size() ||
        h.nfreq != (int)c->in.freqs.size() || h.steps_cap != c->in.steps_cap || h.payload != need - sizeof h ||
        h.step < 0 || h.step > c->in.steps_cap)
        return fail(c, FDTD_E_CHECKSUM, "checkpoint does not match this context");
    if (fnv1a((const unsigned char *)buf + sizeof h, h.payload) != h.checksum)
        return fail(c, FDTD_E_CHECKSUM, "checkpoint checksum mismatch");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned char *p = (const unsigned char *)buf + sizeof h;
    auto put = [&](void *dev, size_t n) -> int {
        if (n) HIPCHK(c, hipMemcpy(dev, p, n, hipMemcpyHostToDevice));
        p += n;
        return FDTD_OK;
    };
    for (int f = 0; f < 6; ++f)
        if ((rc = put(fld(c, f), (size_t)c->field_elems * c->esz))) return rc;
    for (int q = 0; q < 2; ++q)
        if ((rc = put(c->qarr[q], c->qarr_bytes))) return rc;
    for (int q = 0; q < 12; ++q)
        if ((rc = put(psv(c, q), c->psi_elems[q] * c->esz))) return rc;
    if ((rc = put(c->ka.probe_hist, c->in.probes.size() * (size_t)c->in.steps_cap * sizeof(float)))) return rc;
    if ((rc = put(c->ka.dft, c->in.probes.size() * c->in.freqs.size() * sizeof(double2)))) return rc;
    c->steps_done = h.step;
    return FDTD_OK;
}

int fdtd_adjoint_gradient(fdtd_ctx *c, const fdtd_objective *obj, int n_steps, int checkpoint_every,
                          double *J_out, double *grad_out, size_t n_grad)
{
    if (!c || !obj || !J_out || !grad_out) return FDTD_E_INVALID;
    return run_adjoint(c, obj, n_steps, checkpoint_every, J_out, grad_out, n_grad);
}

int fdtd_adjoint_gradient_noalloc(fdtd_ctx *c, const fdtd_objective *obj, int n_steps, double *J_out,
                                  double *grad_out, size_t n_grad)
{
    if (!c || !obj || !J_out || !grad_out) return FDTD_E_INVALID;
    if (!c->scratch)
        return fail(c, FDTD_E_DEVICE, "the static scratch has no device address on device %d", c->device);
    return run_adjoint(c, obj, n_steps, 0, J_out, grad_out, n_grad, c->scratch, FDTD_NOALLOC_SCRATCH_BYTES);
}

}  // extern "C"
