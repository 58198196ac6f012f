// sg_solver.cpp -- the solver object behind the C ABI (include/sgfd.h).
//
// Owns every device buffer of one shot, one HIP stream, an optional hipGraph
// of one time step and the HIP events of timed runs.  The time step is
//   inject(it) -> velocity sweep -> stress sweep -> station gather
// as in propagate() of the reference (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
// ElasticWaveEquation3D.cpp:59-69; 3DGPUwithPML_1DArraySingleGPU_concise/
// ElasticWaveEquation.cu:194-206), all queued asynchronously on the solver's
// stream; the step index lives in device memory so that one captured step
// can be replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgfd.h"
#include "sg_error.hpp"
#include "sg_internal.hpp"
#include "sg_precision.hpp"
#include "sg_worklist.hpp"

namespace sg {
std::string &last_error()
{
    static thread_local std::string e;
    return e;
}
}  // namespace sg

using namespace sg;

namespace {

size_t storage_size(int prec) { return prec == SG_FP64 ? 8 : (prec == SG_FP32 ? 4 : 2); }
size_t compute_size(int prec) { return prec == SG_FP64 ? 8 : 4; }

// bytes of one field component per point: RMW of the full value and of each
// damped split, or -- when every contributing axis damps -- RMW of the splits
// plus a write of the full value (the reference rebuilds it from its splits).
double comp_bytes(int naxes, int ndamped, double sT, double sC)
{
    if (ndamped == naxes) return 2.0 * ndamped * sC + sT;
    return 2.0 * sT + 2.0 * ndamped * sC;
}

// float -> 16-bit storage, round to nearest even
uint16_t to_half_bits(float v)
{
    _Float16 h = (_Float16)v;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
float from_half_bits(uint16_t b)
{
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (float)h;
}
uint16_t to_bf16_bits(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float from_bf16_bits(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// power-of-two scale that maps max|x| into [2^13, 2^14) (FP16 max is 65504)
float pow2_scale(double maxabs)
{
    if (!(maxabs > 0) || !std::isfinite(maxabs)) return 1.0f;
    const int e = (int)std::floor(std::log2(16384.0 / maxabs));
    return std::ldexp(1.0f, std::max(-126, std::min(127, e)));
}

}  // namespace

struct sg_solver {
    sg_config cfg{};
    int prec = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // side streams: the PML work-lists of a sweep run beside the interior
    // list (forked from and joined back to `stream` with events)
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
    int nstreams = 3;
    Geom g{};
    Tile tile{4, 1, 4, 1, 0, 0};

    void *fields[NFIELD] = {};
    void *model[3] = {};
    void *splits = nullptr;  // 24 split arrays of g.nslots compute-type values
    void *tables = nullptr;  // dA[3], dB[3] back to back, compute type
    void *saved = nullptr;   // 3 storage values (TRANSIENT source)
    int *d_step = nullptr;
    double *amp_d = nullptr;
    void *amp_acc = nullptr;
    int namp = 0;
    long long src = -1;

    long long *st_idx = nullptr;
    int *st_comp = nullptr;
    void *st_out = nullptr;
    int nst = 0, ncomp = 0, interval = 1;
    long long st_cap = 0;

    // energy reductions: ENERGY_PART partials per updated z plane, then the
    // results (4 doubles), then the nine max-abs counters (unsigned)
    double *d_energy = nullptr;
    // stability guard (sg_set_stability): Courant limit checked by
    // sg_set_medium, the FP64 energy check every check_interval steps
    double courant = 0.0;        // of the current model (0: none set yet)
    double courant_limit = 1.0;  // 0: no check
    int check_interval = 0;
    double max_growth = 2.0;
    double last_energy = -1.0;   // total energy at the previous check after the source ended
    int src_last = -1;           // last wavelet sample above 1e-6 of its peak
    int round_mode = SG_ROUND_NEAREST;
    void *comp[NFIELD] = {};  // compensation arrays (16-bit storage, see sg_set_compensation)
    unsigned comp_mask = 0;
    long long steps_done = 0;
    // streams of the live batches that step this solver (sg_batch_create);
    // setters, getters and single steps wait for them as for `stream`
    std::vector<hipStream_t> batch_streams;
    std::vector<double> damp[3];
    bool damp_local = false;  // damp[] holds g of d = g * sqrt(mu*b) (sg_set_damping_local)
    void *vs = nullptr;       // local S velocity per split slot (compute type)

    bool use_graph = false;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;

    std::vector<hipEvent_t> events;

    // device copy of the shot descriptor and the tile work-lists (tiled sweeps)
    void *d_shot = nullptr;
    WorkItem *d_wl = nullptr;  // the work-lists back to back
    int wl_off[NLIST + 1] = {0, 0, 0, 0, 0, 0};
    bool wl_valid = false;

    alignas(16) unsigned char shot[sizeof(Shot<double, double>)];
    float fscale[NFIELD];
    float bscale = 1, lscale = 1, mscale = 1;
};

namespace {

// The source point lies in a PML slab (some axis damps there).
bool src_in_pml(const sg_solver *s)
{
    if (s->src < 0 || s->g.pml == 0) return false;
    const Geom &g = s->g;
    const long long i = s->src % g.nx, j = (s->src / g.nx) % g.ny, k = s->src / ((long long)g.nx * g.ny);
    return in_pml((int)i, g.nx, g.pml) || in_pml((int)j, g.ny, g.pml) || in_pml((int)k, g.nz, g.pml);
}

// How the inject kernel adds the source (INJECT_*): the CPU path's transient
// double amount; the CUDA path's float amount kept for good in its unsplit
// interior; and, for a CUDA-path source inside the PML, where that path
// rebuilds sigma from its splits (..._concise/ElasticMacro.h:466-543), the
// float amount dropped after one velocity update.
int inject_mode(const sg_solver *s)
{
    if (s->cfg.source_mode != SG_SOURCE_ACCUMULATE) return INJECT_TRANSIENT;
    return src_in_pml(s) ? INJECT_TRANSIENT_F : INJECT_ACCUMULATE;
}

template <class T, class C>
void fill_shot(sg_solver *s)
{
    Shot<T, C> &h = *reinterpret_cast<Shot<T, C> *>(s->shot);
    std::memset(&h, 0, sizeof h);
    h.g = s->g;
    for (int f = 0; f < NFIELD; ++f) h.f[f] = static_cast<T *>(s->fields[f]);
    h.b = static_cast<const T *>(s->model[0]);
    h.lam = static_cast<const T *>(s->model[1]);
    h.mu = static_cast<const T *>(s->model[2]);
    C *sp = static_cast<C *>(s->splits);
    for (int a = 0; a < 9; ++a) h.sv[a] = sp ? sp + (size_t)a * s->g.nslots : nullptr;
    for (int a = 0; a < NSSPLIT; ++a) h.ss[a] = sp ? sp + (size_t)(9 + a) * s->g.nslots : nullptr;
    C *tb = static_cast<C *>(s->tables);
    const int n3[3] = {s->g.nx, s->g.ny, s->g.nz};
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        h.dA[a] = tb + off;
        h.dB[a] = tb + off + n3[a];
        off += 2 * (size_t)n3[a];
    }
    const double hs[3] = {s->cfg.dx, s->cfg.dy, s->cfg.dz};
    for (int n = 0; n < 5; ++n) {
        h.cx[n] = (C)((double)s->cfg.diff_coef[n] / hs[0]);
        h.cy[n] = (C)((double)s->cfg.diff_coef[n] / hs[1]);
        h.cz[n] = (C)((double)s->cfg.diff_coef[n] / hs[2]);
    }
    h.dt = (C)s->cfg.dt;
    h.src = s->src;
    h.src_saved = (s->src >= 0 && inject_mode(s) != INJECT_ACCUMULATE) ? static_cast<const T *>(s->saved) : nullptr;
    for (int f = 0; f < NFIELD; ++f) {
        h.fscale[f] = s->fscale[f];
        h.finv[f] = 1.0f / s->fscale[f];
    }
    h.bscale = s->bscale;
    h.lscale = s->lscale;
    h.mscale = s->mscale;
    h.round_mode = s->round_mode;
    h.d_step = s->d_step;
    h.vs = s->damp_local ? static_cast<const C *>(s->vs) : nullptr;
    for (int f = 0; f < NFIELD; ++f) {
        h.comp[f] = static_cast<T *>(s->comp[f]);
        h.cscale[f] = s->fscale[f] * COMP_GAIN;
        h.cinv[f] = 1.0f / (s->fscale[f] * COMP_GAIN);
    }
}

size_t shot_size(int prec)
{
    switch (prec) {
    case SG_FP32: return sizeof(Shot<float, float>);
    case SG_FP64: return sizeof(Shot<double, double>);
    case SG_FP16: return sizeof(Shot<_Float16, float>);
    default: return sizeof(Shot<bf16, float>);
    }
}

// rebuilds the host descriptor and copies it to the device (callers hold no
// queued work on the stream: every setter synchronizes first)
int refresh_shot(sg_solver *s)
{
    switch (s->prec) {
    case SG_FP32: fill_shot<float, float>(s); break;
    case SG_FP64: fill_shot<double, double>(s); break;
    case SG_FP16: fill_shot<_Float16, float>(s); break;
    default: fill_shot<bf16, float>(s); break;
    }
    s->graph_valid = false;
    if (s->d_shot) SG_HIP(hipMemcpy(s->d_shot, s->shot, shot_size(s->prec), hipMemcpyHostToDevice));
    return SG_OK;
}

int rebuild_worklists(sg_solver *s)
{
    WorkLists wl;
    const int zc = s->tile.zchunk > 0 ? s->tile.zchunk : auto_zchunk(s->g, s->tile.ty);
    build_worklists(s->g, 0, s->tile.ty, s->tile.pty, zc, wl);
    const std::vector<WorkItem> a = pack_worklists(wl, s->wl_off);
    if (s->d_wl) SG_HIP(hipFree(s->d_wl));
    s->d_wl = nullptr;
    if (!a.empty()) {
        SG_HIP(hipMalloc(&s->d_wl, a.size() * sizeof(WorkItem)));
        SG_HIP(hipMemcpy(s->d_wl, a.data(), a.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    }
    s->wl_valid = true;
    s->graph_valid = false;
    return SG_OK;
}

void drop_graph(sg_solver *s)
{
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
    s->graph_valid = false;
}

int upload_tables(sg_solver *s)
{
    const int n3[3] = {s->g.nx, s->g.ny, s->g.nz};
    const double dt = s->cfg.dt;
    std::vector<double> buf;
    for (int a = 0; a < 3; ++a) {
        std::vector<double> A(n3[a]), B(n3[a]);
        for (int i = 0; i < n3[a]; ++i) {
            const double hd = 0.5 * dt * s->damp[a][i];
            if (s->damp_local) {
                // per-cell damping: the kernels form A, B from dt*g/2 and the local velocity
                A[i] = hd;
                B[i] = 0.0;
            } else {
                A[i] = (1.0 - hd) / (1.0 + hd);
                B[i] = 1.0 / (1.0 + hd);
            }
        }
        buf.insert(buf.end(), A.begin(), A.end());
        buf.insert(buf.end(), B.begin(), B.end());
    }
    if (compute_size(s->prec) == 8) {
        SG_HIP(hipMemcpy(s->tables, buf.data(), buf.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(buf.begin(), buf.end());
        SG_HIP(hipMemcpy(s->tables, f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
    return SG_OK;
}

int enqueue_step(sg_solver *s, hipEvent_t e0, hipEvent_t e1, hipEvent_t e2)
{
    SG_HIP(launch_inject(s->prec, s->shot, s->amp_d, s->amp_acc, s->namp, inject_mode(s), s->d_step, s->saved,
                         s->stream));
    if (e0) SG_HIP(hipEventRecord(e0, s->stream));
    if (s->cfg.kernel == SG_KERNEL_SIMPLE) {
        SG_HIP(launch_velocity_simple(s->prec, s->shot, s->stream));
        if (e1) SG_HIP(hipEventRecord(e1, s->stream));
        SG_HIP(launch_stress_simple(s->prec, s->shot, s->stream));
    } else {
        const Tile &t = s->tile;
        // lists of one sweep write disjoint points and splits and only read
        // what the other sweep writes, so they may run concurrently:
        // interior on the solver stream, x faces on side 0, the rest on side 1
        const int ns = s->nstreams;
        // one stream and (4,1) boxes throughout: the five lists of a sweep in
        // one launch (small grids are launch-bound)
        const bool fused = ns == 1 && tile_fusable(t);
        for (int which = 0; fused && which < 2; ++which) {
            if (which == 1 && e1) SG_HIP(hipEventRecord(e1, s->stream));
            SG_HIP(launch_sweep_fused(which, s->prec, s->comp_mask != 0, t.waves, s->d_shot, s->d_wl, s->wl_off, 0,
                                      s->stream));
        }
        for (int which = 0; !fused && which < 2; ++which) {
            if (which == 1 && e1) SG_HIP(hipEventRecord(e1, s->stream));
            if (ns > 1) {
                SG_HIP(hipEventRecord(s->fork, s->stream));
                for (int q = 0; q < ns - 1; ++q) SG_HIP(hipStreamWaitEvent(s->side[q], s->fork, 0));
            }
            // without PML boxes the interior list runs alone and its FP32
            // kernels load their staging rows straight into LDS (higher
            // occupancy; beside the PML lists that gains nothing)
            const bool lds_rows = s->prec == SG_FP32 && s->comp_mask == 0 && s->wl_off[1] == s->wl_off[NLIST];
            for (int l = 0; l < NLIST; ++l) {
                // list -> stream: 1 stream: all on the solver stream; 2: PML lists on side 0;
                // 3: x faces on side 0, the other PML lists on side 1; 4: edges on side 2
                static const int map[5][NLIST] = {{0}, {0, 0, 0, 0, 0}, {0, 1, 1, 1, 1}, {0, 1, 2, 2, 2}, {0, 1, 2, 2, 3}};
                const int q = map[ns][l];
                hipStream_t st = q == 0 ? s->stream : s->side[q - 1];
                const int dm = (l == 0 && lds_rows) ? DM_INTERIOR_LDS : LIST_DM[l];
                SG_HIP(launch_sweep(which, s->prec, l ? t.pty : t.ty, l ? t.pny : t.ny, dm, s->comp_mask != 0,
                                    s->d_shot, s->d_wl + s->wl_off[l], s->wl_off[l + 1] - s->wl_off[l], 0, st));
            }
            if (ns > 1) {
                for (int q = 0; q < ns - 1; ++q) {
                    SG_HIP(hipEventRecord(s->join[q], s->side[q]));
                    SG_HIP(hipStreamWaitEvent(s->stream, s->join[q], 0));
                }
            }
        }
    }
    if (e2) SG_HIP(hipEventRecord(e2, s->stream));
    if (s->nst > 0)
        SG_HIP(launch_gather(s->prec, s->shot, s->st_idx, s->st_comp, s->nst, s->ncomp, s->interval, s->st_cap,
                             s->st_out, s->d_step, s->stream));
    return SG_OK;
}

int ensure_worklists(sg_solver *s)
{
    if (s->cfg.kernel == SG_KERNEL_TILED && !s->wl_valid) return rebuild_worklists(s);
    return SG_OK;
}

int build_graph(sg_solver *s)
{
    drop_graph(s);
    SG_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(s, nullptr, nullptr, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc != SG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SG_HIP(e);
    s->graph = g;
    SG_HIP(hipGraphInstantiate(&s->gexec, s->graph, nullptr, nullptr, 0));
    s->graph_valid = true;
    return SG_OK;
}

void free_all(sg_solver *s)
{
    drop_graph(s);
    for (auto ev : s->events) (void)hipEventDestroy(ev);
    s->events.clear();
    void *ptrs[] = {s->splits, s->tables, s->saved, s->d_step, s->amp_d, s->amp_acc, s->vs,
                    s->st_idx, s->st_comp, s->st_out, s->d_energy, s->d_shot, s->d_wl};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int f = 0; f < NFIELD; ++f) {
        if (s->fields[f]) (void)hipFree(s->fields[f]);
        if (s->comp[f]) (void)hipFree(s->comp[f]);
    }
    for (int m = 0; m < 3; ++m)
        if (s->model[m]) (void)hipFree(s->model[m]);
    for (int q = 0; q < 3; ++q) {
        if (s->side[q]) (void)hipStreamDestroy(s->side[q]);
        if (s->join[q]) (void)hipEventDestroy(s->join[q]);
    }
    if (s->fork) (void)hipEventDestroy(s->fork);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Device buffers allocated together by a setter: all freed on scope exit
// unless keep() hands them over, so a failed call leaves the solver as it was.
struct DevBuffers {
    std::vector<void *> p;
    int alloc(void **out, size_t bytes, const char *who)
    {
        *out = nullptr;
        const hipError_t e = hipMalloc(out, bytes);
        if (e != hipSuccess) {
            *out = nullptr;
            (void)hipGetLastError();
            return fail(e == hipErrorOutOfMemory ? SG_ERR_NOMEM : SG_ERR_HIP,
                        std::string(who) + ": device allocation of " + std::to_string(bytes) +
                            " bytes failed: " + hipGetErrorString(e));
        }
        p.push_back(*out);
        return SG_OK;
    }
    void keep() { p.clear(); }
    ~DevBuffers()
    {
        for (void *q : p) (void)hipFree(q);
    }
};

// Waits for the solver's stream and for every batch stream that steps this
// solver: a setter or getter must not overlap batched sweeps that read or
// write the same buffers.
int quiesce(sg_solver *s)
{
    SG_HIP(hipStreamSynchronize(s->stream));
    for (hipStream_t b : s->batch_streams) SG_HIP(hipStreamSynchronize(b));
    return SG_OK;
}

#define SG_QUIESCE(s)                     \
    do {                                  \
        if (int rc_q_ = quiesce(s)) return rc_q_; \
    } while (0)

int check_solver(const sg_solver *s)
{
    if (!s) return fail(SG_ERR_INVALID, "solver handle is NULL");
    return SG_OK;
}

// Device scratch of sg_energy_noalloc: the per-plane partial sums and the
// results of the energy reduction, for grids of up to NOALLOC_MAX_PLANES
// updated z planes (512 KiB).  Each GPU holds its own copy for the life of
// the process; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch).
constexpr int NOALLOC_MAX_PLANES = 16384;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[ENERGY_PART * NOALLOC_MAX_PLANES + 4];
std::mutex g_scratch_mutex;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The scratch address on `device` (current): looked up once per device
// (sg_create does it for every device it creates a solver on).
int noalloc_scratch(int device, double **out)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(SG_ERR_UNSUPPORTED, "device index beyond the scratch table (64 devices)");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[device];
    return SG_OK;
}

}  // namespace

extern "C" {

int sg_api_version(void) { return SG_API_VERSION; }

const char *sg_last_error(void) { return sg::last_error().c_str(); }

int sg_create(const sg_config *cfg, sg_solver **out)
{
    if (!cfg || !out) return fail(SG_ERR_INVALID, "sg_create: NULL argument");
    *out = nullptr;
    const sg_config c = *cfg;
    if (c.nx < 11 || c.ny < 11 || c.nz < 11) return fail(SG_ERR_INVALID, "sg_create: each extent must be >= 11");
    if (c.pml < 0 || 2 * c.pml > c.nx - 10 || 2 * c.pml > c.ny - 10 || 2 * c.pml > c.nz - 10)
        return fail(SG_ERR_INVALID, "sg_create: pml must satisfy 0 <= 2*pml <= n-10 on every axis");
    if (!(c.dt > 0) || !(c.dx > 0) || !(c.dy > 0) || !(c.dz > 0))
        return fail(SG_ERR_INVALID, "sg_create: dt, dx, dy, dz must be positive");
    if (c.time_length < 0) return fail(SG_ERR_INVALID, "sg_create: time_length must be >= 0");
    if (c.precision < SG_FP32 || c.precision > SG_BF16) return fail(SG_ERR_INVALID, "sg_create: bad precision");
    if (c.source_mode != SG_SOURCE_TRANSIENT && c.source_mode != SG_SOURCE_ACCUMULATE)
        return fail(SG_ERR_INVALID, "sg_create: bad source_mode");
    if (c.kernel != SG_KERNEL_TILED && c.kernel != SG_KERNEL_SIMPLE)
        return fail(SG_ERR_INVALID, "sg_create: bad kernel");
    // FP64 is the checking build and runs the per-point kernels: its tiled
    // sweeps (256 VGPRs plus AGPR and SGPR spills) took GPU memory faults on
    // small grids whose cause is not found, so they are not offered
    if (c.kernel == SG_KERNEL_TILED && c.precision == SG_FP64)
        return fail(SG_ERR_UNSUPPORTED, "sg_create: FP64 runs the per-point kernels (SG_KERNEL_SIMPLE); "
                                        "the tiled FP64 sweeps are disabled");
    const long long N = (long long)c.nx * c.ny * c.nz;
    // indices are 64-bit; the tiled sweeps address inside one xy plane with
    // 32-bit byte offsets
    if ((long long)c.nx * c.ny * 8 >= (1LL << 32))
        return fail(SG_ERR_INVALID, "sg_create: an xy plane must hold fewer than 2^29 points");

    sg_solver *s = new (std::nothrow) sg_solver();
    if (!s) return fail(SG_ERR_NOMEM, "sg_create: host allocation failed");
    s->cfg = c;
    s->prec = c.precision;
    for (int f = 0; f < NFIELD; ++f) s->fscale[f] = 1.0f;
    // below ~4 M points a step is launch-bound and the fork/join of side
    // streams costs more than the overlap gains (measured at 100^3: 8.5
    // against 6.8 Gpt/s; at 256^3 three streams give 21.4 against 17.5)
    s->nstreams = N < (4LL << 20) ? 1 : 3;

    Geom &g = s->g;
    g.nx = c.nx;
    g.ny = c.ny;
    g.nz = c.nz;
    g.pml = c.pml;
    g.ux = c.nx - 2 * HALO;
    g.uy = c.ny - 2 * HALO;
    g.uz = c.nz - 2 * HALO;
    g.sy = c.nx;
    g.sz = (long long)c.nx * c.ny;
    split_layout(g);
    // the tiled sweeps address inside one split array with 32-bit byte offsets
    if (c.kernel == SG_KERNEL_TILED && g.nslots * (long long)compute_size(c.precision) >= (1LL << 32)) {
        delete s;
        return fail(SG_ERR_UNSUPPORTED, "sg_create: PML split arrays too large for the tiled sweeps "
                                        "(slots * element size >= 4 GiB); use SG_KERNEL_SIMPLE");
    }

    auto bail = [&](int rc) {
        free_all(s);
        delete s;
        return rc;
    };
    int rc;
#define SG_TRY(call)                      \
    do {                                  \
        hipError_t e_ = (call);           \
        if (e_ != hipSuccess) {           \
            rc = fail(e_ == hipErrorOutOfMemory ? SG_ERR_NOMEM : SG_ERR_HIP, \
                      std::string("sg_create: ") + #call + ": " + hipGetErrorString(e_)); \
            return bail(rc);              \
        }                                 \
    } while (0)

    SG_TRY(hipGetDevice(&s->device));
    {
        // the scratch address of sg_energy_noalloc on this device
        double *scr = nullptr;
        if ((rc = noalloc_scratch(s->device, &scr)) != SG_OK) return bail(rc);
    }
    SG_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (int q = 0; q < 3; ++q) {
        SG_TRY(hipStreamCreateWithFlags(&s->side[q], hipStreamNonBlocking));
        SG_TRY(hipEventCreateWithFlags(&s->join[q], hipEventDisableTiming));
    }
    SG_TRY(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming));
    const size_t sT = storage_size(s->prec), sC = compute_size(s->prec);
    for (int f = 0; f < NFIELD; ++f) {
        SG_TRY(hipMalloc(&s->fields[f], N * sT));
        SG_TRY(hipMemsetAsync(s->fields[f], 0, N * sT, s->stream));
    }
    for (int m = 0; m < 3; ++m) {
        SG_TRY(hipMalloc(&s->model[m], N * sT));
        SG_TRY(hipMemsetAsync(s->model[m], 0, N * sT, s->stream));
    }
    if (g.nslots > 0) {
        SG_TRY(hipMalloc(&s->splits, (size_t)g.nslots * 24 * sC));
        SG_TRY(hipMemsetAsync(s->splits, 0, (size_t)g.nslots * 24 * sC, s->stream));
    }
    SG_TRY(hipMalloc(&s->tables, 2 * (size_t)(c.nx + c.ny + c.nz) * sC));
    SG_TRY(hipMalloc(&s->saved, 8 * sT));
    SG_TRY(hipMemsetAsync(s->saved, 0, 8 * sT, s->stream));
    SG_TRY(hipMalloc(&s->d_step, sizeof(int)));
    SG_TRY(hipMemsetAsync(s->d_step, 0, sizeof(int), s->stream));
    SG_TRY(hipMalloc(&s->d_energy, (size_t)(ENERGY_PART * g.uz + 4) * sizeof(double) + NFIELD * sizeof(unsigned)));
    SG_TRY(hipMalloc(&s->d_shot, sizeof(Shot<double, double>)));
    SG_TRY(hipStreamSynchronize(s->stream));
#undef SG_TRY

    // default damping: the reference CPU formula (tau 3.4, V 2000, R 1e-3)
    const int n3[3] = {c.nx, c.ny, c.nz};
    const double h3[3] = {c.dx, c.dy, c.dz};
    for (int a = 0; a < 3; ++a) {
        s->damp[a].assign(n3[a], 0.0);
        if (c.pml > 0) sg_damping_profile(n3[a], c.pml, h3[a], SG_DAMP_REFERENCE, 2000.0, 1e-3, 3.4, s->damp[a].data());
    }
    if ((rc = upload_tables(s)) != SG_OK) return bail(rc);
    if ((rc = refresh_shot(s)) != SG_OK) return bail(rc);
    *out = s;
    return SG_OK;
}

int sg_destroy(sg_solver *s)
{
    if (!s) return SG_OK;
    if (!s->batch_streams.empty())
        return fail(SG_ERR_STATE, "sg_destroy: the solver belongs to a live batch; destroy the batch first");
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return SG_OK;
}

static int fill_vs(sg_solver *s);

int sg_set_medium(sg_solver *s, const float *b, const float *lam, const float *mu)
{
    if (int rc = check_solver(s)) return rc;
    if (!b || !lam || !mu) return fail(SG_ERR_INVALID, "sg_set_medium: NULL array");
    const size_t N = (size_t)s->g.nx * s->g.ny * s->g.nz;
    const float *src[3] = {b, lam, mu};
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    double courant = 0.0;
    if (int rc = sg_courant_number((long long)N, b, lam, mu, s->cfg.dt, s->cfg.dx, s->cfg.dy, s->cfg.dz,
                                   s->cfg.diff_coef, &courant))
        return fail(rc, std::string("sg_set_medium: ") + sg_last_error());
    if (s->courant_limit > 0 && courant > s->courant_limit) {
        char m[256];
        snprintf(m, sizeof m, "sg_set_medium: unstable model: Courant number %.6g exceeds the limit %.6g "
                 "(dt * max vp * sum|c_n| * sqrt(sum 1/h^2); reduce dt or raise the limit with sg_set_stability)",
                 courant, s->courant_limit);
        return fail(SG_ERR_INVALID, m);
    }
    s->courant = courant;
    if (s->prec == SG_FP32) {
        for (int m = 0; m < 3; ++m) SG_HIP(hipMemcpy(s->model[m], src[m], N * 4, hipMemcpyHostToDevice));
    } else if (s->prec == SG_FP64) {
        std::vector<double> d(N);
        for (int m = 0; m < 3; ++m) {
            for (size_t i = 0; i < N; ++i) d[i] = src[m][i];
            SG_HIP(hipMemcpy(s->model[m], d.data(), N * 8, hipMemcpyHostToDevice));
        }
    } else {
        // 16-bit storage: one power-of-two scale per model array
        float *sc[3] = {&s->bscale, &s->lscale, &s->mscale};
        std::vector<uint16_t> h(N);
        for (int m = 0; m < 3; ++m) {
            double mx = 0;
            for (size_t i = 0; i < N; ++i) mx = std::max(mx, (double)std::fabs(src[m][i]));
            *sc[m] = pow2_scale(mx);
            for (size_t i = 0; i < N; ++i)
                h[i] = s->prec == SG_FP16 ? to_half_bits(src[m][i] * *sc[m]) : to_bf16_bits(src[m][i] * *sc[m]);
            SG_HIP(hipMemcpy(s->model[m], h.data(), N * 2, hipMemcpyHostToDevice));
        }
        if (int rc = refresh_shot(s)) return rc;
    }
    return fill_vs(s);
}

int sg_set_medium_velocity(sg_solver *s, const float *vp, const float *vs, const float *rho)
{
    if (int rc = check_solver(s)) return rc;
    const long long N = (long long)s->g.nx * s->g.ny * s->g.nz;
    std::vector<float> b, lam, mu;
    try {
        b.resize((size_t)N);
        lam.resize((size_t)N);
        mu.resize((size_t)N);
    } catch (const std::bad_alloc &) {
        return fail(SG_ERR_NOMEM, "sg_set_medium_velocity: host allocation failed");
    }
    if (int rc = sg_elastic_from_velocity(N, vp, vs, rho, b.data(), lam.data(), mu.data())) return rc;
    return sg_set_medium(s, b.data(), lam.data(), mu.data());
}

int sg_set_field_scales(sg_solver *s, const float scales[9])
{
    if (int rc = check_solver(s)) return rc;
    if (!scales) return fail(SG_ERR_INVALID, "sg_set_field_scales: NULL");
    for (int f = 0; f < NFIELD; ++f) {
        const float v = scales[f];
        int e;
        if (!(v > 0) || !std::isfinite(v) || std::frexp(v, &e) != 0.5f)
            return fail(SG_ERR_INVALID, "sg_set_field_scales: scales must be positive powers of two");
    }
    if (s->steps_done != 0) return fail(SG_ERR_STATE, "sg_set_field_scales: set scales before stepping (or after sg_reset)");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    for (int f = 0; f < NFIELD; ++f) s->fscale[f] = s->prec == SG_FP16 || s->prec == SG_BF16 ? scales[f] : 1.0f;
    return refresh_shot(s);
}

int sg_get_field_scales(const sg_solver *s, float out[12])
{
    if (!s || !out) return fail(SG_ERR_INVALID, "sg_get_field_scales: NULL");
    for (int f = 0; f < NFIELD; ++f) out[f] = s->fscale[f];
    out[9] = s->bscale;
    out[10] = s->lscale;
    out[11] = s->mscale;
    return SG_OK;
}

int sg_set_damping(sg_solver *s, const double *dx, const double *dy, const double *dz)
{
    if (int rc = check_solver(s)) return rc;
    const double *d3[3] = {dx, dy, dz};
    const int n3[3] = {s->g.nx, s->g.ny, s->g.nz};
    for (int a = 0; a < 3; ++a) {
        if (!d3[a]) return fail(SG_ERR_INVALID, "sg_set_damping: NULL profile");
        for (int i = 0; i < n3[a]; ++i)
            if (!std::isfinite(d3[a][i]) || d3[a][i] < 0) return fail(SG_ERR_INVALID, "sg_set_damping: d must be finite and >= 0");
    }
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    for (int a = 0; a < 3; ++a) s->damp[a].assign(d3[a], d3[a] + n3[a]);
    const bool was_local = s->damp_local;
    s->damp_local = false;
    if (int rc = upload_tables(s)) return rc;
    return was_local ? refresh_shot(s) : SG_OK;
}

// the local S velocity of every split slot from the medium on the device
static int fill_vs(sg_solver *s)
{
    if (!s->damp_local || s->g.nslots == 0) return SG_OK;
    if (!s->vs) SG_HIP(hipMalloc(&s->vs, (size_t)s->g.nslots * compute_size(s->prec)));
    if (int rc = refresh_shot(s)) return rc;
    SG_HIP(launch_fill_vs(s->prec, s->shot, s->vs, s->stream));
    SG_QUIESCE(s);
    return SG_OK;
}

int sg_set_damping_local(sg_solver *s, const double *gx, const double *gy, const double *gz)
{
    if (int rc = check_solver(s)) return rc;
    const double *g3[3] = {gx, gy, gz};
    const int n3[3] = {s->g.nx, s->g.ny, s->g.nz};
    for (int a = 0; a < 3; ++a) {
        if (!g3[a]) return fail(SG_ERR_INVALID, "sg_set_damping_local: NULL profile");
        for (int i = 0; i < n3[a]; ++i)
            if (!std::isfinite(g3[a][i]) || g3[a][i] < 0)
                return fail(SG_ERR_INVALID, "sg_set_damping_local: g must be finite and >= 0");
    }
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    for (int a = 0; a < 3; ++a) s->damp[a].assign(g3[a], g3[a] + n3[a]);
    s->damp_local = true;
    if (int rc = upload_tables(s)) return rc;
    return fill_vs(s);
}

int sg_set_source(sg_solver *s, int x, int y, int z, const float *wavelet, int n)
{
    if (int rc = check_solver(s)) return rc;
    const Geom &g = s->g;
    if (x < HALO || x >= g.nx - HALO || y < HALO || y >= g.ny - HALO || z < HALO || z >= g.nz - HALO)
        return fail(SG_ERR_RANGE, "sg_set_source: source must lie in the updated region [5, n-5)");
    if (n < 0 || (n > 0 && !wavelet)) return fail(SG_ERR_INVALID, "sg_set_source: bad wavelet");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    // new buffers first; the solver changes only once they are complete
    DevBuffers nb;
    double *amp_d = nullptr;
    void *amp_acc = nullptr;
    if (n > 0) {
        // TRANSIENT: pow(dt, 2) * s(it) in double (ElasticWaveEquation3D.cpp:123)
        std::vector<double> ad(n);
        const double dt = (double)s->cfg.dt;
        for (int i = 0; i < n; ++i) ad[i] = dt * dt * (double)wavelet[i];
        const size_t sC = compute_size(s->prec);
        if (int rc = nb.alloc((void **)&amp_d, n * sizeof(double), "sg_set_source")) return rc;
        if (int rc = nb.alloc(&amp_acc, n * sC, "sg_set_source")) return rc;
        SG_HIP(hipMemcpy(amp_d, ad.data(), n * sizeof(double), hipMemcpyHostToDevice));
        // ACCUMULATE: deltaT * deltaT * sourceValue in the compute type (..._concise/.cu:216)
        if (sC == 8) {
            SG_HIP(hipMemcpy(amp_acc, ad.data(), n * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> a(n);
            const float dtf = s->cfg.dt;
            for (int i = 0; i < n; ++i) a[i] = dtf * dtf * wavelet[i];
            SG_HIP(hipMemcpy(amp_acc, a.data(), n * 4, hipMemcpyHostToDevice));
        }
    }
    nb.keep();
    if (s->amp_d) (void)hipFree(s->amp_d);
    if (s->amp_acc) (void)hipFree(s->amp_acc);
    s->amp_d = amp_d;
    s->amp_acc = amp_acc;
    s->namp = n;
    {
        float peak = 0.0f;
        for (int i = 0; i < n; ++i) peak = std::max(peak, std::fabs(wavelet[i]));
        s->src_last = -1;
        for (int i = 0; i < n; ++i)
            if (std::fabs(wavelet[i]) > 1e-6f * peak) s->src_last = i;
    }
    s->src = x + (long long)y * g.sy + (long long)z * g.sz;
    return refresh_shot(s);
}

int sg_set_source_sample(sg_solver *s, int it, float value)
{
    if (int rc = check_solver(s)) return rc;
    if (it < 0 || it >= s->namp) return fail(SG_ERR_RANGE, "sg_set_source_sample: index outside the wavelet");
    SG_HIP(hipSetDevice(s->device));
    const double dt = (double)s->cfg.dt;
    const double ad = dt * dt * (double)value;
    // after every step already queued on this solver or a batch holding it
    SG_QUIESCE(s);
    SG_HIP(hipMemcpy(s->amp_d + it, &ad, sizeof ad, hipMemcpyHostToDevice));
    if (compute_size(s->prec) == 8) {
        SG_HIP(hipMemcpy(static_cast<double *>(s->amp_acc) + it, &ad, 8, hipMemcpyHostToDevice));
    } else {
        const float af = s->cfg.dt * s->cfg.dt * value;
        SG_HIP(hipMemcpy(static_cast<float *>(s->amp_acc) + it, &af, 4, hipMemcpyHostToDevice));
    }
    return SG_OK;
}

int sg_set_stations(sg_solver *s, const int *xyz, int n, unsigned comp_mask, int interval)
{
    if (int rc = check_solver(s)) return rc;
    if (n < 0 || (n > 0 && !xyz) || interval < 1 || (comp_mask & ~0x1FFu))
        return fail(SG_ERR_INVALID, "sg_set_stations: bad arguments");
    if (n > 0 && (comp_mask & 0x1FFu) == 0) return fail(SG_ERR_INVALID, "sg_set_stations: empty component mask");
    const Geom &g = s->g;
    std::vector<long long> idx(n);
    for (int i = 0; i < n; ++i) {
        const int x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        if (x < 0 || x >= g.nx || y < 0 || y >= g.ny || z < 0 || z >= g.nz)
            return fail(SG_ERR_RANGE, "sg_set_stations: station " + std::to_string(i) + " outside the grid");
        idx[i] = x + (long long)y * g.sy + (long long)z * g.sz;
    }
    std::vector<int> comp;
    for (int f = 0; f < NFIELD; ++f)
        if (comp_mask & (1u << f)) comp.push_back(f);
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    // new buffers first; the solver changes only once they are complete
    DevBuffers nb;
    long long *st_idx = nullptr;
    int *st_comp = nullptr;
    void *st_out = nullptr;
    long long cap = 0;
    if (n > 0) {
        const long long steps = std::max(1, s->cfg.time_length);
        cap = (steps + interval - 1) / interval;
        const size_t bytes = (size_t)cap * n * comp.size() * compute_size(s->prec);
        if (int rc = nb.alloc((void **)&st_idx, n * sizeof(long long), "sg_set_stations")) return rc;
        if (int rc = nb.alloc((void **)&st_comp, comp.size() * sizeof(int), "sg_set_stations")) return rc;
        if (int rc = nb.alloc(&st_out, bytes, "sg_set_stations")) return rc;
        SG_HIP(hipMemcpy(st_idx, idx.data(), n * sizeof(long long), hipMemcpyHostToDevice));
        SG_HIP(hipMemcpy(st_comp, comp.data(), comp.size() * sizeof(int), hipMemcpyHostToDevice));
        SG_HIP(launch_fill_zero(st_out, (long long)bytes, s->stream));
        SG_HIP(hipStreamSynchronize(s->stream));
    }
    nb.keep();
    if (s->st_idx) (void)hipFree(s->st_idx);
    if (s->st_comp) (void)hipFree(s->st_comp);
    if (s->st_out) (void)hipFree(s->st_out);
    s->st_idx = st_idx;
    s->st_comp = st_comp;
    s->st_out = st_out;
    s->nst = n;
    s->ncomp = (int)comp.size();
    s->interval = interval;
    s->st_cap = cap;
    s->graph_valid = false;
    return SG_OK;
}

// The stability check after the step just completed: FP64 energies of the
// current state; SG_ERR_UNSTABLE for a non-finite wavefield, or for energy
// growth beyond max_growth between two checks once the source has stopped.
static int stability_check(sg_solver *s)
{
    double *res = s->d_energy + (size_t)ENERGY_PART * s->g.uz;
    double e[3];
    SG_HIP(launch_energy(s->prec, s->shot, s->d_energy, res, s->stream));
    SG_HIP(hipMemcpyAsync(e, res, sizeof e, hipMemcpyDeviceToHost, s->stream));
    SG_HIP(hipStreamSynchronize(s->stream));
    char m[320];
    if (!std::isfinite(e[0]) || !std::isfinite(e[1]) || !std::isfinite(e[2])) {
        snprintf(m, sizeof m, "sg_step: unstable: a wavefield holds a non-finite value after step %lld",
                 s->steps_done);
        return fail(SG_ERR_UNSTABLE, m);
    }
    const double tot = e[0] + e[2];
    if (s->steps_done > (long long)s->src_last + 1) {
        if (s->last_energy > 0 && tot > s->max_growth * s->last_energy) {
            snprintf(m, sizeof m, "sg_step: unstable: energy grew from %.6g to %.6g (x%.3g, limit x%.3g) in the "
                     "%d steps to step %lld, after the source stopped", s->last_energy, tot, tot / s->last_energy,
                     s->max_growth, s->check_interval, s->steps_done);
            return fail(SG_ERR_UNSTABLE, m);
        }
        s->last_energy = tot;
    }
    return SG_OK;
}

static int enqueue_steps(sg_solver *s, int m)
{
    if (s->use_graph) {
        if (!s->graph_valid)
            if (int rc = build_graph(s)) return rc;
        for (int i = 0; i < m; ++i) SG_HIP(hipGraphLaunch(s->gexec, s->stream));
    } else {
        for (int i = 0; i < m; ++i)
            if (int rc = enqueue_step(s, nullptr, nullptr, nullptr)) return rc;
    }
    return SG_OK;
}

int sg_step(sg_solver *s, int nsteps)
{
    if (int rc = check_solver(s)) return rc;
    if (nsteps < 0) return fail(SG_ERR_INVALID, "sg_step: nsteps < 0");
    SG_HIP(hipSetDevice(s->device));
    // a shot stepped on its own after batched steps: those finish first
    for (hipStream_t b : s->batch_streams) SG_HIP(hipStreamSynchronize(b));
    if (int rc = ensure_worklists(s)) return rc;
    for (int done = 0; done < nsteps;) {
        int m = nsteps - done;
        const int ci = s->check_interval;
        if (ci > 0) m = (int)std::min<long long>(m, ci - s->steps_done % ci);
        if (int rc = enqueue_steps(s, m)) return rc;
        s->steps_done += m;
        done += m;
        if (ci > 0 && s->steps_done % ci == 0)
            if (int rc = stability_check(s)) return rc;
    }
    return SG_OK;
}

int sg_run(sg_solver *s)
{
    if (int rc = check_solver(s)) return rc;
    if (int rc = sg_step(s, s->cfg.time_length)) return rc;
    return sg_synchronize(s);
}

int sg_synchronize(sg_solver *s)
{
    if (int rc = check_solver(s)) return rc;
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    return SG_OK;
}

int sg_reset(sg_solver *s)
{
    if (int rc = check_solver(s)) return rc;
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    const size_t N = (size_t)s->g.nx * s->g.ny * s->g.nz, sT = storage_size(s->prec);
    for (int f = 0; f < NFIELD; ++f) {
        SG_HIP(hipMemsetAsync(s->fields[f], 0, N * sT, s->stream));
        if (s->comp[f]) SG_HIP(hipMemsetAsync(s->comp[f], 0, N * sT, s->stream));
    }
    if (s->splits) SG_HIP(hipMemsetAsync(s->splits, 0, (size_t)s->g.nslots * 24 * compute_size(s->prec), s->stream));
    SG_HIP(hipMemsetAsync(s->d_step, 0, sizeof(int), s->stream));
    if (s->st_out)
        SG_HIP(hipMemsetAsync(s->st_out, 0, (size_t)s->st_cap * s->nst * s->ncomp * compute_size(s->prec), s->stream));
    SG_QUIESCE(s);
    s->steps_done = 0;
    s->last_energy = -1.0;
    return SG_OK;
}

long long sg_steps_done(const sg_solver *s) { return s ? s->steps_done : -1; }

int sg_get_stations(sg_solver *s, float *out, long long count, long long *n_out)
{
    if (int rc = check_solver(s)) return rc;
    const long long rows = std::min<long long>(s->st_cap, (s->steps_done + s->interval - 1) / s->interval);
    const long long avail = rows * s->nst * s->ncomp;
    if (n_out) *n_out = avail;
    if (!out || count <= 0 || avail == 0) return SG_OK;
    const long long n = std::min(count, avail);
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    if (compute_size(s->prec) == 4) {
        SG_HIP(hipMemcpy(out, s->st_out, n * 4, hipMemcpyDeviceToHost));
    } else {
        std::vector<double> d(n);
        SG_HIP(hipMemcpy(d.data(), s->st_out, n * 8, hipMemcpyDeviceToHost));
        for (long long i = 0; i < n; ++i) out[i] = (float)d[i];
    }
    return SG_OK;
}

static int get_field_any(sg_solver *s, int field, float *outf, double *outd)
{
    if (int rc = check_solver(s)) return rc;
    if (field < 0 || field >= NFIELD) return fail(SG_ERR_RANGE, "sg_get_field: bad field id");
    const size_t N = (size_t)s->g.nx * s->g.ny * s->g.nz;
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    if (s->prec == SG_FP32) {
        if (outf) {
            SG_HIP(hipMemcpy(outf, s->fields[field], N * 4, hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(N);
            SG_HIP(hipMemcpy(t.data(), s->fields[field], N * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) outd[i] = t[i];
        }
    } else if (s->prec == SG_FP64) {
        if (outd) {
            SG_HIP(hipMemcpy(outd, s->fields[field], N * 8, hipMemcpyDeviceToHost));
        } else {
            std::vector<double> t(N);
            SG_HIP(hipMemcpy(t.data(), s->fields[field], N * 8, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i) outf[i] = (float)t[i];
        }
    } else {
        std::vector<uint16_t> t(N), c;
        SG_HIP(hipMemcpy(t.data(), s->fields[field], N * 2, hipMemcpyDeviceToHost));
        if (s->comp[field]) {
            c.resize(N);
            SG_HIP(hipMemcpy(c.data(), s->comp[field], N * 2, hipMemcpyDeviceToHost));
        }
        const float inv = 1.0f / s->fscale[field], cinv = 1.0f / (s->fscale[field] * COMP_GAIN);
        auto dec = [&](uint16_t b) { return s->prec == SG_FP16 ? from_half_bits(b) : from_bf16_bits(b); };
        for (size_t i = 0; i < N; ++i) {
            float v = dec(t[i]) * inv;
            if (!c.empty()) v += dec(c[i]) * cinv;
            if (outf) outf[i] = v;
            else outd[i] = v;
        }
    }
    return SG_OK;
}

int sg_get_field(sg_solver *s, int field, float *out)
{
    if (!out) return fail(SG_ERR_INVALID, "sg_get_field: out is NULL");
    return get_field_any(s, field, out, nullptr);
}

int sg_get_field_f64(sg_solver *s, int field, double *out)
{
    if (!out) return fail(SG_ERR_INVALID, "sg_get_field_f64: out is NULL");
    return get_field_any(s, field, nullptr, out);
}

int sg_energy(sg_solver *s, double out[2])
{
    if (int rc = check_solver(s)) return rc;
    if (!out) return fail(SG_ERR_INVALID, "sg_energy: out is NULL");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    double *res = s->d_energy + (size_t)ENERGY_PART * s->g.uz;
    SG_HIP(launch_energy(s->prec, s->shot, s->d_energy, res, s->stream));
    SG_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SG_QUIESCE(s);
    return SG_OK;
}

int sg_energy_full(sg_solver *s, double out[3])
{
    if (int rc = check_solver(s)) return rc;
    if (!out) return fail(SG_ERR_INVALID, "sg_energy_full: out is NULL");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    double *res = s->d_energy + (size_t)ENERGY_PART * s->g.uz;
    SG_HIP(launch_energy(s->prec, s->shot, s->d_energy, res, s->stream));
    SG_HIP(hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SG_QUIESCE(s);
    return SG_OK;
}

int sg_energy_noalloc(sg_solver *s, double out[2])
{
    if (int rc = check_solver(s)) return rc;
    if (!out) return fail(SG_ERR_INVALID, "sg_energy_noalloc: out is NULL");
    if (s->g.uz > NOALLOC_MAX_PLANES)
        return fail(SG_ERR_UNSUPPORTED, "sg_energy_noalloc: the grid has " + std::to_string(s->g.uz) +
                                            " updated z planes; the scratch holds " +
                                            std::to_string(NOALLOC_MAX_PLANES) + " (use sg_energy)");
    SG_HIP(hipSetDevice(s->device));
    double *scr = nullptr;
    if (int rc = noalloc_scratch(s->device, &scr)) return rc;
    SG_QUIESCE(s);
    // every partial and result is written before it is read: nothing to clear
    double *res = scr + (size_t)ENERGY_PART * s->g.uz;
    SG_HIP(launch_energy(s->prec, s->shot, scr, res, s->stream));
    SG_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    SG_HIP(hipStreamSynchronize(s->stream));
    return SG_OK;
}

int sg_set_stability(sg_solver *s, const sg_stability *st)
{
    if (int rc = check_solver(s)) return rc;
    if (!st) return fail(SG_ERR_INVALID, "sg_set_stability: NULL");
    if (!(st->courant_limit >= 0) || !std::isfinite(st->courant_limit) || st->check_interval < 0 ||
        !(st->max_growth > 1) || !std::isfinite(st->max_growth))
        return fail(SG_ERR_INVALID, "sg_set_stability: need courant_limit >= 0, check_interval >= 0, max_growth > 1");
    s->courant_limit = st->courant_limit;
    s->check_interval = st->check_interval;
    s->max_growth = st->max_growth;
    s->last_energy = -1.0;
    return SG_OK;
}

int sg_get_stability(const sg_solver *s, sg_stability *st)
{
    if (!s || !st) return fail(SG_ERR_INVALID, "sg_get_stability: NULL");
    st->courant_limit = s->courant_limit;
    st->check_interval = s->check_interval;
    st->max_growth = s->max_growth;
    return SG_OK;
}

int sg_courant(const sg_solver *s, double *courant)
{
    if (!s || !courant) return fail(SG_ERR_INVALID, "sg_courant: NULL");
    *courant = s->courant;
    return SG_OK;
}

int sg_set_compensation(sg_solver *s, unsigned field_mask)
{
    if (int rc = check_solver(s)) return rc;
    if (field_mask >> NFIELD) return fail(SG_ERR_INVALID, "sg_set_compensation: mask has bits above the nine fields");
    if (field_mask && s->prec != SG_FP16 && s->prec != SG_BF16)
        return fail(SG_ERR_UNSUPPORTED, "sg_set_compensation: only FP16/BF16 storage is compensated");
    if (field_mask && s->cfg.kernel != SG_KERNEL_TILED)
        return fail(SG_ERR_UNSUPPORTED, "sg_set_compensation: needs the tiled sweeps");
    if (s->steps_done != 0) return fail(SG_ERR_STATE, "sg_set_compensation: set it before stepping (or after sg_reset)");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    const size_t N = (size_t)s->g.nx * s->g.ny * s->g.nz, sT = storage_size(s->prec);
    for (int f = 0; f < NFIELD; ++f) {
        const bool on = (field_mask >> f) & 1u;
        if (on && !s->comp[f]) {
            hipError_t e = hipMalloc(&s->comp[f], N * sT);
            if (e != hipSuccess) {
                s->comp[f] = nullptr;
                return fail(e == hipErrorOutOfMemory ? SG_ERR_NOMEM : SG_ERR_HIP,
                            std::string("sg_set_compensation: ") + hipGetErrorString(e));
            }
        } else if (!on && s->comp[f]) {
            SG_HIP(hipFree(s->comp[f]));
            s->comp[f] = nullptr;
        }
        if (s->comp[f]) SG_HIP(hipMemsetAsync(s->comp[f], 0, N * sT, s->stream));
    }
    SG_QUIESCE(s);
    s->comp_mask = field_mask;
    s->graph_valid = false;
    return refresh_shot(s);
}

int sg_set_streams(sg_solver *s, int nstreams)
{
    if (int rc = check_solver(s)) return rc;
    if (nstreams < 1 || nstreams > 4) return fail(SG_ERR_INVALID, "sg_set_streams: nstreams must be 1 to 4");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    s->nstreams = nstreams;
    s->graph_valid = false;
    return SG_OK;
}

int sg_get_streams(const sg_solver *s, int *nstreams)
{
    if (int rc = check_solver(s)) return rc;
    if (!nstreams) return fail(SG_ERR_INVALID, "sg_get_streams: null output");
    *nstreams = s->nstreams;
    return SG_OK;
}

int sg_set_rounding(sg_solver *s, int mode)
{
    if (int rc = check_solver(s)) return rc;
    if (mode != SG_ROUND_NEAREST && mode != SG_ROUND_STOCHASTIC)
        return fail(SG_ERR_INVALID, "sg_set_rounding: mode must be SG_ROUND_NEAREST or SG_ROUND_STOCHASTIC");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    s->round_mode = mode;
    return refresh_shot(s);
}

int sg_field_maxabs(sg_solver *s, float out[9])
{
    if (int rc = check_solver(s)) return rc;
    if (!out) return fail(SG_ERR_INVALID, "sg_field_maxabs: out is NULL");
    SG_HIP(hipSetDevice(s->device));
    SG_QUIESCE(s);
    unsigned *d = reinterpret_cast<unsigned *>(s->d_energy + (size_t)ENERGY_PART * s->g.uz + 4);
    unsigned h[NFIELD];
    SG_HIP(launch_fill_zero(d, NFIELD * sizeof(unsigned), s->stream));
    SG_HIP(launch_maxabs(s->prec, s->shot, d, s->stream));
    SG_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    SG_QUIESCE(s);
    for (int f = 0; f < NFIELD; ++f) std::memcpy(&out[f], &h[f], sizeof(float));
    return SG_OK;
}

int sg_step_timed(sg_solver *s, int nsteps, double *ms_vel, double *ms_str, double *ms_total)
{
    if (int rc = check_solver(s)) return rc;
    if (nsteps < 0) return fail(SG_ERR_INVALID, "sg_step_timed: nsteps < 0");
    SG_HIP(hipSetDevice(s->device));
    if (int rc = ensure_worklists(s)) return rc;
    const int chunk = std::min(nsteps, 2048);
    const size_t need = 3 * (size_t)chunk + 2;
    while (s->events.size() < need) {
        hipEvent_t e;
        SG_HIP(hipEventCreate(&e));
        s->events.push_back(e);
    }
    hipEvent_t t0 = s->events[3 * chunk], t1 = s->events[3 * chunk + 1];
    double v = 0, st = 0;
    SG_HIP(hipEventRecord(t0, s->stream));
    for (int done = 0; done < nsteps;) {
        const int m = std::min(chunk, nsteps - done);
        for (int i = 0; i < m; ++i)
            if (int rc = enqueue_step(s, s->events[3 * i], s->events[3 * i + 1], s->events[3 * i + 2])) return rc;
        s->steps_done += m;
        SG_HIP(hipEventSynchronize(s->events[3 * (m - 1) + 2]));
        for (int i = 0; i < m; ++i) {
            float a = 0, b = 0;
            SG_HIP(hipEventElapsedTime(&a, s->events[3 * i], s->events[3 * i + 1]));
            SG_HIP(hipEventElapsedTime(&b, s->events[3 * i + 1], s->events[3 * i + 2]));
            v += a;
            st += b;
        }
        done += m;
    }
    SG_HIP(hipEventRecord(t1, s->stream));
    SG_HIP(hipEventSynchronize(t1));
    float tot = 0;
    SG_HIP(hipEventElapsedTime(&tot, t0, t1));
    if (ms_vel) *ms_vel = v;
    if (ms_str) *ms_str = st;
    if (ms_total) *ms_total = tot;
    return SG_OK;
}

int sg_set_graph(sg_solver *s, int enable)
{
    if (int rc = check_solver(s)) return rc;
    s->use_graph = enable != 0;
    if (!s->use_graph) drop_graph(s);
    s->graph_valid = false;
    return SG_OK;
}

int sg_set_tile(sg_solver *s, const sg_tile *tile)
{
    if (int rc = check_solver(s)) return rc;
    if (!tile) return fail(SG_ERR_INVALID, "sg_set_tile: NULL tile");
    if (!tile_shape_ok(tile->tile_y, tile->rows_per_thread) ||
        !tile_shape_ok(tile->pml_tile_y, tile->pml_rows_per_thread))
        return fail(SG_ERR_INVALID,
                    "sg_set_tile: (tile_y, rows_per_thread) must be (4,1), (8,1), (8,2), (16,1) or (16,2)");
    if (tile->z_chunk < 0) return fail(SG_ERR_INVALID, "sg_set_tile: z_chunk must be >= 0");
    if (!waves_ok(tile->waves_per_simd)) return fail(SG_ERR_INVALID, "sg_set_tile: waves_per_simd must be 0, 2, 3 or 4");
    if (tile->waves_per_simd != 0 && s->prec != SG_FP32)
        return fail(SG_ERR_UNSUPPORTED, "sg_set_tile: waves_per_simd bounds exist for FP32 storage only");
    s->tile = Tile{tile->tile_y, tile->rows_per_thread, tile->pml_tile_y, tile->pml_rows_per_thread, tile->z_chunk,
                   tile->waves_per_simd};
    s->graph_valid = false;
    s->wl_valid = false;
    return SG_OK;
}

int sg_get_tile(const sg_solver *s, sg_tile *tile)
{
    if (int rc = check_solver(s)) return rc;
    if (!tile) return fail(SG_ERR_INVALID, "sg_get_tile: NULL tile");
    *tile = sg_tile{s->tile.ty, s->tile.ny, s->tile.pty, s->tile.pny, s->tile.zchunk, s->tile.waves};
    return SG_OK;
}

int sg_autotune(sg_solver *s, int steps, sg_tile *best_tile, double *best_ms)
{
    if (int rc = check_solver(s)) return rc;
    if (steps < 1) return fail(SG_ERR_INVALID, "sg_autotune: steps < 1");
    if (s->cfg.kernel != SG_KERNEL_TILED) return fail(SG_ERR_STATE, "sg_autotune: the solver uses the per-point kernels");
    // Three passes, each candidate timed over `steps` steps from the current
    // state: interior shape x z-chunk length with the PML shape fixed, then
    // the PML shape, then the one-launch form with each register bound.
    // The state is reset afterwards.
    const int uz = s->g.uz;
    const int shapes[][2] = {{4, 1}, {8, 1}, {16, 1}, {8, 2}, {16, 2}};
    const int zcs[] = {0, 8, 10, 12, 16, 20, 24, 32, 48};
    double best = 1e300;
    Tile bt = s->tile;
    const int ns0 = s->nstreams;
    int bns = ns0;
    const bool graph = s->use_graph;
    s->use_graph = false;
    auto trial = [&](const Tile &t, int ns) -> int {
        s->tile = t;
        s->nstreams = ns;
        s->wl_valid = false;
        double v = 0, st = 0, tot = 0;
        if (int rc = sg_step_timed(s, steps, &v, &st, &tot)) return rc;
        if (v + st < best) {
            best = v + st;
            bt = t;
            bns = ns;
        }
        return SG_OK;
    };
    int rc = SG_OK;
    const Tile start = s->tile;
    for (const auto &sh : shapes)
        for (int zc : zcs)
            if (!rc && zc <= uz) rc = trial(Tile{sh[0], sh[1], start.pty, start.pny, zc, start.waves}, ns0);
    const Tile mid = bt;
    for (const auto &sh : shapes)
        if (!rc && (sh[0] != mid.pty || sh[1] != mid.pny))
            rc = trial(Tile{mid.ty, mid.ny, sh[0], sh[1], mid.zchunk, mid.waves}, ns0);
    // the one-launch-per-sweep form on one stream, each register bound
    // (separate instantiations; FP32 storage)
    const int wvs[] = {0, 2, 3, 4};
    for (int w : wvs)
        if (!rc && (w == 0 || s->prec == SG_FP32) && s->comp_mask == 0) rc = trial(Tile{4, 1, 4, 1, bt.zchunk, w}, 1);
    s->use_graph = graph;
    s->tile = bt;
    s->nstreams = bns;
    s->graph_valid = false;
    s->wl_valid = false;
    if (rc) return rc;
    if (best_tile) *best_tile = sg_tile{bt.ty, bt.ny, bt.pty, bt.pny, bt.zchunk, bt.waves};
    if (best_ms) *best_ms = best / steps;
    return sg_reset(s);
}

int sg_sweep_bytes(const sg_solver *s, double *vbytes, double *sbytes)
{
    if (!s) return fail(SG_ERR_INVALID, "solver handle is NULL");
    const Geom &g = s->g;
    const double sT = (double)storage_size(s->prec), sC = (double)compute_size(s->prec);
    const double P2 = 2.0 * g.pml;
    const double cnt[3][2] = {{g.ux - P2, P2}, {g.uy - P2, P2}, {g.uz - P2, P2}};
    double vb = 0, sb = 0;
    for (int m = 0; m < 8; ++m) {
        const int dx = m & 1, dy = (m >> 1) & 1, dz = (m >> 2) & 1;
        const double pts = cnt[0][dx] * cnt[1][dy] * cnt[2][dz];
        if (pts <= 0) continue;
        const int nd = dx + dy + dz;
        // velocity: 6 stresses + buoyancy read, 3 components with 3 axes each
        const double v = 7 * sT + 3 * comp_bytes(3, nd, sT, sC);
        // stress: 3 velocities + lambda + mu read, 3 normal (3 axes), 3 shear (2 axes)
        const double st = 5 * sT + 3 * comp_bytes(3, nd, sT, sC) + comp_bytes(2, dx + dy, sT, sC) +
                          comp_bytes(2, dx + dz, sT, sC) + comp_bytes(2, dy + dz, sT, sC);
        vb += pts * v;
        sb += pts * st;
    }
    if (vbytes) *vbytes = vb;
    if (sbytes) *sbytes = sb;
    return SG_OK;
}

int sg_copy_bandwidth(long long bytes, int reps, double *gbps)
{
    if (bytes < 1024 || reps < 1 || !gbps) return fail(SG_ERR_INVALID, "sg_copy_bandwidth: bad arguments");
    bytes &= ~15LL;
    float *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SG_OK;
    auto cleanup = [&]() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    };
    hipError_t e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_fill_zero(a, bytes, st);
    if (e == hipSuccess) e = launch_copy(a, b, bytes / 4, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy(a, b, bytes / 4, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess)
        rc = fail(SG_ERR_HIP, std::string("sg_copy_bandwidth: ") + hipGetErrorString(e));
    else
        *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    cleanup();
    return rc;
}

}  // extern "C"

// ---- accessors for the batch module ----------------------------------------
namespace sg {
const void *solver_shot(const sg_solver *s) { return s->shot; }
int solver_prec(const sg_solver *s) { return s->prec; }
int solver_kernel(const sg_solver *s) { return s->cfg.kernel; }
int solver_device(const sg_solver *s) { return s->device; }
hipStream_t solver_stream(const sg_solver *s) { return s->stream; }
const Geom &solver_geom(const sg_solver *s) { return s->g; }
void solver_add_steps(sg_solver *s, long long n) { s->steps_done += n; }
void solver_batch_ref(sg_solver *s, hipStream_t batch_stream, int d)
{
    auto &v = s->batch_streams;
    if (d > 0) v.push_back(batch_stream);
    else {
        auto it = std::find(v.begin(), v.end(), batch_stream);
        if (it != v.end()) v.erase(it);
    }
}
void solver_aux(const sg_solver *s, AuxDesc *a)
{
    a->amp_d = s->amp_d;
    a->amp_acc = static_cast<const float *>(s->amp_acc);
    a->namp = s->namp;
    a->inject = inject_mode(s);
    a->d_step = s->d_step;
    a->saved = static_cast<float *>(s->saved);
    a->st_idx = s->st_idx;
    a->st_comp = s->st_comp;
    a->nst = s->nst;
    a->ncomp = s->ncomp;
    a->interval = s->interval;
    a->cap = s->st_cap;
    a->st_out = static_cast<float *>(s->st_out);
}
}  // namespace sg
