// sg_batch.cpp -- many independent shots advanced together.
//
// The boxes of all shots form one work-list per damping mask (interior, x,
// y, z faces, edges and corners), stored back to back, so small and uneven
// shots share the GPU instead of running one under-filled launch each.
// With (4,1) boxes throughout each half-step is ONE launch over all of them:
// a grid-stride kernel whose blocks pick each item's damping-mask code, its
// registers bounded to a chosen number of waves per SIMD, the grid one block
// per item or a persistent grid.  Other box shapes run one launch per list
// (the default, 16-row interior boxes: measured faster at the 32-shot size,
// since one launch carries the registers of the edge and corner code), the
// PML lists on two side streams.  Source injection and station recording of
// all shots are one launch each.  A step is captured in a hipGraph and
// replayed; sg_batch_autotune picks the form, z-chunk, bound and grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/sgfd.h"
#include "sg_error.hpp"
#include "sg_internal.hpp"
#include "sg_worklist.hpp"

namespace sg {
const void *solver_shot(const sg_solver *s);
int solver_prec(const sg_solver *s);
int solver_kernel(const sg_solver *s);
hipStream_t solver_stream(const sg_solver *s);
const Geom &solver_geom(const sg_solver *s);
void solver_aux(const sg_solver *s, AuxDesc *a);
void solver_add_steps(sg_solver *s, long long n);
void solver_batch_ref(sg_solver *s, hipStream_t batch_stream, int d);
int solver_device(const sg_solver *s);
}  // namespace sg

using namespace sg;

struct sg_batch {
    std::vector<sg_solver *> shots;
    int device = 0;
    hipStream_t stream = nullptr;
    // PML work-lists run beside the interior list: x faces on side[0], the
    // other PML lists on side[1] (fork/join with events)
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
    void *d_shots = nullptr;
    AuxDesc *d_aux = nullptr;
    WorkItem *d_items = nullptr;
    int off[NLIST + 1] = {0, 0, 0, 0, 0, 0};  // work-lists back to back (each padded to a multiple of 8)
    int grid = 0;
    Tile tile{16, 1, 4, 1, 24, 0};
    int max_gather = 0;
    // the shot and auxiliary descriptors as last uploaded to d_shots/d_aux
    std::vector<unsigned char> shot_host;
    std::vector<AuxDesc> aux_host;
    bool use_graph = true;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

void drop_graph(sg_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

// Work-lists over all shots (one per damping mask, each interleaved for the
// XCD round-robin), stored back to back.
int build_items(sg_batch *b)
{
    WorkLists wl;
    for (size_t si = 0; si < b->shots.size(); ++si)
        build_worklists(solver_geom(b->shots[si]), (int)si, b->tile.ty, b->tile.pty, b->tile.zchunk, wl);
    const std::vector<WorkItem> all = pack_worklists(wl, b->off);
    if (b->d_items) SG_HIP(hipFree(b->d_items));
    b->d_items = nullptr;
    SG_HIP(hipMalloc(&b->d_items, all.size() * sizeof(WorkItem)));
    SG_HIP(hipMemcpy(b->d_items, all.data(), all.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
    b->graph_valid = false;
    return SG_OK;
}

// Upload the shots' descriptors when they differ from the last upload (or
// when forced).  They hold device pointers that sg_set_stations and
// sg_set_source reallocate and sg_set_damping_local adds, so every step
// checks them; kernels still in flight read the old copy, hence the wait.
// The captured graph is dropped: the gather launch depends on the stations.
int sync_descriptors(sg_batch *b, bool force)
{
    const size_t n = b->shots.size(), ss = sizeof(Shot<float, float>);
    std::vector<unsigned char> sh(n * ss);
    std::vector<AuxDesc> aux(n);
    std::memset(static_cast<void *>(aux.data()), 0, n * sizeof(AuxDesc));  // padding compares equal
    int max_gather = 0;
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(sh.data() + i * ss, solver_shot(b->shots[i]), ss);
        solver_aux(b->shots[i], &aux[i]);
        max_gather = std::max(max_gather, aux[i].nst * aux[i].ncomp);
    }
    if (!force && sh == b->shot_host &&
        std::memcmp(static_cast<const void *>(aux.data()), static_cast<const void *>(b->aux_host.data()),
                    n * sizeof(AuxDesc)) == 0)
        return SG_OK;
    SG_HIP(hipStreamSynchronize(b->stream));
    SG_HIP(hipMemcpy(b->d_shots, sh.data(), sh.size(), hipMemcpyHostToDevice));
    SG_HIP(hipMemcpy(b->d_aux, aux.data(), n * sizeof(AuxDesc), hipMemcpyHostToDevice));
    b->shot_host.swap(sh);
    b->aux_host.swap(aux);
    b->max_gather = max_gather;
    drop_graph(b);
    return SG_OK;
}

int enqueue(sg_batch *b)
{
    const int n = (int)b->shots.size();
    SG_HIP(launch_inject_batch(b->d_shots, b->d_aux, n, b->stream));
    const Tile &t = b->tile;
    if (tile_fusable(t)) {
        // one launch per half-step over every shot's work-lists
        for (int which = 0; which < 2; ++which)
            SG_HIP(launch_sweep_fused(which, SG_FP32, 0, t.waves, b->d_shots, b->d_items, b->off, b->grid,
                                      b->stream));
        SG_HIP(launch_gather_batch(b->d_shots, b->d_aux, n, b->max_gather, b->stream));
        return SG_OK;
    }
    for (int which = 0; which < 2; ++which) {
        SG_HIP(hipEventRecord(b->fork, b->stream));
        for (int q = 0; q < 2; ++q) SG_HIP(hipStreamWaitEvent(b->side[q], b->fork, 0));
        for (int l = 0; l < NLIST; ++l) {
            const int cnt = b->off[l + 1] - b->off[l];
            hipStream_t st = l == 0 ? b->stream : (l == 1 ? b->side[0] : b->side[1]);
            SG_HIP(launch_sweep(which, SG_FP32, l ? t.pty : t.ty, l ? t.pny : t.ny, LIST_DM[l], 0, b->d_shots,
                                b->d_items + b->off[l], cnt, b->grid > 0 ? std::min(b->grid, cnt) : 0, st));
        }
        for (int q = 0; q < 2; ++q) {
            SG_HIP(hipEventRecord(b->join[q], b->side[q]));
            SG_HIP(hipStreamWaitEvent(b->stream, b->join[q], 0));
        }
    }
    SG_HIP(launch_gather_batch(b->d_shots, b->d_aux, n, b->max_gather, b->stream));
    return SG_OK;
}

void free_batch(sg_batch *b)
{
    drop_graph(b);
    if (b->d_shots) (void)hipFree(b->d_shots);
    if (b->d_aux) (void)hipFree(b->d_aux);
    if (b->d_items) (void)hipFree(b->d_items);
    for (int q = 0; q < 2; ++q) {
        if (b->side[q]) (void)hipStreamDestroy(b->side[q]);
        if (b->join[q]) (void)hipEventDestroy(b->join[q]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

}  // namespace

extern "C" {

int sg_batch_create(sg_solver *const *shots, int n, sg_batch **out)
{
    if (!out) return fail(SG_ERR_INVALID, "sg_batch_create: out is NULL");
    *out = nullptr;
    if (!shots || n < 1) return fail(SG_ERR_INVALID, "sg_batch_create: need at least one shot");
    if (!shots[0]) return fail(SG_ERR_INVALID, "sg_batch_create: NULL shot");
    const int dev = solver_device(shots[0]);
    for (int i = 0; i < n; ++i) {
        if (!shots[i]) return fail(SG_ERR_INVALID, "sg_batch_create: NULL shot");
        if (solver_prec(shots[i]) != SG_FP32) return fail(SG_ERR_UNSUPPORTED, "sg_batch_create: batching is FP32 only");
        if (solver_device(shots[i]) != dev) return fail(SG_ERR_INVALID, "sg_batch_create: shots on different devices");
        // the batched sweeps are the tiled ones, whose limits sg_create
        // checks only for tiled solvers
        if (solver_kernel(shots[i]) != SG_KERNEL_TILED)
            return fail(SG_ERR_UNSUPPORTED, "sg_batch_create: shots must use the tiled sweeps (SG_KERNEL_TILED)");
        for (int j = 0; j < i; ++j)
            if (shots[j] == shots[i])
                return fail(SG_ERR_INVALID, "sg_batch_create: shot " + std::to_string(i) + " is listed twice");
    }
    sg_batch *b = new (std::nothrow) sg_batch();
    if (!b) return fail(SG_ERR_NOMEM, "sg_batch_create: host allocation failed");
    b->shots.assign(shots, shots + n);
    b->device = dev;
    auto bail = [&](int rc) {
        free_batch(b);
        delete b;
        return rc;
    };
    hipError_t e = hipSetDevice(dev);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = hipStreamSynchronize(solver_stream(shots[i]));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int q = 0; q < 2 && e == hipSuccess; ++q) {
        e = hipStreamCreateWithFlags(&b->side[q], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[q], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&b->d_shots, (size_t)n * sizeof(Shot<float, float>));
    if (e == hipSuccess) e = hipMalloc(&b->d_aux, n * sizeof(AuxDesc));
    if (e != hipSuccess)
        return bail(fail(e == hipErrorOutOfMemory ? SG_ERR_NOMEM : SG_ERR_HIP,
                         std::string("sg_batch_create: ") + hipGetErrorString(e)));
    if (int rc = sync_descriptors(b, true)) return bail(rc);
    if (int rc = build_items(b)) return bail(rc);
    for (auto *s : b->shots) solver_batch_ref(s, b->stream, +1);
    *out = b;
    return SG_OK;
}

int sg_batch_set_launch(sg_batch *b, const sg_tile *tile, int grid, int use_graph)
{
    if (!b) return fail(SG_ERR_INVALID, "sg_batch_set_launch: NULL batch");
    if (!tile) return fail(SG_ERR_INVALID, "sg_batch_set_launch: NULL tile");
    if (!tile_shape_ok(tile->tile_y, tile->rows_per_thread) ||
        !tile_shape_ok(tile->pml_tile_y, tile->pml_rows_per_thread))
        return fail(SG_ERR_INVALID,
                    "sg_batch_set_launch: (tile_y, rows_per_thread) must be (4,1), (8,1), (8,2), (16,1) or (16,2)");
    if (tile->z_chunk < 1 || grid < 0) return fail(SG_ERR_INVALID, "sg_batch_set_launch: bad z_chunk or grid");
    if (!waves_ok(tile->waves_per_simd))
        return fail(SG_ERR_INVALID, "sg_batch_set_launch: waves_per_simd must be 0, 2, 3 or 4");
    SG_HIP(hipSetDevice(b->device));
    SG_HIP(hipStreamSynchronize(b->stream));
    b->tile = Tile{tile->tile_y, tile->rows_per_thread, tile->pml_tile_y, tile->pml_rows_per_thread, tile->z_chunk,
                   tile->waves_per_simd};
    b->grid = grid;
    b->use_graph = use_graph != 0;
    drop_graph(b);
    return build_items(b);
}

int sg_batch_step(sg_batch *b, int nsteps)
{
    if (!b) return fail(SG_ERR_INVALID, "sg_batch_step: NULL batch");
    if (nsteps < 0) return fail(SG_ERR_INVALID, "sg_batch_step: nsteps < 0");
    SG_HIP(hipSetDevice(b->device));
    // work a shot queued on its own stream comes first
    for (auto *s : b->shots) SG_HIP(hipStreamSynchronize(solver_stream(s)));
    if (int rc = sync_descriptors(b, false)) return rc;
    if (b->use_graph) {
        if (!b->graph_valid) {
            drop_graph(b);
            SG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(b);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            SG_HIP(e);
            b->graph = g;
            SG_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            b->graph_valid = true;
        }
        for (int i = 0; i < nsteps; ++i) SG_HIP(hipGraphLaunch(b->gexec, b->stream));
    } else {
        for (int i = 0; i < nsteps; ++i)
            if (int rc = enqueue(b)) return rc;
    }
    for (auto *s : b->shots) solver_add_steps(s, nsteps);
    return SG_OK;
}

int sg_batch_step_timed(sg_batch *b, int nsteps, double *ms)
{
    if (!b) return fail(SG_ERR_INVALID, "sg_batch_step_timed: NULL batch");
    SG_HIP(hipSetDevice(b->device));
    hipEvent_t e0, e1;
    SG_HIP(hipEventCreate(&e0));
    SG_HIP(hipEventCreate(&e1));
    SG_HIP(hipEventRecord(e0, b->stream));
    int rc = sg_batch_step(b, nsteps);
    SG_HIP(hipEventRecord(e1, b->stream));
    SG_HIP(hipEventSynchronize(e1));
    float t = 0;
    SG_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = t;
    return rc;
}

int sg_batch_autotune(sg_batch *b, int steps, sg_tile *best_tile, int *best_grid, double *best_ms)
{
    if (!b) return fail(SG_ERR_INVALID, "sg_batch_autotune: NULL batch");
    if (steps < 1) return fail(SG_ERR_INVALID, "sg_batch_autotune: steps < 1");
    SG_HIP(hipSetDevice(b->device));
    int cus = 0;
    SG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    const Tile t0 = b->tile;
    const int g0 = b->grid;
    const bool graph0 = b->use_graph;
    Tile bt = t0;
    int bg = g0;
    double best = 1e300;
    auto trial = [&](const Tile &t, int grid) -> int {
        const sg_tile st{t.ty, t.ny, t.pty, t.pny, t.zchunk, t.waves};
        if (int rc = sg_batch_set_launch(b, &st, grid, graph0 ? 1 : 0)) return rc;
        if (int rc = sg_batch_step(b, 1)) return rc;  // capture outside the timing
        double ms = 0;
        if (int rc = sg_batch_step_timed(b, steps, &ms)) return rc;
        if (ms < best) {
            best = ms;
            bt = t;
            bg = grid;
        }
        return SG_OK;
    };
    int rc = SG_OK;
    // the one-launch form: z-chunk x register bound x grid (one block per
    // item, or 2 or 4 persistent blocks per CU)
    const int zcs[] = {16, 24, 32};
    const int wvs[] = {0, 2, 3, 4};
    const int grids[] = {0, 2 * cus, 4 * cus};
    for (int zc : zcs)
        for (int w : wvs)
            for (int g : grids)
                if (!rc) rc = trial(Tile{4, 1, 4, 1, zc, w}, g);
    // one launch per list with taller interior boxes
    if (!rc) rc = trial(Tile{8, 1, 4, 1, 24, 0}, 0);
    if (!rc) rc = trial(Tile{16, 1, 4, 1, 24, 0}, 0);
    const sg_tile bst{bt.ty, bt.ny, bt.pty, bt.pny, bt.zchunk, bt.waves};
    const sg_tile st0{t0.ty, t0.ny, t0.pty, t0.pny, t0.zchunk, t0.waves};
    const int rc2 = sg_batch_set_launch(b, rc ? &st0 : &bst, rc ? g0 : bg, graph0 ? 1 : 0);
    if (rc) return rc;
    if (rc2) return rc2;
    SG_HIP(hipStreamSynchronize(b->stream));
    for (auto *s : b->shots)
        if (int r = sg_reset(s)) return r;
    if (best_tile) *best_tile = bst;
    if (best_grid) *best_grid = bg;
    if (best_ms) *best_ms = best / steps;
    return SG_OK;
}

int sg_batch_synchronize(sg_batch *b)
{
    if (!b) return fail(SG_ERR_INVALID, "sg_batch_synchronize: NULL batch");
    SG_HIP(hipSetDevice(b->device));
    SG_HIP(hipStreamSynchronize(b->stream));
    return SG_OK;
}

int sg_batch_destroy(sg_batch *b)
{
    if (!b) return SG_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto *s : b->shots) solver_batch_ref(s, b->stream, -1);
    free_batch(b);
    delete b;
    return SG_OK;
}

}  // extern "C"
