// sg_internal.hpp -- device-side views and launch entry points shared by the
// HIP kernels (sg_kernels.hip, sg_tiled.hip) and the solver (sg_solver.cpp).
//
// Data layout in HBM (one shot):
//   * 9 wavefields and 3 model arrays, x-fastest over the total grid
//     (index i + j*sy + k*sz, sy = nx, sz = nx*ny), like the reference CUDA
//     INDEX macro (3DGPUwithPML_1DArraySingleGPU_concise/ElasticMacro.h:22).
//   * PML split components only where a split exists: a slab store of three
//     sections, each starting on a 64-element boundary: the 2P z-slab planes
//     (rows of pitch nx, indexed by absolute i and j, so split rows line up
//     with field rows), the 2P y-slab rows of each z-interior plane (pitch
//     nx), and the x-slabs of the y/z interior with the low and high slab
//     stored apart (P values per row, rows of one slab contiguous).  Every
//     split array spans the whole store; a split is touched only at points
//     where its axis damps.
//   * per-axis damping factor tables A(i) = (1 - dt d/2)/(1 + dt d/2) and
//     B(i) = 1/(1 + dt d/2), formed in double on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {

constexpr int HALO = 5;  // half stencil width: updated points are [5, n-5)

enum Field { VX = 0, VY, VZ, SXX, SYY, SZZ, SXY, SXZ, SYZ, NFIELD };

// Stress split slots: normal components carry x,y,z splits, shear components
// only the two axes that contribute (the reference's third shear split is
// identically zero: 3DCPUwithPML.../ElasticWaveEquation3D.cpp:233,237,241).
enum StressSplit {
    XX_X = 0, XX_Y, XX_Z, YY_X, YY_Y, YY_Z, ZZ_X, ZZ_Y, ZZ_Z,
    XY_X, XY_Y, XZ_X, XZ_Z, YZ_Y, YZ_Z, NSSPLIT
};

struct Geom {
    int nx, ny, nz;      // total extents
    int pml;             // PML thickness
    int ux, uy, uz;      // updated extents (n - 10)
    long long sy, sz;    // strides in elements
    long long zsec;      // start of the y-slab section of the split store
    long long ysec;      // start of the x-slab section
    long long xhalf;     // size of one x slab (low, then high)
    long long nslots;    // total split slots
};

inline long long align64(long long n) { return (n + 63) / 64 * 64; }

// section sizes of the split store
inline void split_layout(Geom &g)
{
    const long long P = g.pml;
    g.zsec = align64(2 * P * g.ny * (long long)g.nx);
    g.ysec = g.zsec + align64((g.uz - 2 * P) * 2 * P * (long long)g.nx);
    g.xhalf = align64((g.uz - 2 * P) * (g.uy - 2 * P) * P);
    g.nslots = P > 0 ? g.ysec + 2 * g.xhalf : 0;
}

__host__ __device__ inline bool in_pml(int i, int n, int pml)
{
    return i < HALO + pml || i >= n - HALO - pml;
}

// Split-store slot of an updated PML point (i, j, k) (absolute indices).
__host__ __device__ inline long long split_slot(const Geom &g, int i, int j, int k)
{
    const int P = g.pml;
    const int kk = k - HALO;
    if (kk < P || kk >= g.uz - P) {
        const int kz = kk < P ? kk : kk - (g.uz - 2 * P);
        return ((long long)kz * g.ny + j) * g.nx + i;
    }
    const int jj = j - HALO;
    if (jj < P || jj >= g.uy - P) {
        const int jy = jj < P ? jj : jj - (g.uy - 2 * P);
        return g.zsec + ((long long)(kk - P) * (2 * P) + jy) * g.nx + i;
    }
    const int ii = i - HALO;
    const bool high = ii >= P;
    const int ix = high ? ii - (g.ux - P) : ii;
    return g.ysec + (high ? g.xhalf : 0) + ((long long)(kk - P) * (g.uy - 2 * P) + (jj - P)) * P + ix;
}

// Everything one sweep needs about one shot, passed by value to kernels (or
// read from device memory by the batched kernels).
template <class T, class C>
struct Shot {
    Geom g;
    T *f[NFIELD];
    const T *b, *lam, *mu;
    C *sv[9];           // velocity splits: component*3 + axis
    C *ss[NSSPLIT];     // stress splits
    const C *dA[3];     // damping factor tables per axis
    const C *dB[3];
    C cx[5], cy[5], cz[5];  // c_n / h per axis
    C dt;
    long long src;      // linear index of the source point, -1 if none
    const T *src_saved; // TRANSIENT mode: sxx/syy/szz at src before injection
    // storage scales (FP16/BF16: stored = value * scale); 1 otherwise
    float fscale[NFIELD];
    float finv[NFIELD];
    float bscale, lscale, mscale;  // model storage scales (value = stored / scale)
    int round_mode;                // 16-bit storage: 0 nearest even, 1 stochastic
    const int *d_step;             // steps started (advanced by the source kernel)
    // 16-bit storage, compensated time update: per field (or null) the
    // rounding residual of the last store, stored as residual * cscale
    T *comp[NFIELD];
    float cscale[NFIELD], cinv[NFIELD];
    // per-cell damping (the CUDA reference's form, d = g(axis index) *
    // sqrt(mu*b) of the cell): the local S velocity per split slot, or null;
    // when set, dA holds dt*g/2 per axis and A, B are formed per point
    const C *vs;
};

// residual scale of compensation arrays relative to the field scale: the
// residual of a 16-bit store is below half an ulp, 2^-11 of the value
constexpr float COMP_GAIN = 1024.0f;

// Box shapes of the tiled sweeps (block = 64 x ty/ny threads).
struct Tile {
    int ty, ny;    // interior boxes: y rows per box, y rows per thread
    int pty, pny;  // PML boxes
    int zchunk;    // z planes per box (0 = automatic)
    int waves;     // register bound of the one-launch form (waves per SIMD, 0 none)
};

// (4,1) boxes throughout: the five lists of a sweep can run as one launch
inline bool tile_fusable(const Tile &t) { return t.ty == 4 && t.ny == 1 && t.pty == 4 && t.pny == 1; }
inline bool waves_ok(int w) { return w == 0 || w == 2 || w == 3 || w == 4; }

inline bool tile_shape_ok(int ty, int ny)
{
    return (ty == 4 && ny == 1) || (ty == 8 && (ny == 1 || ny == 2)) || (ty == 16 && (ny == 1 || ny == 2));
}

// Work item of the batched launches: one (shot, x tile, y tile, z chunk).
struct WorkItem {
    int shot;    // -1: padding entry
    int i0, i1;  // x range [i0, i1), at most 64 wide
    int j0, j1;  // y range [j0, j1), at most TY high
    int k0, k1;  // z range [k0, k1)
};

// Per-shot state of the small kernels (source, recording) for batched steps.
struct AuxDesc {
    const double *amp_d;
    const float *amp_acc;
    int namp, inject;  // INJECT_*
    int *d_step;
    float *saved;
    const long long *st_idx;
    const int *st_comp;
    int nst, ncomp, interval;
    long long cap;
    float *st_out;
};

}  // namespace sg

// ---- launch entry points (defined in the .hip files) -----------------------
// prec: 0 fp32, 1 fp64, 2 fp16, 3 bf16.  `shot` points to a host Shot<T,C>
// of the matching instantiation.
namespace sg {
// per-point sweeps (host Shot passed by value)
hipError_t launch_velocity_simple(int prec, const void *shot, hipStream_t st);
hipError_t launch_stress_simple(int prec, const void *shot, hipStream_t st);
// tiled sweeps over a work-list: which 0 = velocity, 1 = stress; shots and
// items in device memory; dm = damping mask of the list's boxes (0 interior,
// 1/2/4 x/y/z PML face, 7 edges and corners, DM_INTERIOR_LDS interior boxes
// whose staging rows are loaded straight into LDS, FP32 only); comp != 0
// selects the compensated 16-bit instantiations; grid 0 = one block per item
constexpr int DM_INTERIOR_LDS = 8;
hipError_t launch_sweep(int which, int prec, int ty, int ny, int dm, int comp, const void *shots_dev,
                        const WorkItem *items, int nitems, int grid, hipStream_t st);
hipError_t sweep_fp32(int which, int ty, int ny, int dm, int comp, const void *shots, const WorkItem *items, int nitems,
                      int grid, hipStream_t st);
hipError_t sweep_fp64(int which, int ty, int ny, int dm, int comp, const void *shots, const WorkItem *items, int nitems,
                      int grid, hipStream_t st);
hipError_t sweep_fp16(int which, int ty, int ny, int dm, int comp, const void *shots, const WorkItem *items, int nitems,
                      int grid, hipStream_t st);
hipError_t sweep_bf16(int which, int ty, int ny, int dm, int comp, const void *shots, const WorkItem *items, int nitems,
                      int grid, hipStream_t st);

// all work-lists of a sweep in one launch, (4,1) boxes (off: NLIST+1 list
// offsets into items); waves: register bound of the launch in waves per
// SIMD (0 none; 2, 3, 4 for FP32 storage)
hipError_t sweep_fused_fp32(int which, int comp, int waves, const void *shots, const WorkItem *items, const int *off, int grid,
                            hipStream_t st);
hipError_t sweep_fused_fp64(int which, int comp, int waves, const void *shots, const WorkItem *items, const int *off, int grid,
                            hipStream_t st);
hipError_t sweep_fused_fp16(int which, int comp, int waves, const void *shots, const WorkItem *items, const int *off, int grid,
                            hipStream_t st);
hipError_t sweep_fused_bf16(int which, int comp, int waves, const void *shots, const WorkItem *items, const int *off, int grid,
                            hipStream_t st);
hipError_t launch_sweep_fused(int which, int prec, int comp, int waves, const void *shots_dev, const WorkItem *items,
                              const int *off, int grid, hipStream_t st);

// work-lists: interior, x face, y face, z face, edges/corners
constexpr int NLIST = 5;
constexpr int LIST_DM[NLIST] = {0, 1, 2, 4, 7};
// Source injection modes: TRANSIENT adds pow(dt,2)*s in double and the stress
// sweep restarts from the saved sigma at the source; ACCUMULATE adds
// dt*dt*s in the compute type and keeps it; TRANSIENT_F adds dt*dt*s and
// restarts from the saved sigma.
enum { INJECT_TRANSIENT = 0, INJECT_ACCUMULATE = 1, INJECT_TRANSIENT_F = 2 };

hipError_t launch_inject(int prec, const void *shot, const double *amp_d, const void *amp_acc,
                         int namp, int mode, int *d_step, void *saved, hipStream_t st);
hipError_t launch_gather(int prec, const void *shot, const long long *idx, const int *comp, int nst,
                         int ncomp, int interval, long long capacity, void *out, const int *d_step,
                         hipStream_t st);
// FP64 energies (kinetic, sum of stresses squared, strain energy) into
// out3[0..2]; scratch holds ENERGY_PART doubles per updated z plane (nz - 10
// planes); out3 may lie inside the caller's scratch beyond those
constexpr int ENERGY_PART = 4;
hipError_t launch_energy(int prec, const void *shot, double *scratch, double *out3, hipStream_t st);
// local S velocity sqrt(mu*b) at every split slot (per-cell damping)
hipError_t launch_fill_vs(int prec, const void *shot, void *vs, hipStream_t st);
// largest |value| per wavefield (float bits, atomicMax into out9[9], which
// the caller zeroes)
hipError_t launch_maxabs(int prec, const void *shot, unsigned *out9, hipStream_t st);
hipError_t launch_copy(const float *src, float *dst, long long n, hipStream_t st);
hipError_t launch_fill_zero(void *p, long long bytes, hipStream_t st);
hipError_t launch_inject_batch(const void *shots_dev, const AuxDesc *aux, int nshots, hipStream_t st);
hipError_t launch_gather_batch(const void *shots_dev, const AuxDesc *aux, int nshots, int max_n, hipStream_t st);
}  // namespace sg
