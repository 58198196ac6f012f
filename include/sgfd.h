/*
 * sgfd.h -- C ABI of the MI355X (gfx950) staggered-grid elastic solver.
 *
 * One solver = one 3-D velocity-stress model on one GPU: 10th-order
 * staggered-grid derivatives, leapfrog time stepping, split-field PML, an
 * explosive point source on sxx/syy/szz and station recording.  All device
 * memory is owned by the solver; every call returns an sg_status and never
 * exits the process.  A solver owns one HIP stream; different solvers are
 * independent (no global device state), so several may coexist.
 *
 * Reference interfaces replaced (paths relative to the reference root,
 * CONCISE = 3DGPUwithPML_1DArraySingleGPU_concise,
 * CPUPML  = 3DCPUwithPML_3DArray_GridIncludePMLandOrder):
 *   ElasticWaveEquation ctor  CONCISE/ElasticWaveEquation.h:41-46, .cu:8-45
 *   ElasticWaveEquation3D ctor CPUPML/ElasticWaveEquation3D.h:13-18
 *   run()                     CONCISE/ElasticWaveEquation.cu:140-185
 *   propagate(it, s)          CONCISE/ElasticWaveEquation.cu:194-206
 *   velocityUpdate kernel     CONCISE/ElasticWaveEquation.cu:237-517
 *   stressUpdate kernel       CONCISE/ElasticWaveEquation.cu:519-820
 *   locateSource kernel       CONCISE/ElasticWaveEquation.cu:209-224
 *   recordStationData kernel  CONCISE/ElasticWaveEquation.cu:226-235
 *   sourceFunction            CPUPML/ElasticWaveEquation3D.cpp:114-119
 *   calculateDamping          CPUPML/ElasticWaveEquation3D.cpp:253-300
 *
 * Layout conventions: arrays are x-fastest over the TOTAL grid (index
 * i + j*nx + k*nx*ny) like CONCISE/ElasticMacro.h:22, where the total extent
 * includes the 5-cell stencil halo (never updated, kept at zero) and the PML.
 * Updated points: [5, n-5) per axis; PML: the first/last `pml` of those.
 */
#ifndef SGFD_H
#define SGFD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_API_VERSION 2

typedef struct sg_solver sg_solver;
typedef struct sg_batch sg_batch;

/* Status codes (the reference exits via CHECK, common.h:6-16; this returns). */
typedef enum sg_status {
    SG_OK = 0,
    SG_ERR_INVALID = -1,  /* bad argument or configuration */
    SG_ERR_RANGE = -2,    /* index (source, station, field) outside the grid */
    SG_ERR_NOMEM = -3,    /* device or host allocation failed */
    SG_ERR_HIP = -4,      /* a HIP runtime call failed (see sg_last_error) */
    SG_ERR_STATE = -5,    /* call not valid in the current state */
    SG_ERR_UNSUPPORTED = -6,
    SG_ERR_UNSTABLE = -7  /* the stability check stopped stepping (sg_set_stability) */
} sg_status;

typedef enum sg_precision {
    SG_FP32 = 0,  /* float storage, float arithmetic */
    SG_FP64 = 1,  /* double storage, double arithmetic (the error yardstick) */
    SG_FP16 = 2,  /* IEEE half storage with per-field power-of-two scales, float arithmetic */
    SG_BF16 = 3   /* bfloat16 storage with per-field power-of-two scales, float arithmetic */
} sg_precision;

/* Source semantics.  TRANSIENT follows the reference CPU path: the injected
 * dt^2*s(it) is seen by that step's velocity update only, because the stress
 * update rebuilds sigma from its splits (CPUPML/ElasticWaveEquation3D.cpp:
 * 121-130 then 215/220/225).  ACCUMULATE follows the reference CUDA path,
 * whose unsplit interior keeps every injection (CONCISE/.cu:216-218, 256-258),
 * with the float add dt*dt*s of locateSource. */
typedef enum sg_source_mode { SG_SOURCE_TRANSIENT = 0, SG_SOURCE_ACCUMULATE = 1 } sg_source_mode;

/* Field ids; also the bit positions of station component masks. */
typedef enum sg_field {
    SG_VX = 0, SG_VY, SG_VZ, SG_SXX, SG_SYY, SG_SZZ, SG_SXY, SG_SXZ, SG_SYZ, SG_NFIELDS
} sg_field;

/* Kernel variants for the two half-step sweeps. */
typedef enum sg_kernel { SG_KERNEL_TILED = 0, SG_KERNEL_SIMPLE = 1 } sg_kernel;

/* Rounding of 16-bit wavefield stores (FP16/BF16 solvers; FP32/FP64 ignore
 * it): nearest even, or stochastic (unbiased; the random bits come from a
 * hash of point, step and field, so results are reproducible and independent
 * of launch shape).  No counterpart in the reference (FP32 only). */
typedef enum sg_rounding { SG_ROUND_NEAREST = 0, SG_ROUND_STOCHASTIC = 1 } sg_rounding;

/* Damping profile formulas for sg_damping_profile(). */
typedef enum sg_damping_kind {
    SG_DAMP_REFERENCE = 0, /* log10(1/R)*tau*V/(P*h)*(dist/P)^2  (CPUPML/ElasticMacro.h:21-23) */
    SG_DAMP_QUADRATIC = 1, /* -3*V*ln(R)/(2*P*h)*(dist/P)^2 */
    SG_DAMP_CUDA = 2       /* tau*V/h*(8/15 - 0.03*P + P^2/1500)*(dist/P)^2
                              (3DGPUwithPML_1DArraySingleGPU_concise/ElasticMacro.h:26-27) */
} sg_damping_kind;

/* Ricker delay conventions for sg_ricker(). */
typedef enum sg_ricker_delay {
    SG_RICKER_CEIL_DELAY = 0, /* t0 = ceil(1/(f0*dt)) steps (CPUPML/ElasticWaveEquation3D.cpp:116) */
    SG_RICKER_NO_DELAY = 1    /* peak at t = 1/f0 */
} sg_ricker_delay;

/* Solver configuration: GRID + pml + Delta + DifferenceCoefficient + timeLength
 * of CONCISE/ElasticWaveEquation.h:41-46 (GPUDim is replaced by `kernel` and
 * the tile choice; the source and stations are set by their own calls). */
typedef struct sg_config {
    int nx, ny, nz;        /* total extents incl. PML and the 5-cell halo */
    int pml;               /* PML thickness in cells (0 = no absorbing layer) */
    float dt, dx, dy, dz;  /* time step [s] and spacings [m] */
    float diff_coef[5];    /* signed staggered coefficients c1..c5 */
    int time_length;       /* steps run by sg_run(); also the station buffer length */
    int precision;         /* sg_precision */
    int source_mode;       /* sg_source_mode */
    int kernel;            /* sg_kernel */
} sg_config;

/* ---- lifetime ------------------------------------------------------------ */

/* Allocates device state (fields, PML splits, tables) on the current HIP
 * device.  Fields start at zero; damping defaults to the reference CPU
 * formula (tau 3.4, V 2000, R 1e-3) until sg_set_damping is called.
 * SG_FP64 runs SG_KERNEL_SIMPLE only (SG_ERR_UNSUPPORTED for the tiled
 * sweeps). */
int sg_create(const sg_config *cfg, sg_solver **out);
/* Frees the solver (NULL: no-op).  SG_ERR_STATE, and nothing freed, while a
 * batch holds it: destroy the batch first. */
int sg_destroy(sg_solver *s);

/* Host-side copy of the model (Medium buoyancy/lambda/mu of CONCISE/
 * ElasticMacro.h:379-384, copied in like run()'s H2D at .cu:145-147).
 * x-fastest over the total grid, n = nx*ny*nz values each.  b and mu must be
 * positive, 3 lambda + 2 mu positive.  A model whose Courant number
 * (sg_courant) exceeds the stability limit is refused with SG_ERR_INVALID
 * (the reference runs it to garbage). */
int sg_set_medium(sg_solver *s, const float *buoyancy, const float *lambda, const float *mu);
/* The model as P velocity, S velocity and density (the reference's main
 * programs form b = 1/rho, lambda and mu from a velocity and a density,
 * CPUPML/main.cpp:34-36, CONCISE/main.cu:53-55): sg_elastic_from_velocity
 * below, then sg_set_medium. */
int sg_set_medium_velocity(sg_solver *s, const float *vp, const float *vs, const float *rho);

/* Per-axis damping d(i) [1/s] for i in [0, n) of that axis (values outside
 * the PML slabs are ignored).  The update factors (1 -/+ dt*d/2) are formed
 * in double on the host. */
int sg_set_damping(sg_solver *s, const double *d_x, const double *d_y, const double *d_z);
/* Per-cell damping of the CUDA reference path
 * (3DGPUwithPML_1DArraySingleGPU_concise/ElasticWaveEquation.cu:285-288 and
 * ElasticMacro.h:21,26-27): d_a(i,j,k) = g_a(index on axis a) * sqrt(mu*b)
 * of the cell, i.e. the damping formula evaluated with the local S velocity.
 * g_a comes from sg_damping_profile(..., v = 1, ...) (the CUDA path:
 * SG_DAMP_REFERENCE with tau 4.3).  Uses the medium set by sg_set_medium
 * (and follows later sg_set_medium calls); sg_set_damping switches back to
 * per-axis damping. */
int sg_set_damping_local(sg_solver *s, const double *g_x, const double *g_y, const double *g_z);

/* Point source at absolute index (x, y, z) (SOURCELOCATION) with wavelet
 * samples s(it), it < n (steps beyond n inject nothing).  The injected
 * amount is pow(dt,2)*s(it) in double (TRANSIENT) or dt*dt*s(it) in the
 * storage precision (ACCUMULATE). */
int sg_set_source(sg_solver *s, int x, int y, int z, const float *wavelet, int n);

/* Overwrites wavelet sample `it` (it < n of sg_set_source): the source value
 * argument of propagate(timeIndex, sourceValue) (CONCISE/.cu:194). */
int sg_set_source_sample(sg_solver *s, int it, float value);

/* Stations (STATIONLOCATION, generalised to many) and the recorded
 * components (bit f of comp_mask records field f).  Recorded after the
 * stress update of every `interval`-th step (interval >= 1), step-major:
 * [record][station][component].  Replaces any previous station set. */
int sg_set_stations(sg_solver *s, const int *xyz, int n, unsigned comp_mask, int interval);

/* FP16/BF16 storage: power-of-two scale per wavefield (stored = value *
 * scale), ids as sg_field; default 1.  Must be set before the first step.
 * The model arrays get their scales from sg_set_medium (max |x| mapped into
 * [2^13, 2^14)).  sg_get_field_scales returns the 9 field scales followed by
 * the buoyancy, lambda and mu scales. */
int sg_set_field_scales(sg_solver *s, const float scales[9]);
/* Rounding mode of 16-bit wavefield stores (sg_rounding).  Fields with a
 * compensation array always round to nearest. */
int sg_set_rounding(sg_solver *s, int mode);
/* Compensated time update for 16-bit storage (FP16/BF16 solvers, tiled
 * sweeps): for each field whose bit (1 << sg_field) is set, a second 16-bit
 * array holds the rounding residual of the last store, added back at the
 * next update of that point, so rounding errors do not accumulate over time
 * steps (stencil neighbours still read the 16-bit value).  Costs two more
 * 16-bit accesses per point per step for each field.  Call before stepping;
 * 0 turns it off.  No counterpart in the reference. */
int sg_set_compensation(sg_solver *s, unsigned field_mask);
int sg_get_field_scales(const sg_solver *s, float out[12]);

/* ---- stability guard ----------------------------------------------------- */

/* Largest normalised Courant number of the current model (0 before
 * sg_set_medium): dt * max(vp) * (|c1|+...+|c5|) * sqrt(1/dx^2 + 1/dy^2 +
 * 1/dz^2), vp = sqrt((lambda + 2 mu) b) per point.  The staggered leapfrog
 * scheme is stable at <= 1 (10th order, dx = dy = dz: vp dt/dx <= 0.4385).
 * No counterpart in the reference, which never checks. */
int sg_courant(const sg_solver *s, double *courant);

/* courant_limit: sg_set_medium refuses a model above it (default 1; 0 = no
 * check).  check_interval: every that many steps of sg_step/sg_run (counted
 * by sg_steps_done) an FP64 energy check runs and sg_step waits for it
 * (default 0 = off); it returns SG_ERR_UNSTABLE, with the remaining steps not
 * run, when a wavefield holds a non-finite value, or when, after the last
 * wavelet sample above 1e-6 of the wavelet's peak, the total energy
 * (kinetic + strain, sg_energy_full) grew by more than max_growth (> 1,
 * default 2) since the previous check.  Batched steps are not checked.  No
 * counterpart in the reference. */
typedef struct sg_stability {
    double courant_limit;
    int check_interval;
    double max_growth;
} sg_stability;
int sg_set_stability(sg_solver *s, const sg_stability *st);
int sg_get_stability(const sg_solver *s, sg_stability *st);

/* ---- stepping --------------------------------------------------------------- */

/* Advances nsteps (source -> velocity -> stress -> record per step), queued on
 * the solver's stream; returns without waiting (except for the stability
 * check, when enabled: sg_set_stability). */
int sg_step(sg_solver *s, int nsteps);
/* time_length steps from the current state, then waits (run(), minus its
 * file output). */
int sg_run(sg_solver *s);
int sg_synchronize(sg_solver *s);
/* Zeroes fields, splits and the step counter (the model is kept). */
int sg_reset(sg_solver *s);
/* Steps completed (queued) so far. */
long long sg_steps_done(const sg_solver *s);

/* ---- results ------------------------------------------------------------- */

/* Copies up to `count` recorded station values (float; waits for the stream).
 * Returns the number of values available in *n_out. */
int sg_get_stations(sg_solver *s, float *out, long long count, long long *n_out);
/* Copies one full field (x-fastest, total grid) to host float / double. */
int sg_get_field(sg_solver *s, int field, float *out);
int sg_get_field_f64(sg_solver *s, int field, double *out);
/* FP64 reductions over the updated region: out[0] = sum rho*|v|^2/2,
 * out[1] = sum of the six stress components squared.  Used to check that
 * the PML drains energy after the source stops. */
int sg_energy(sg_solver *s, double out[2]);
/* The same reduction with a third value: out[2] = the strain energy
 * sum sigma:sigma/(4 mu) - lambda (tr sigma)^2 / (4 mu (3 lambda + 2 mu))
 * (sigma:sigma counts each shear component twice), so out[0] + out[2] is the
 * elastic energy the stability check follows.  Every sum is taken in a fixed
 * order (per z plane, then over the planes): the result does not depend on
 * scheduling. */
int sg_energy_full(sg_solver *s, double out[3]);
/* sg_energy without a library-allocated work buffer: the reduction's partial
 * sums live in a fixed device array of the library (512 KiB per GPU, held
 * until the process ends), so the call allocates nothing.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  Same result, bit
 * for bit, as sg_energy.  Limit: grids of at most 16384 updated z planes
 * (nz - 10); larger grids get SG_ERR_UNSUPPORTED.  All calls on one GPU
 * share the array: calls that overlap on one GPU (from several threads)
 * race and must be serialised by the caller. */
int sg_energy_noalloc(sg_solver *s, double out[2]);
/* Largest |value| of each wavefield (vx vy vz sxx syy szz sxy sxz syz),
 * unscaled.  Run an FP32 solver for a pilot interval and read this to pick
 * the power-of-two storage scales of an FP16/BF16 solver
 * (sg_set_field_scales).  No counterpart in the reference. */
int sg_field_maxabs(sg_solver *s, float out[9]);

/* ---- timing and tuning -------------------------------------------------------- */

/* Like sg_step but brackets each sweep with HIP events on the solver stream;
 * returns the summed milliseconds of the velocity sweeps, the stress sweeps
 * and the whole interval (waits for the stream). */
int sg_step_timed(sg_solver *s, int nsteps, double *ms_velocity, double *ms_stress, double *ms_total);
/* Captures one time step into a hipGraph and replays it (1 = on, 0 = off). */
int sg_set_graph(sg_solver *s, int enable);
/* Box shapes of the tiled sweeps.  A box is at most 64 x tile_y points in x, y
 * and z_chunk planes deep; a block of 64 x tile_y/rows_per_thread threads
 * computes it.  Interior boxes (no damped point) and PML boxes have their
 * own shape.  (tile_y, rows_per_thread) is one of (4,1) (8,1) (8,2) (16,1)
 * (16,2); z_chunk 0 = automatic (the PML thickness P when 8 <= P <= 32,
 * so every box marches as many planes as the whole z PML slabs; else 24,
 * fewer on grids with few columns).  waves_per_simd bounds the registers of
 * the one-launch-per-sweep form (the five work-lists in one kernel, used by
 * a one-stream solver with (4,1) boxes throughout and by batches with (4,1)
 * boxes), so that at least that many waves fit on a SIMD: 0 = no bound, or
 * 2, 3, 4 (FP32 storage only).  Neither shape nor bound changes a result. */
typedef struct sg_tile {
    int tile_y;
    int rows_per_thread;
    int pml_tile_y;
    int pml_rows_per_thread;
    int z_chunk;
    int waves_per_simd;
} sg_tile;
int sg_set_tile(sg_solver *s, const sg_tile *tile);
int sg_get_tile(const sg_solver *s, sg_tile *tile);
/* Streams the work-lists of one sweep are spread over: 1 = one after the
 * other on the solver stream (with (4,1) boxes for the interior and the PML,
 * all five lists as one launch per sweep); 2 = PML lists beside the interior
 * list; 3 = x faces and the other PML lists on two side streams.  Default: 3,
 * or 1 for grids below 4 Mi points (launch-bound).  The lists write disjoint
 * points, so results do not change. */
int sg_set_streams(sg_solver *s, int nstreams);
int sg_get_streams(const sg_solver *s, int *nstreams);
/* Autotuner: times `steps` steps for a set of interior shapes, PML shapes
 * and z-chunk lengths, then (FP32 storage) the one-launch-per-sweep form on
 * one stream with each register bound (waves_per_simd 0, 2, 3, 4); keeps the
 * fastest (written to *best, together with the stream count it needs, set
 * as by sg_set_streams; per-step sweep time in *best_ms) and RESETS the
 * wavefields (sg_reset): call it before the real run. */
int sg_autotune(sg_solver *s, int steps, sg_tile *best, double *best_ms);
/* Algorithmic HBM bytes of one velocity and one stress sweep (the lean
 * split-PML word count times the storage element size). */
int sg_sweep_bytes(const sg_solver *s, double *velocity_bytes, double *stress_bytes);
/* Device copy of `bytes` between two solver-owned scratch buffers, timed with
 * HIP events: the measured-copy yardstick for bandwidth fractions. */
int sg_copy_bandwidth(long long bytes, int reps, double *gbps);

/* ---- batched shots ------------------------------------------------------- */

/* Groups existing FP32 solvers (any sizes) so one launch per half-step
 * advances all of them over a cross-shot tile work-list.  The solvers stay
 * owned by the caller and must outlive the batch.  A shot's source,
 * stations, damping or medium may change between batch steps: each
 * sg_batch_step re-reads the shots' descriptors and re-uploads them when
 * they changed, after any work queued on the shots' own streams.  Before a
 * shot is stepped or read on its own, call sg_batch_synchronize. */
int sg_batch_create(sg_solver *const *shots, int n, sg_batch **out);
/* Launch shape of the batched sweeps: box shapes (as sg_set_tile; z_chunk
 * must be > 0), grid size (0 = one block per work item; smaller = persistent
 * blocks striding over the list) and whether one step is replayed from a
 * hipGraph.  With (4,1) boxes throughout each half-step is ONE launch over
 * the work-lists of all shots, its registers bounded by waves_per_simd;
 * other shapes run one launch per work-list (the default: (16,1) interior
 * and (4,1) PML boxes, 24-plane chunks, one block per item, graph on). */
int sg_batch_set_launch(sg_batch *b, const sg_tile *tile, int grid, int use_graph);
/* Batch autotuner: times `steps` batched steps for z-chunk lengths, register
 * bounds and grid sizes of the one-launch form (and two per-list shapes),
 * keeps the fastest (its tile and grid written to *best and *best_grid, the
 * time per step to *best_ms) and RESETS every shot (sg_reset). */
int sg_batch_autotune(sg_batch *b, int steps, sg_tile *best, int *best_grid, double *best_ms);
int sg_batch_step(sg_batch *b, int nsteps);
/* sg_batch_step bracketed by HIP events on the batch stream; waits. */
int sg_batch_step_timed(sg_batch *b, int nsteps, double *ms);
int sg_batch_synchronize(sg_batch *b);
int sg_batch_destroy(sg_batch *b);

/* ---- host numerics (FP64, no device work) --------------------------------- */

/* s(it) of sourceFunction in the reference's float arithmetic. */
int sg_ricker(float f0, float dt, int nt, int delay_mode, float *out);
/* d(i), i < n, for one axis of total extent n (sg_damping_kind). */
int sg_damping_profile(int n, int pml, double h, int kind, double v, double r, double tau, double *out);
/* Coefficients: kind 0 = the reference literals (CPUPML/ElasticMacro.h:65-69),
 * kind 1 = the exact rationals 19845/16384, -735/8192, 567/40960,
 * -405/229376, 35/294912. */
int sg_diff_coef(int kind, float out[5]);

/* Unit-variance Gaussian-correlated random field (correlation exp(-r^2/a^2),
 * a = corr cells), seeded and reproducible, clipped to +-clip; x-fastest. */
int sg_gaussian_field(int nx, int ny, int nz, double corr, unsigned long long seed, double clip, float *out);

/* Raw float32 model files of the reference's MATLAB tools (V.dat,
 * randomMedium/RandomMedia_3_d.m:101-103, fwrite(Fid, V, 'float')): MATLAB
 * writes V(i,j,k) in column order, i fastest, which is this library's
 * x-fastest layout with x = i.  Reading requires the file to hold exactly n
 * floats. */
int sg_read_raw_f32(const char *path, long long n, float *out);
int sg_write_raw_f32(const char *path, long long n, const float *in);

/* The normalised Courant number of a model given as n points of b, lambda,
 * mu (sg_courant; what sg_set_medium checks), on the host in double.  The
 * same validity checks as sg_set_medium. */
int sg_courant_number(long long n, const float *buoyancy, const float *lambda, const float *mu, float dt, float dx,
                      float dy, float dz, const float diff_coef[5], double *courant);

/* b = 1/rho, mu = rho*vs^2, lambda = rho*vp^2 - 2*mu for n points, formed in
 * double and rounded once to float.  rho and vs must be positive and finite,
 * vp finite. */
int sg_elastic_from_velocity(long long n, const float *vp, const float *vs, const float *rho, float *buoyancy,
                             float *lambda, float *mu);

/* ---- random media on the GPU (spectral method) ------------------------------ */

/* Spectrum of sg_random_medium, indexed by the signed bin index s of each
 * axis (s = m for m <= (n-1)/2, else m - n):
 *   REFERENCE:  (a0 a1 a2) / (1 + (kx^2 a0^2)^p + (ky^2 a1^2)^p + (kz^2 a2^2)^p),
 *               k = s / l, p = order: the form of randomMedium/RandomMedia_3_d.m:29
 *               (a = b = c = 10, p = 10, l = 50 there); the field is |psi|
 *               as in the script (:66);
 *   GAUSSIAN:   exp(-(kx^2 a0^2 + ky^2 a1^2 + kz^2 a2^2) / 4), k = 2 pi s / n
 *               (radians per cell, a in cells): correlation exp(-(x/a0)^2 - ...);
 *   VON_KARMAN: (1 + kx^2 a0^2 + ky^2 a1^2 + kz^2 a2^2)^-(order + 3/2)
 *               (order 0.5: exponential correlation);
 * the Gaussian and von Karman fields are Re psi. */
typedef enum sg_spectrum_kind {
    SG_SPECTRUM_REFERENCE = 0,
    SG_SPECTRUM_GAUSSIAN = 1,
    SG_SPECTRUM_VON_KARMAN = 2
} sg_spectrum_kind;

typedef struct sg_spectrum {
    int kind;     /* sg_spectrum_kind */
    double a[3];  /* correlation lengths along x, y, z */
    double order; /* REFERENCE: exponent p; VON_KARMAN: order (> 0) */
    double l;     /* REFERENCE: wavenumber scale */
} sg_spectrum;

/* Seeded random medium by the reference's spectral method
 * (randomMedium/RandomMedia_3_d.m:20-93) on the current HIP device:
 * psi = ifft3(sqrt(P) e^{i theta}) with theta uniform from a counter-based
 * generator keyed by seed and bin, the field's mean removed, scaled to
 * standard deviation eps: out = v0 (1 + delta), or delta itself when
 * v0 <= 0.  nz x ny x nx floats, x fastest, into host memory. */
int sg_random_medium(int nx, int ny, int nz, const sg_spectrum *spec, unsigned long long seed, double v0, double eps,
                     float *out);
/* The same, with the device time of the build (HIP events around the
 * spectrum, transform, moments and scaling; the FFT plan and the copy of the
 * result to the host excluded) in milliseconds. */
int sg_random_medium_timed(int nx, int ny, int nz, const sg_spectrum *spec, unsigned long long seed, double v0,
                           double eps, float *out, double *device_ms);

/* Last error message of the calling thread ("" if none). */
const char *sg_last_error(void);
int sg_api_version(void);

#ifdef __cplusplus
}
#endif
#endif /* SGFD_H */
