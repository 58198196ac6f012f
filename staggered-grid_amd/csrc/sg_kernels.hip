// sg_kernels.hip -- per-point ("simple") sweeps and the small kernels of a
// time step: source injection, station gather, FP64 energy, copy yardstick.
//
// The simple sweeps compute one grid point per thread straight from global
// memory.  They are the plain statement of the update that the tiled sweeps
// (sg_tiled.hip) must reproduce, and a selectable kernel variant.
//
// Update being computed (3DCPUwithPML_3DArray_GridIncludePMLandOrder/
// ElasticWaveEquation3D.cpp:134-249, with the interior written unsplit as in
// 3DGPUwithPML_1DArraySingleGPU_concise/ElasticWaveEquation.cu:244-282,
// 526-585):
//   vx += dt*bx*(D+x sxx + D-y sxy + D-z sxz), bx = (b(i)+b(i+1))/2, ...
//   sxx += dt*((lam+2mu) D-x vx + lam (D-y vy + D-z vz)), ...
//   sxy += dt*muxy*(D+x vy + D+y vx), muxy = 4/(1/mu + ...) over 4 nodes, ...
// At PML points every damped axis a of a component keeps a split
//   f_a <- A(a) f_a + B(a) inc_a,   A = (1-dt d/2)/(1+dt d/2), B = 1/(1+dt d/2)
// and the component is rebuilt as (f - sum old damped splits) + undamped
// increments + sum new damped splits, or as the plain sum of its splits when
// every contributing axis is damped.
#include "sg_internal.hpp"
#include "sg_precision.hpp"
#include "sg_stencil.hpp"

namespace sg {

// ---------------------------------------------------------------------------
// simple sweeps
// ---------------------------------------------------------------------------

// field value at o (with its compensation residual when there is one)
template <class T, class C>
__device__ __forceinline__ C ldf(const Shot<T, C> &s, int f, long long o)
{
    C v = Cvt<T, C>::ld(s.f[f][o], s.finv[f]);
    if (sizeof(T) == 2 && s.comp[f]) v += Cvt<T, C>::ld(s.comp[f][o], s.cinv[f]);
    return v;
}

template <class T, class C>
__device__ __forceinline__ C dfwd_g(const Shot<T, C> &s, int f, long long o, long long st, const C *c)
{
    const T *p = s.f[f];
    const float inv = s.finv[f];
    auto L = [&](long long q) { return Cvt<T, C>::ld(p[q], inv); };
    C r = c[0] * (L(o + st) - L(o));
    r += c[1] * (L(o + 2 * st) - L(o - st));
    r += c[2] * (L(o + 3 * st) - L(o - 2 * st));
    r += c[3] * (L(o + 4 * st) - L(o - 3 * st));
    r += c[4] * (L(o + 5 * st) - L(o - 4 * st));
    return r;
}

template <class T, class C>
__device__ __forceinline__ C dbwd_g(const Shot<T, C> &s, int f, long long o, long long st, const C *c)
{
    const T *p = s.f[f];
    const float inv = s.finv[f];
    auto L = [&](long long q) { return Cvt<T, C>::ld(p[q], inv); };
    C r = c[0] * (L(o) - L(o - st));
    r += c[1] * (L(o + st) - L(o - 2 * st));
    r += c[2] * (L(o + 2 * st) - L(o - 3 * st));
    r += c[3] * (L(o + 3 * st) - L(o - 4 * st));
    r += c[4] * (L(o + 4 * st) - L(o - 5 * st));
    return r;
}

// per-cell damping: A, B from dt*g/2 (in A) and the slot's local S velocity
template <class T, class C, int NA>
__device__ __forceinline__ void local_damping(const Shot<T, C> &s, long long slot, C (&A)[NA], C (&B)[NA])
{
    if (!s.vs) return;
    const C v = s.vs[slot];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const C hd = A[a] * v;
        B[a] = (C)1 / ((C)1 + hd);
        A[a] = ((C)1 - hd) * B[a];
    }
}

// store of an updated field value (stochastic rounding when enabled for
// 16-bit storage, keyed by point, step and field)
template <class T, class C>
__device__ __forceinline__ T store_val(const Shot<T, C> &s, int f, long long o, C v)
{
    if (sizeof(T) == 2 && s.round_mode)
        return CvtSR<T, C>::st(v, s.fscale[f], sr_bits((uint32_t)o, (uint32_t)*s.d_step, (uint32_t)f));
    return Cvt<T, C>::st(v, s.fscale[f]);
}

template <class T, class C>
__global__ __launch_bounds__(256) void k_velocity_simple(const Shot<T, C> s)
{
    const Geom &g = s.g;
    const int i = blockIdx.x * 64 + threadIdx.x + HALO;
    const int j = blockIdx.y * 4 + threadIdx.y + HALO;
    const int k = blockIdx.z + HALO;
    if (i >= g.nx - HALO || j >= g.ny - HALO) return;
    const long long sy = g.sy, sz = g.sz;
    const long long o = i + j * sy + k * sz;

    C d[3][3];
    d[0][0] = dfwd_g(s, SXX, o, 1, s.cx);
    d[0][1] = dbwd_g(s, SXY, o, sy, s.cy);
    d[0][2] = dbwd_g(s, SXZ, o, sz, s.cz);
    d[1][0] = dbwd_g(s, SXY, o, 1, s.cx);
    d[1][1] = dfwd_g(s, SYY, o, sy, s.cy);
    d[1][2] = dbwd_g(s, SYZ, o, sz, s.cz);
    d[2][0] = dbwd_g(s, SXZ, o, 1, s.cx);
    d[2][1] = dbwd_g(s, SYZ, o, sy, s.cy);
    d[2][2] = dfwd_g(s, SZZ, o, sz, s.cz);

    const float ib = 1.0f / s.bscale;
    const C b0 = Cvt<T, C>::ld(s.b[o], ib);
    C db[3];
    db[0] = s.dt * ((b0 + Cvt<T, C>::ld(s.b[o + 1], ib)) * (C)0.5);
    db[1] = s.dt * ((b0 + Cvt<T, C>::ld(s.b[o + sy], ib)) * (C)0.5);
    db[2] = s.dt * ((b0 + Cvt<T, C>::ld(s.b[o + sz], ib)) * (C)0.5);

    const bool px = in_pml(i, g.nx, g.pml), py = in_pml(j, g.ny, g.pml), pz = in_pml(k, g.nz, g.pml);
    if (!(px || py || pz)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            C v = ldf(s, VX + c, o);
            v += db[c] * (d[c][0] + d[c][1] + d[c][2]);
            s.f[VX + c][o] = store_val(s, VX + c, o, v);
        }
        return;
    }
    const long long slot = split_slot(g, i, j, k);
    const bool dm[3] = {px, py, pz};
    C A[3] = {s.dA[0][i], s.dA[1][j], s.dA[2][k]};
    C B[3] = {s.dB[0][i], s.dB[1][j], s.dB[2][k]};
    local_damping(s, slot, A, B);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const C inc[3] = {db[c] * d[c][0], db[c] * d[c][1], db[c] * d[c][2]};
        C *const sp[3] = {s.sv[3 * c], s.sv[3 * c + 1], s.sv[3 * c + 2]};
        const C v = pml_update<C, 3>(ldf(s, VX + c, o), inc, dm, sp, slot, A, B);
        s.f[VX + c][o] = store_val(s, VX + c, o, v);
    }
}

template <class T, class C>
__global__ __launch_bounds__(256) void k_stress_simple(const Shot<T, C> s)
{
    const Geom &g = s.g;
    const int i = blockIdx.x * 64 + threadIdx.x + HALO;
    const int j = blockIdx.y * 4 + threadIdx.y + HALO;
    const int k = blockIdx.z + HALO;
    if (i >= g.nx - HALO || j >= g.ny - HALO) return;
    const long long sy = g.sy, sz = g.sz;
    const long long o = i + j * sy + k * sz;

    const C dX = dbwd_g(s, VX, o, 1, s.cx);
    const C dY = dbwd_g(s, VY, o, sy, s.cy);
    const C dZ = dbwd_g(s, VZ, o, sz, s.cz);
    const C dxy_x = dfwd_g(s, VY, o, 1, s.cx), dxy_y = dfwd_g(s, VX, o, sy, s.cy);
    const C dxz_x = dfwd_g(s, VZ, o, 1, s.cx), dxz_z = dfwd_g(s, VX, o, sz, s.cz);
    const C dyz_y = dfwd_g(s, VZ, o, sy, s.cy), dyz_z = dfwd_g(s, VY, o, sz, s.cz);

    const float il = 1.0f / s.lscale, im = 1.0f / s.mscale;
    const C lam = Cvt<T, C>::ld(s.lam[o], il);
    auto M = [&](long long q) { return Cvt<T, C>::ld(s.mu[q], im); };
    const C mu = M(o);
    const C r0 = (C)1 / mu, rx = (C)1 / M(o + 1), ry = (C)1 / M(o + sy), rz = (C)1 / M(o + sz);
    const C rxy = (C)1 / M(o + 1 + sy), rxz = (C)1 / M(o + 1 + sz), ryz = (C)1 / M(o + sy + sz);
    const C muxy = (C)4 / (((r0 + rx) + ry) + rxy);
    const C muxz = (C)4 / (((r0 + rx) + rz) + rxz);
    const C muyz = (C)4 / (((r0 + ry) + rz) + ryz);

    const C dl = s.dt * lam, dl2m = s.dt * (lam + (C)2 * mu);
    const C dmxy = s.dt * muxy, dmxz = s.dt * muxz, dmyz = s.dt * muyz;

    C old[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) old[c] = ldf(s, SXX + c, o);
    if (o == s.src && s.src_saved) {
        // TRANSIENT source: the stress update starts from sigma before injection
#pragma unroll
        for (int c = 0; c < 3; ++c) old[c] = Cvt<T, C>::ld(s.src_saved[c], s.finv[SXX + c]);
    }

    const bool px = in_pml(i, g.nx, g.pml), py = in_pml(j, g.ny, g.pml), pz = in_pml(k, g.nz, g.pml);
    C nw[6];
    if (!(px || py || pz)) {
        nw[0] = old[0] + (dl2m * dX + dl * (dY + dZ));
        nw[1] = old[1] + (dl2m * dY + dl * (dX + dZ));
        nw[2] = old[2] + (dl2m * dZ + dl * (dX + dY));
        nw[3] = old[3] + dmxy * (dxy_x + dxy_y);
        nw[4] = old[4] + dmxz * (dxz_x + dxz_z);
        nw[5] = old[5] + dmyz * (dyz_y + dyz_z);
    } else {
        const long long slot = split_slot(g, i, j, k);
        const bool dm[3] = {px, py, pz};
        C A[3] = {s.dA[0][i], s.dA[1][j], s.dA[2][k]};
        C B[3] = {s.dB[0][i], s.dB[1][j], s.dB[2][k]};
        local_damping(s, slot, A, B);
        {
            const C inc[3] = {dl2m * dX, dl * dY, dl * dZ};
            C *const sp[3] = {s.ss[XX_X], s.ss[XX_Y], s.ss[XX_Z]};
            nw[0] = pml_update<C, 3>(old[0], inc, dm, sp, slot, A, B);
        }
        {
            const C inc[3] = {dl * dX, dl2m * dY, dl * dZ};
            C *const sp[3] = {s.ss[YY_X], s.ss[YY_Y], s.ss[YY_Z]};
            nw[1] = pml_update<C, 3>(old[1], inc, dm, sp, slot, A, B);
        }
        {
            const C inc[3] = {dl * dX, dl * dY, dl2m * dZ};
            C *const sp[3] = {s.ss[ZZ_X], s.ss[ZZ_Y], s.ss[ZZ_Z]};
            nw[2] = pml_update<C, 3>(old[2], inc, dm, sp, slot, A, B);
        }
        {
            const C inc[2] = {dmxy * dxy_x, dmxy * dxy_y};
            const bool m2[2] = {px, py};
            C *const sp[2] = {s.ss[XY_X], s.ss[XY_Y]};
            const C A2[2] = {A[0], A[1]}, B2[2] = {B[0], B[1]};
            nw[3] = pml_update<C, 2>(old[3], inc, m2, sp, slot, A2, B2);
        }
        {
            const C inc[2] = {dmxz * dxz_x, dmxz * dxz_z};
            const bool m2[2] = {px, pz};
            C *const sp[2] = {s.ss[XZ_X], s.ss[XZ_Z]};
            const C A2[2] = {A[0], A[2]}, B2[2] = {B[0], B[2]};
            nw[4] = pml_update<C, 2>(old[4], inc, m2, sp, slot, A2, B2);
        }
        {
            const C inc[2] = {dmyz * dyz_y, dmyz * dyz_z};
            const bool m2[2] = {py, pz};
            C *const sp[2] = {s.ss[YZ_Y], s.ss[YZ_Z]};
            const C A2[2] = {A[1], A[2]}, B2[2] = {B[1], B[2]};
            nw[5] = pml_update<C, 2>(old[5], inc, m2, sp, slot, A2, B2);
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) s.f[SXX + c][o] = store_val(s, SXX + c, o, nw[c]);
}

// ---------------------------------------------------------------------------
// source injection (one thread): also advances the device step counter
// ---------------------------------------------------------------------------
// TRANSIENT: sigma = float(double(sigma) + pow(dt,2)*s(it)) for sxx/syy/szz
// after saving the pre-injection values (3DCPUwithPML.../ElasticWaveEquation3D.cpp:
// 123-125).  ACCUMULATE: sigma += dt*dt*s(it) in the storage precision
// (3DGPUwithPML_1DArraySingleGPU_concise/ElasticWaveEquation.cu:216-218).
template <class T, class C>
__global__ void k_inject(const Shot<T, C> s, const double *amp_d, const C *amp_acc, int namp, int mode,
                         int *d_step, T *saved)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int it = *d_step;
    *d_step = it + 1;
    if (s.src < 0) return;
    for (int c = 0; c < 3; ++c) {
        const int f = SXX + c;
        T *p = s.f[f] + s.src;
        T *pc = (sizeof(T) == 2 && s.comp[f]) ? s.comp[f] + s.src : nullptr;
        const C cur = ldf(s, f, s.src);
        C nv = cur;
        if (mode == INJECT_ACCUMULATE) {
            if (it < namp) nv = cur + amp_acc[it];
        } else {
            saved[c] = *p;
            if (pc) saved[3 + c] = *pc;
            if (it < namp) nv = mode == INJECT_TRANSIENT_F ? cur + amp_acc[it] : (C)((double)cur + amp_d[it]);
        }
        if (it < namp) {
            const T q = Cvt<T, C>::st(nv, s.fscale[f]);
            *p = q;
            if (pc) *pc = Cvt<T, C>::st(nv - Cvt<T, C>::ld(q, s.finv[f]), s.cscale[f]);
        }
    }
}

// ---------------------------------------------------------------------------
// station gather: one thread per (station, component) after the stress sweep
// ---------------------------------------------------------------------------
template <class T, class C>
__global__ void k_gather(const Shot<T, C> s, const long long *idx, const int *comp, int nst, int ncomp,
                         int interval, long long cap, C *out, const int *d_step)
{
    const int it = *d_step - 1;
    if (it < 0 || it % interval) return;
    const long long row = it / interval;
    if (row >= cap) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nst * ncomp) return;
    const int st = t / ncomp, q = t - st * ncomp;
    const int f = comp[q];
    out[row * (long long)nst * ncomp + t] = ldf(s, f, idx[st]);
}

// batched FP32 forms of the two kernels above: one shot per thread / per grid row
__global__ void k_inject_batch(const Shot<float, float> *shots, const AuxDesc *aux, int nshots)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nshots) return;
    const Shot<float, float> &s = shots[t];
    const AuxDesc &a = aux[t];
    const int it = *a.d_step;
    *a.d_step = it + 1;
    if (s.src < 0) return;
    for (int c = 0; c < 3; ++c) {
        float *p = s.f[SXX + c] + s.src;
        if (a.inject == INJECT_ACCUMULATE) {
            if (it < a.namp) *p = *p + a.amp_acc[it];
        } else {
            a.saved[c] = *p;
            if (it < a.namp) *p = a.inject == INJECT_TRANSIENT_F ? *p + a.amp_acc[it] : (float)((double)*p + a.amp_d[it]);
        }
    }
}

__global__ void k_gather_batch(const Shot<float, float> *shots, const AuxDesc *aux, int nshots)
{
    const int sh = blockIdx.y;
    if (sh >= nshots) return;
    const AuxDesc &a = aux[sh];
    const int n = a.nst * a.ncomp;
    const int it = *a.d_step - 1;
    if (n == 0 || it < 0 || it % a.interval) return;
    const long long row = it / a.interval;
    if (row >= a.cap) return;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int st = t / a.ncomp, q = t - st * a.ncomp;
        a.st_out[row * (long long)n + t] = shots[sh].f[a.st_comp[q]][a.st_idx[st]];
    }
}

// ---------------------------------------------------------------------------
// FP64 energy over the updated region
// ---------------------------------------------------------------------------
// sqrt(mu*b) of every damped updated point into its split slot
// (..._concise/ElasticWaveEquation.cu:285: Vs_wave = sqrt(mu * buoyancy))
template <class T, class C>
__global__ __launch_bounds__(256) void k_fill_vs(const Shot<T, C> s, C *vs)
{
    const Geom &g = s.g;
    const long long n = (long long)g.ux * g.uy * g.uz;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.ux) + HALO;
        const long long r = t / g.ux;
        const int j = (int)(r % g.uy) + HALO;
        const int k = (int)(r / g.uy) + HALO;
        if (!(in_pml(i, g.nx, g.pml) || in_pml(j, g.ny, g.pml) || in_pml(k, g.nz, g.pml))) continue;
        const long long o = i + j * g.sy + k * g.sz;
        const C m = Cvt<T, C>::ld(s.mu[o], 1.0f / s.mscale), b = Cvt<T, C>::ld(s.b[o], 1.0f / s.bscale);
        vs[split_slot(g, i, j, k)] = sqrt(m * b);
    }
}

// Largest |value| of each of the nine wavefields (unscaled), as float bits
// (non-negative floats order like unsigned integers).
template <class T, class C>
__global__ __launch_bounds__(256) void k_maxabs(const Shot<T, C> s, unsigned *out9)
{
    const long long n = (long long)s.g.nx * s.g.ny * s.g.nz;
    float m[NFIELD];
#pragma unroll
    for (int f = 0; f < NFIELD; ++f) m[f] = 0.0f;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int f = 0; f < NFIELD; ++f) m[f] = fmaxf(m[f], fabsf((float)Cvt<T, C>::ld(s.f[f][t], s.finv[f])));
    }
    __shared__ float red[NFIELD][256];
#pragma unroll
    for (int f = 0; f < NFIELD; ++f) red[f][threadIdx.x] = m[f];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int f = 0; f < NFIELD; ++f) red[f][threadIdx.x] = fmaxf(red[f][threadIdx.x], red[f][threadIdx.x + w]);
        __syncthreads();
    }
    if ((int)threadIdx.x < NFIELD) atomicMax(&out9[threadIdx.x], __float_as_uint(red[threadIdx.x][0]));
}

// FP64 energies over the updated region, summed in a fixed order (one block
// per z plane of the updated region, then one block over the planes), so a
// result does not depend on scheduling:
//   e[0] = sum rho |v|^2 / 2   (kinetic, rho = 1/b)
//   e[1] = sum of the six stress components squared
//   e[2] = sum sigma:sigma / (4 mu) - lambda (tr sigma)^2 / (4 mu (3 lambda + 2 mu))
//          (strain energy of the isotropic medium; sigma:sigma counts each
//          shear component twice)
// part: ENERGY_PART doubles per plane; block (64, 16).
template <class T, class C>
__global__ __launch_bounds__(1024) void k_energy_planes(const Shot<T, C> s, double *part)
{
    const Geom &g = s.g;
    const int k = (int)blockIdx.x + HALO;
    const float ib = 1.0f / s.bscale, il = 1.0f / s.lscale, im = 1.0f / s.mscale;
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
    for (int j = HALO + (int)threadIdx.y; j < g.ny - HALO; j += (int)blockDim.y) {
        for (int i = HALO + (int)threadIdx.x; i < g.nx - HALO; i += (int)blockDim.x) {
            const long long o = i + j * g.sy + k * g.sz;
            const double b = (double)Cvt<T, C>::ld(s.b[o], ib);
            const double lam = (double)Cvt<T, C>::ld(s.lam[o], il);
            const double mu = (double)Cvt<T, C>::ld(s.mu[o], im);
            double v2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double v = (double)ldf(s, VX + c, o);
                v2 += v * v;
            }
            double n2 = 0.0, sh2 = 0.0, tr = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double v = (double)ldf(s, SXX + c, o);
                n2 += v * v;
                tr += v;
            }
            for (int c = 3; c < 6; ++c) {
                const double v = (double)ldf(s, SXX + c, o);
                sh2 += v * v;
            }
            e0 += 0.5 * v2 / b;
            e1 += n2 + sh2;
            e2 += (n2 + 2.0 * sh2) / (4.0 * mu) - lam * tr * tr / (4.0 * mu * (3.0 * lam + 2.0 * mu));
        }
    }
    __shared__ double red[3][1024];
    const int t = (int)(threadIdx.x + threadIdx.y * blockDim.x);
    red[0][t] = e0;
    red[1][t] = e1;
    red[2][t] = e2;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 3; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t < 3) part[(size_t)blockIdx.x * ENERGY_PART + t] = red[t][0];
}

// out[0..2] = the plane partials summed in plane order by a fixed tree
__global__ __launch_bounds__(256) void k_energy_sum(const double *part, int nplanes, double *out)
{
    __shared__ double red[3][256];
    const int t = (int)threadIdx.x;
    double e[3] = {0.0, 0.0, 0.0};
    for (int p = t; p < nplanes; p += 256)
        for (int q = 0; q < 3; ++q) e[q] += part[(size_t)p * ENERGY_PART + q];
    for (int q = 0; q < 3; ++q) red[q][t] = e[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 3; ++q) red[q][t] += red[q][t + w];
        __syncthreads();
    }
    if (t < 3) out[t] = red[t][0];
}

// streaming copy: one pass, each thread loads four 8-byte values of its
// block's contiguous 8 KiB, then stores them (measured on one MI355X, 1 GiB
// each way: 6.16 TB/s; 4-byte loads 6.06, 16-byte loads 5.66, a grid-stride
// loop 4.2-4.5)
typedef float v2f __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy(const v2f *__restrict__ src, v2f *__restrict__ dst, long long n2)
{
    const long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
    v2f v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long long t = base + q * 256;
        if (t < n2) v[q] = src[t];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const long long t = base + q * 256;
        if (t < n2) dst[t] = v[q];
    }
}

__global__ __launch_bounds__(256) void k_zero(uint32_t *p, long long n)
{
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        p[t] = 0u;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

template <class T, class C>
static hipError_t vel_simple(const void *shot, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    dim3 grid((s.g.ux + 63) / 64, (s.g.uy + 3) / 4, s.g.uz), block(64, 4);
    hipLaunchKernelGGL((k_velocity_simple<T, C>), grid, block, 0, st, s);
    return hipGetLastError();
}

template <class T, class C>
static hipError_t str_simple(const void *shot, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    dim3 grid((s.g.ux + 63) / 64, (s.g.uy + 3) / 4, s.g.uz), block(64, 4);
    hipLaunchKernelGGL((k_stress_simple<T, C>), grid, block, 0, st, s);
    return hipGetLastError();
}



#define SG_DISPATCH(prec, FN, ...)                               \
    switch (prec) {                                              \
    case 0: return FN<float, float>(__VA_ARGS__);                \
    case 1: return FN<double, double>(__VA_ARGS__);              \
    case 2: return FN<_Float16, float>(__VA_ARGS__);             \
    case 3: return FN<bf16, float>(__VA_ARGS__);                 \
    default: return hipErrorInvalidValue;                        \
    }

template <class T, class C>
static hipError_t vel_any(const void *shot, hipStream_t st)
{
    return vel_simple<T, C>(shot, st);
}

template <class T, class C>
static hipError_t str_any(const void *shot, hipStream_t st)
{
    return str_simple<T, C>(shot, st);
}

hipError_t launch_velocity_simple(int prec, const void *shot, hipStream_t st)
{
    SG_DISPATCH(prec, vel_any, shot, st)
}

hipError_t launch_stress_simple(int prec, const void *shot, hipStream_t st)
{
    SG_DISPATCH(prec, str_any, shot, st)
}

template <class T, class C>
static hipError_t inject_t(const void *shot, const double *amp_d, const void *amp_acc, int namp, int mode,
                           int *d_step, void *saved, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    hipLaunchKernelGGL((k_inject<T, C>), dim3(1), dim3(64), 0, st, s, amp_d, static_cast<const C *>(amp_acc), namp,
                       mode, d_step, static_cast<T *>(saved));
    return hipGetLastError();
}

hipError_t launch_sweep(int which, int prec, int ty, int ny, int dm, int comp, const void *shots_dev,
                        const WorkItem *items, int nitems, int grid, hipStream_t st)
{
    switch (prec) {
    case 0: return sweep_fp32(which, ty, ny, dm, comp, shots_dev, items, nitems, grid, st);
    case 1: return sweep_fp64(which, ty, ny, dm, comp, shots_dev, items, nitems, grid, st);
    case 2: return sweep_fp16(which, ty, ny, dm, comp, shots_dev, items, nitems, grid, st);
    case 3: return sweep_bf16(which, ty, ny, dm, comp, shots_dev, items, nitems, grid, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep_fused(int which, int prec, int comp, int waves, const void *shots_dev, const WorkItem *items,
                              const int *off, int grid, hipStream_t st)
{
    switch (prec) {
    case 0: return sweep_fused_fp32(which, comp, waves, shots_dev, items, off, grid, st);
    case 1: return sweep_fused_fp64(which, comp, waves, shots_dev, items, off, grid, st);
    case 2: return sweep_fused_fp16(which, comp, waves, shots_dev, items, off, grid, st);
    case 3: return sweep_fused_bf16(which, comp, waves, shots_dev, items, off, grid, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_inject(int prec, const void *shot, const double *amp_d, const void *amp_acc, int namp,
                         int mode, int *d_step, void *saved, hipStream_t st)
{
    SG_DISPATCH(prec, inject_t, shot, amp_d, amp_acc, namp, mode, d_step, saved, st)
}

template <class T, class C>
static hipError_t gather_t(const void *shot, const long long *idx, const int *comp, int nst, int ncomp,
                           int interval, long long cap, void *out, const int *d_step, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    const int n = nst * ncomp;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_gather<T, C>), dim3((n + 255) / 256), dim3(256), 0, st, s, idx, comp, nst, ncomp,
                       interval, cap, static_cast<C *>(out), d_step);
    return hipGetLastError();
}

hipError_t launch_gather(int prec, const void *shot, const long long *idx, const int *comp, int nst, int ncomp,
                         int interval, long long capacity, void *out, const int *d_step, hipStream_t st)
{
    SG_DISPATCH(prec, gather_t, shot, idx, comp, nst, ncomp, interval, capacity, out, d_step, st)
}

template <class T, class C>
static hipError_t energy_t(const void *shot, double *scratch, double *out3, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    const int np = s.g.nz - 2 * HALO;
    hipLaunchKernelGGL((k_energy_planes<T, C>), dim3(np), dim3(64, 16), 0, st, s, scratch);
    hipLaunchKernelGGL(k_energy_sum, dim3(1), dim3(256), 0, st, (const double *)scratch, np, out3);
    return hipGetLastError();
}

hipError_t launch_energy(int prec, const void *shot, double *scratch, double *out3, hipStream_t st)
{
    SG_DISPATCH(prec, energy_t, shot, scratch, out3, st)
}

template <class T, class C>
static hipError_t fill_vs_t(const void *shot, void *vs, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    hipLaunchKernelGGL((k_fill_vs<T, C>), dim3(2048), dim3(256), 0, st, s, static_cast<C *>(vs));
    return hipGetLastError();
}

hipError_t launch_fill_vs(int prec, const void *shot, void *vs, hipStream_t st)
{
    SG_DISPATCH(prec, fill_vs_t, shot, vs, st)
}

template <class T, class C>
static hipError_t maxabs_t(const void *shot, unsigned *out9, hipStream_t st)
{
    const Shot<T, C> &s = *static_cast<const Shot<T, C> *>(shot);
    hipLaunchKernelGGL((k_maxabs<T, C>), dim3(1024), dim3(256), 0, st, s, out9);
    return hipGetLastError();
}

hipError_t launch_maxabs(int prec, const void *shot, unsigned *out9, hipStream_t st)
{
    SG_DISPATCH(prec, maxabs_t, shot, out9, st)
}

hipError_t launch_inject_batch(const void *shots_dev, const AuxDesc *aux, int nshots, hipStream_t st)
{
    hipLaunchKernelGGL(k_inject_batch, dim3((nshots + 63) / 64), dim3(64), 0, st,
                       static_cast<const Shot<float, float> *>(shots_dev), aux, nshots);
    return hipGetLastError();
}

hipError_t launch_gather_batch(const void *shots_dev, const AuxDesc *aux, int nshots, int max_n, hipStream_t st)
{
    if (max_n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_batch, dim3((max_n + 255) / 256, nshots), dim3(256), 0, st,
                       static_cast<const Shot<float, float> *>(shots_dev), aux, nshots);
    return hipGetLastError();
}

hipError_t launch_copy(const float *src, float *dst, long long n, hipStream_t st)
{
    const long long n2 = n / 2, blocks = (n2 + 1023) / 1024;
    if (blocks <= 0) return hipSuccess;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, st, reinterpret_cast<const v2f *>(src),
                       reinterpret_cast<v2f *>(dst), n2);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, long long bytes, hipStream_t st)
{
    hipLaunchKernelGGL(k_zero, dim3(4096), dim3(256), 0, st, static_cast<uint32_t *>(p), bytes / 4);
    return hipGetLastError();
}

}  // namespace sg
