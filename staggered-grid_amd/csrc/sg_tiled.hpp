// sg_tiled.hpp -- the tiled velocity and stress sweeps (included by sg_kernels.hip).
//
// Mapping to gfx950:
//  * a block of 256 threads (64 x 4) owns an xy tile of 64 x TY points and
//    marches through a chunk of z planes; TY = 4*NY, each thread computes NY
//    consecutive y rows of one x column;
//  * the current z plane of every field that needs x or y derivatives is
//    staged in LDS with its radius-5 halo (x rows of 74, y columns of TY+10),
//    loaded with x-contiguous (coalesced) global reads;
//  * fields that need z derivatives are kept in per-thread register queues of
//    ten planes, so every value is read from HBM once per sweep and the z taps
//    cost no LDS traffic;
//  * the PML is handled per (tile, plane): tiles whose points are all interior
//    take the unsplit update with no branches, the others call the split
//    update with per-point damping flags;
//  * blocks are ordered so that each XCD (blockIdx % 8 on MI355X) receives a
//    contiguous band of tiles, keeping halo re-reads inside one L2.
#pragma once

namespace sg {

constexpr int TX = 64;          // x points per tile = threads in x
constexpr int XW = TX + 10;     // LDS row width with the x halo
constexpr int NTHR = 256;

struct TileMap {
    int nbx, nby, nbz, zc;
};

// XCD-aware decode of a linear block id (grid size padded to a multiple of 8).
__device__ __forceinline__ bool decode_block(const TileMap &m, int &bx, int &by, int &bz)
{
    const int nb = gridDim.x;
    const int per = nb >> 3;
    const int b = blockIdx.x;
    const int logical = (b & 7) * per + (b >> 3);
    const int nreal = m.nbx * m.nby * m.nbz;
    if (logical >= nreal) return false;
    bx = logical % m.nbx;
    const int r = logical / m.nbx;
    by = r % m.nby;
    bz = r / m.nby;
    return true;
}

template <class T, class C>
__device__ __forceinline__ C gload(const T *p, long long o, bool ok, float inv)
{
    return ok ? Cvt<T, C>::ld(p[o], inv) : (C)0;
}

// 5-pair staggered sums over an LDS line: forward at position c uses
// f[c+n] - f[c+1-n], backward uses f[c+n-1] - f[c-n], n = 1..5.
template <class C>
__device__ __forceinline__ C lfwd(const C *f, int c, int st, const C *k)
{
    C r = k[0] * (f[(c + 1) * st] - f[c * st]);
    r += k[1] * (f[(c + 2) * st] - f[(c - 1) * st]);
    r += k[2] * (f[(c + 3) * st] - f[(c - 2) * st]);
    r += k[3] * (f[(c + 4) * st] - f[(c - 3) * st]);
    r += k[4] * (f[(c + 5) * st] - f[(c - 4) * st]);
    return r;
}

template <class C>
__device__ __forceinline__ C lbwd(const C *f, int c, int st, const C *k)
{
    C r = k[0] * (f[c * st] - f[(c - 1) * st]);
    r += k[1] * (f[(c + 1) * st] - f[(c - 2) * st]);
    r += k[2] * (f[(c + 2) * st] - f[(c - 3) * st]);
    r += k[3] * (f[(c + 3) * st] - f[(c - 4) * st]);
    r += k[4] * (f[(c + 4) * st] - f[(c - 5) * st]);
    return r;
}

// queue variants: q[m] holds plane k - base + m
template <class C>
__device__ __forceinline__ C qfwd(const C *q, int c, const C *k)  // c = index of plane k
{
    return lfwd<C>(q, c, 1, k);
}
template <class C>
__device__ __forceinline__ C qbwd(const C *q, int c, const C *k)
{
    return lbwd<C>(q, c, 1, k);
}

// ---------------------------------------------------------------------------
// velocity sweep
// ---------------------------------------------------------------------------
template <class C, int NY>
struct VelLds {
    static constexpr int TY = 4 * NY, YH = TY + 10;
    C sxx[TY][XW];
    C sxz[TY][XW];
    C syy[YH][TX];
    C syz[YH][TX];
    C sxy[YH][XW];
    C b[TY + 1][TX + 1];
};

template <class C, int NY>
struct StrLds {
    static constexpr int TY = 4 * NY, YH = TY + 10;
    C vx[YH][XW];
    C vy[YH][XW];
    C vz[YH][XW];
    C rm[2][TY + 1][TX + 1];  // 1/mu, planes k and k+1 (ring)
};

// One velocity tile: columns [i0, i0+64) x rows [j0, j0+TY), planes [k0, k1).
template <class T, class C, int NY>
__device__ __forceinline__ void velocity_tile(const Shot<T, C> &s, int bx, int by, int k0, int k1,
                                              VelLds<C, NY> &L)
{
    constexpr int TY = 4 * NY;
    auto &l_sxx = L.sxx;
    auto &l_sxz = L.sxz;
    auto &l_syy = L.syy;
    auto &l_syz = L.syz;
    auto &l_sxy = L.sxy;
    auto &l_b = L.b;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = HALO + bx * TX, j0 = HALO + by * TY;
    const int i = i0 + tx;
    const long long sy = g.sy, sz = g.sz;
    const bool xin = i < g.nx;              // inside the array (loads)
    const bool xup = i < g.nx - HALO;        // updated column

    // thread rows
    int jr[NY];
    bool yin[NY], yup[NY];
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        jr[r] = j0 + ty * NY + r;
        yin[r] = jr[r] < g.ny;
        yup[r] = jr[r] < g.ny - HALO;
    }
    // x-halo column served by this thread (tx < 10), and the b halo (tx == 10)
    const int hc = tx < 5 ? tx : TX + tx;    // LDS column of the halo (0..4, 69..73)
    const int hi = i0 - HALO + hc;           // its x index
    const bool hx = tx < 10 && hi >= 0 && hi < g.nx;

    const float fi_xx = s.finv[SXX], fi_yy = s.finv[SYY], fi_zz = s.finv[SZZ];
    const float fi_xy = s.finv[SXY], fi_xz = s.finv[SXZ], fi_yz = s.finv[SYZ];
    const float ib = 1.0f / s.bscale;

    // z queues: qxz/qyz[m] = plane k-5+m, qzz[m] = plane k-4+m, qb[m] = plane k+m
    C qxz[NY][10], qyz[NY][10], qzz[NY][10], qb[NY][2];
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const bool ok = xin && yin[r];
        const long long o = i + jr[r] * sy;
#pragma unroll
        for (int q = 0; q < 10; ++q) {
            const long long oz = o + (long long)(k0 - 5 + q) * sz;
            qxz[r][q] = gload<T, C>(s.f[SXZ], oz, ok, fi_xz);
            qyz[r][q] = gload<T, C>(s.f[SYZ], oz, ok, fi_yz);
            qzz[r][q] = gload<T, C>(s.f[SZZ], oz + sz, ok, fi_zz);
        }
        qb[r][0] = gload<T, C>(s.b, o + (long long)k0 * sz, ok, ib);
        qb[r][1] = gload<T, C>(s.b, o + (long long)(k0 + 1) * sz, ok, ib);
    }

    const bool tile_xpml = (i0 < HALO + g.pml) || (i0 + TX - 1 >= g.nx - HALO - g.pml);
    const bool tile_ypml = (j0 < HALO + g.pml) || (j0 + TY - 1 >= g.ny - HALO - g.pml);
    const bool px = in_pml(i, g.nx, g.pml);

    for (int k = k0; k < k1; ++k) {
        const long long kz = (long long)k * sz;
        // ---- stage plane k in LDS
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            const bool ok = xin && yin[r];
            const long long o = i + jr[r] * sy + kz;
            l_sxx[lr][tx + 5] = gload<T, C>(s.f[SXX], o, ok, fi_xx);
            l_sxy[lr + 5][tx + 5] = gload<T, C>(s.f[SXY], o, ok, fi_xy);
            l_syy[lr + 5][tx] = gload<T, C>(s.f[SYY], o, ok, fi_yy);
            l_sxz[lr][tx + 5] = qxz[r][5];
            l_syz[lr + 5][tx] = qyz[r][5];
            l_b[lr][tx] = qb[r][0];
            if (tx < 10) {
                const bool okh = hx && yin[r];
                const long long oh = hi + jr[r] * sy + kz;
                l_sxx[lr][hc] = gload<T, C>(s.f[SXX], oh, okh, fi_xx);
                l_sxy[lr + 5][hc] = gload<T, C>(s.f[SXY], oh, okh, fi_xy);
                l_sxz[lr][hc] = gload<T, C>(s.f[SXZ], oh, okh, fi_xz);
            } else if (tx == 10) {
                const bool okh = (i0 + TX) < g.nx && yin[r];
                l_b[lr][TX] = gload<T, C>(s.b, (i0 + TX) + jr[r] * sy + kz, okh, ib);
            }
        }
        // y halo rows: 0..4 -> j0-5..j0-1, 5..9 -> j0+TY..j0+TY+4
        for (int h = ty; h < 10; h += 4) {
            const int lr = h < 5 ? h : TY + h;
            const int jh = j0 - HALO + lr;
            const bool ok = xin && jh >= 0 && jh < g.ny;
            const long long o = i + jh * sy + kz;
            l_syy[lr][tx] = gload<T, C>(s.f[SYY], o, ok, fi_yy);
            l_syz[lr][tx] = gload<T, C>(s.f[SYZ], o, ok, fi_yz);
            l_sxy[lr][tx + 5] = gload<T, C>(s.f[SXY], o, ok, fi_xy);
        }
        if (ty == 0) {
            const int jb = j0 + TY;
            l_b[TY][tx] = gload<T, C>(s.b, i + jb * sy + kz, xin && jb < g.ny, ib);
        }
        __syncthreads();

        // ---- update
        const bool pz = in_pml(k, g.nz, g.pml);
        const bool tile_pml = pz || tile_xpml || tile_ypml;
        const C A0 = s.dA[0][min(i, g.nx - 1)], B0 = s.dB[0][min(i, g.nx - 1)];
        const C A2 = s.dA[2][k], B2 = s.dB[2][k];
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            // derivatives
            C d[3][3];
            d[0][0] = lfwd<C>(&l_sxx[lr][0], tx + 5, 1, s.cx);
            d[0][1] = lbwd<C>(&l_sxy[0][tx + 5], lr + 5, XW, s.cy);
            d[0][2] = qbwd<C>(qxz[r], 5, s.cz);
            d[1][0] = lbwd<C>(&l_sxy[lr + 5][0], tx + 5, 1, s.cx);
            d[1][1] = lfwd<C>(&l_syy[0][tx], lr + 5, TX, s.cy);
            d[1][2] = qbwd<C>(qyz[r], 5, s.cz);
            d[2][0] = lbwd<C>(&l_sxz[lr][0], tx + 5, 1, s.cx);
            d[2][1] = lbwd<C>(&l_syz[0][tx], lr + 5, TX, s.cy);
            d[2][2] = qfwd<C>(qzz[r], 4, s.cz);
            const C b0 = l_b[lr][tx];
            C db[3];
            db[0] = s.dt * ((b0 + l_b[lr][tx + 1]) * (C)0.5);
            db[1] = s.dt * ((b0 + l_b[lr + 1][tx]) * (C)0.5);
            db[2] = s.dt * ((b0 + qb[r][1]) * (C)0.5);
            if (!(xup && yup[r])) continue;
            const long long o = i + jr[r] * sy + kz;
            if (!tile_pml) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    C v = Cvt<T, C>::ld(s.f[VX + c][o], s.finv[VX + c]);
                    v += db[c] * (d[c][0] + d[c][1] + d[c][2]);
                    s.f[VX + c][o] = Cvt<T, C>::st(v, s.fscale[VX + c]);
                }
            } else {
                const bool py = in_pml(jr[r], g.ny, g.pml);
                if (!(px || py || pz)) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        C v = Cvt<T, C>::ld(s.f[VX + c][o], s.finv[VX + c]);
                        v += db[c] * (d[c][0] + d[c][1] + d[c][2]);
                        s.f[VX + c][o] = Cvt<T, C>::st(v, s.fscale[VX + c]);
                    }
                } else {
                    const long long slot = split_slot(g, i, jr[r], k);
                    const bool dm[3] = {px, py, pz};
                    const C A[3] = {A0, s.dA[1][jr[r]], A2};
                    const C B[3] = {B0, s.dB[1][jr[r]], B2};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const C inc[3] = {db[c] * d[c][0], db[c] * d[c][1], db[c] * d[c][2]};
                        C *const sp[3] = {s.sv[3 * c], s.sv[3 * c + 1], s.sv[3 * c + 2]};
                        const C v = pml_update<C, 3>(Cvt<T, C>::ld(s.f[VX + c][o], s.finv[VX + c]), inc, dm, sp,
                                                     slot, A, B);
                        s.f[VX + c][o] = Cvt<T, C>::st(v, s.fscale[VX + c]);
                    }
                }
            }
        }
        __syncthreads();

        // ---- advance the z queues
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const bool ok = xin && yin[r];
            const long long o = i + jr[r] * sy;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                qxz[r][q] = qxz[r][q + 1];
                qyz[r][q] = qyz[r][q + 1];
                qzz[r][q] = qzz[r][q + 1];
            }
            const bool okz5 = ok && (k + 5) < g.nz, okz6 = ok && (k + 6) < g.nz, okz2 = ok && (k + 2) < g.nz;
            qxz[r][9] = gload<T, C>(s.f[SXZ], o + (long long)(k + 5) * sz, okz5, fi_xz);
            qyz[r][9] = gload<T, C>(s.f[SYZ], o + (long long)(k + 5) * sz, okz5, fi_yz);
            qzz[r][9] = gload<T, C>(s.f[SZZ], o + (long long)(k + 6) * sz, okz6, fi_zz);
            qb[r][0] = qb[r][1];
            qb[r][1] = gload<T, C>(s.b, o + (long long)(k + 2) * sz, okz2, ib);
        }
    }
}

template <class T, class C, int NY>
__global__ __launch_bounds__(NTHR) void k_velocity_tiled(const Shot<T, C> s, const TileMap m)
{
    __shared__ VelLds<C, NY> L;
    int bx, by, bz;
    if (!decode_block(m, bx, by, bz)) return;
    const int k0 = HALO + bz * m.zc;
    velocity_tile<T, C, NY>(s, bx, by, k0, min(k0 + m.zc, s.g.nz - HALO), L);
}

// ---------------------------------------------------------------------------
// stress sweep
// ---------------------------------------------------------------------------
template <class T, class C, int NY>
__device__ __forceinline__ void stress_tile(const Shot<T, C> &s, int bx, int by, int k0, int k1,
                                            StrLds<C, NY> &L)
{
    constexpr int TY = 4 * NY;
    auto &l_vx = L.vx;
    auto &l_vy = L.vy;
    auto &l_vz = L.vz;
    auto &l_rm = L.rm;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = HALO + bx * TX, j0 = HALO + by * TY;
    const int i = i0 + tx;
    const long long sy = g.sy, sz = g.sz;
    const bool xin = i < g.nx;
    const bool xup = i < g.nx - HALO;
    int jr[NY];
    bool yin[NY], yup[NY];
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        jr[r] = j0 + ty * NY + r;
        yin[r] = jr[r] < g.ny;
        yup[r] = jr[r] < g.ny - HALO;
    }
    const int hc = tx < 5 ? tx : TX + tx;
    const int hi = i0 - HALO + hc;
    const bool hx = tx < 10 && hi >= 0 && hi < g.nx;
    const float fi_vx = s.finv[VX], fi_vy = s.finv[VY], fi_vz = s.finv[VZ];
    const float il = 1.0f / s.lscale, im = 1.0f / s.mscale;

    // z queues: qvx/qvy[m] = plane k-4+m, qvz[m] = plane k-5+m; mu at plane k
    C qvx[NY][10], qvy[NY][10], qvz[NY][10], muk[NY];
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const bool ok = xin && yin[r];
        const long long o = i + jr[r] * sy;
#pragma unroll
        for (int q = 0; q < 10; ++q) {
            const long long oz = o + (long long)(k0 - 5 + q) * sz;
            qvx[r][q] = gload<T, C>(s.f[VX], oz + sz, ok, fi_vx);
            qvy[r][q] = gload<T, C>(s.f[VY], oz + sz, ok, fi_vy);
            qvz[r][q] = gload<T, C>(s.f[VZ], oz, ok, fi_vz);
        }
        muk[r] = gload<T, C>(s.mu, o + (long long)k0 * sz, ok, im);
    }
    // 1/mu planes k0 and k0+1 into the ring (a zero mu outside the array gives
    // inf, never read by an updated point)
    auto stage_rmu = [&](int kp) {
        const int slot = kp & 1;
        const long long kz = (long long)kp * sz;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            const bool ok = xin && yin[r];
            const C v = gload<T, C>(s.mu, i + jr[r] * sy + kz, ok, im);
            l_rm[slot][lr][tx] = ok ? (C)1 / v : (C)0;
            if (tx == 10) {
                const bool okh = (i0 + TX) < g.nx && yin[r];
                const C w = gload<T, C>(s.mu, (i0 + TX) + jr[r] * sy + kz, okh, im);
                l_rm[slot][lr][TX] = okh ? (C)1 / w : (C)0;
            }
        }
        if (ty == 0) {
            const int jb = j0 + TY;
            const bool ok = xin && jb < g.ny;
            const C v = gload<T, C>(s.mu, i + jb * sy + kz, ok, im);
            l_rm[slot][TY][tx] = ok ? (C)1 / v : (C)0;
            if (tx == 10) {
                const bool okh = (i0 + TX) < g.nx && jb < g.ny;
                const C w = gload<T, C>(s.mu, (i0 + TX) + jb * sy + kz, okh, im);
                l_rm[slot][TY][TX] = okh ? (C)1 / w : (C)0;
            }
        }
    };
    stage_rmu(k0);
    stage_rmu(k0 + 1);

    const bool tile_xpml = (i0 < HALO + g.pml) || (i0 + TX - 1 >= g.nx - HALO - g.pml);
    const bool tile_ypml = (j0 < HALO + g.pml) || (j0 + TY - 1 >= g.ny - HALO - g.pml);
    const bool px = in_pml(i, g.nx, g.pml);
    const int src_k = s.src >= 0 ? (int)(s.src / sz) : -1;

    for (int k = k0; k < k1; ++k) {
        const long long kz = (long long)k * sz;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            l_vx[lr + 5][tx + 5] = qvx[r][4];
            l_vy[lr + 5][tx + 5] = qvy[r][4];
            l_vz[lr + 5][tx + 5] = qvz[r][5];
            if (tx < 10) {
                const bool okh = hx && yin[r];
                const long long oh = hi + jr[r] * sy + kz;
                l_vx[lr + 5][hc] = gload<T, C>(s.f[VX], oh, okh, fi_vx);
                l_vy[lr + 5][hc] = gload<T, C>(s.f[VY], oh, okh, fi_vy);
                l_vz[lr + 5][hc] = gload<T, C>(s.f[VZ], oh, okh, fi_vz);
            }
        }
        for (int h = ty; h < 10; h += 4) {
            const int lr = h < 5 ? h : TY + h;
            const int jh = j0 - HALO + lr;
            const bool ok = xin && jh >= 0 && jh < g.ny;
            const long long o = i + jh * sy + kz;
            l_vx[lr][tx + 5] = gload<T, C>(s.f[VX], o, ok, fi_vx);
            l_vy[lr][tx + 5] = gload<T, C>(s.f[VY], o, ok, fi_vy);
            l_vz[lr][tx + 5] = gload<T, C>(s.f[VZ], o, ok, fi_vz);
        }
        __syncthreads();

        const bool pz = in_pml(k, g.nz, g.pml);
        const bool tile_pml = pz || tile_xpml || tile_ypml;
        const C A0 = s.dA[0][min(i, g.nx - 1)], B0 = s.dB[0][min(i, g.nx - 1)];
        const C A2 = s.dA[2][k], B2 = s.dB[2][k];
        const int s0 = k & 1, s1 = s0 ^ 1;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            const C dX = lbwd<C>(&l_vx[lr + 5][0], tx + 5, 1, s.cx);
            const C dY = lbwd<C>(&l_vy[0][tx + 5], lr + 5, XW, s.cy);
            const C dZ = qbwd<C>(qvz[r], 5, s.cz);
            const C dxy_x = lfwd<C>(&l_vy[lr + 5][0], tx + 5, 1, s.cx);
            const C dxy_y = lfwd<C>(&l_vx[0][tx + 5], lr + 5, XW, s.cy);
            const C dxz_x = lfwd<C>(&l_vz[lr + 5][0], tx + 5, 1, s.cx);
            const C dxz_z = qfwd<C>(qvx[r], 4, s.cz);
            const C dyz_y = lfwd<C>(&l_vz[0][tx + 5], lr + 5, XW, s.cy);
            const C dyz_z = qfwd<C>(qvy[r], 4, s.cz);
            const C r0 = l_rm[s0][lr][tx], rx = l_rm[s0][lr][tx + 1], ry = l_rm[s0][lr + 1][tx];
            const C rxy = l_rm[s0][lr + 1][tx + 1];
            const C rz = l_rm[s1][lr][tx], rxz = l_rm[s1][lr][tx + 1], ryz = l_rm[s1][lr + 1][tx];
            if (!(xup && yup[r])) continue;
            const long long o = i + jr[r] * sy + kz;
            const C lam = Cvt<T, C>::ld(s.lam[o], il);
            const C mu = muk[r];
            const C muxy = (C)4 / (((r0 + rx) + ry) + rxy);
            const C muxz = (C)4 / (((r0 + rx) + rz) + rxz);
            const C muyz = (C)4 / (((r0 + ry) + rz) + ryz);
            const C dl = s.dt * lam, dl2m = s.dt * (lam + (C)2 * mu);
            const C dmxy = s.dt * muxy, dmxz = s.dt * muxz, dmyz = s.dt * muyz;
            C old[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) old[c] = Cvt<T, C>::ld(s.f[SXX + c][o], s.finv[SXX + c]);
            if (k == src_k && o == s.src && s.src_saved) {
#pragma unroll
                for (int c = 0; c < 3; ++c) old[c] = Cvt<T, C>::ld(s.src_saved[c], s.finv[SXX + c]);
            }
            C nw[6];
            const bool py = in_pml(jr[r], g.ny, g.pml);
            if (!tile_pml || !(px || py || pz)) {
                nw[0] = old[0] + (dl2m * dX + dl * (dY + dZ));
                nw[1] = old[1] + (dl2m * dY + dl * (dX + dZ));
                nw[2] = old[2] + (dl2m * dZ + dl * (dX + dY));
                nw[3] = old[3] + dmxy * (dxy_x + dxy_y);
                nw[4] = old[4] + dmxz * (dxz_x + dxz_z);
                nw[5] = old[5] + dmyz * (dyz_y + dyz_z);
            } else {
                const long long slot = split_slot(g, i, jr[r], k);
                const bool dm[3] = {px, py, pz};
                const C A[3] = {A0, s.dA[1][jr[r]], A2};
                const C B[3] = {B0, s.dB[1][jr[r]], B2};
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
                    const C A2_[2] = {A[0], A[1]}, B2_[2] = {B[0], B[1]};
                    nw[3] = pml_update<C, 2>(old[3], inc, m2, sp, slot, A2_, B2_);
                }
                {
                    const C inc[2] = {dmxz * dxz_x, dmxz * dxz_z};
                    const bool m2[2] = {px, pz};
                    C *const sp[2] = {s.ss[XZ_X], s.ss[XZ_Z]};
                    const C A2_[2] = {A[0], A[2]}, B2_[2] = {B[0], B[2]};
                    nw[4] = pml_update<C, 2>(old[4], inc, m2, sp, slot, A2_, B2_);
                }
                {
                    const C inc[2] = {dmyz * dyz_y, dmyz * dyz_z};
                    const bool m2[2] = {py, pz};
                    C *const sp[2] = {s.ss[YZ_Y], s.ss[YZ_Z]};
                    const C A2_[2] = {A[1], A[2]}, B2_[2] = {B[1], B[2]};
                    nw[5] = pml_update<C, 2>(old[5], inc, m2, sp, slot, A2_, B2_);
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) s.f[SXX + c][o] = Cvt<T, C>::st(nw[c], s.fscale[SXX + c]);
        }
        __syncthreads();

        // ring: plane k+2 replaces plane k; queues advance
        if (k + 1 < k1) stage_rmu(k + 2);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const bool ok = xin && yin[r];
            const long long o = i + jr[r] * sy;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                qvx[r][q] = qvx[r][q + 1];
                qvy[r][q] = qvy[r][q + 1];
                qvz[r][q] = qvz[r][q + 1];
            }
            const bool okz6 = ok && (k + 6) < g.nz, okz5 = ok && (k + 5) < g.nz, okz1 = ok && (k + 1) < g.nz;
            qvx[r][9] = gload<T, C>(s.f[VX], o + (long long)(k + 6) * sz, okz6, fi_vx);
            qvy[r][9] = gload<T, C>(s.f[VY], o + (long long)(k + 6) * sz, okz6, fi_vy);
            qvz[r][9] = gload<T, C>(s.f[VZ], o + (long long)(k + 5) * sz, okz5, fi_vz);
            muk[r] = gload<T, C>(s.mu, o + (long long)(k + 1) * sz, okz1, im);
        }
    }
}

template <class T, class C, int NY>
__global__ __launch_bounds__(NTHR) void k_stress_tiled(const Shot<T, C> s, const TileMap m)
{
    __shared__ StrLds<C, NY> L;
    int bx, by, bz;
    if (!decode_block(m, bx, by, bz)) return;
    const int k0 = HALO + bz * m.zc;
    stress_tile<T, C, NY>(s, bx, by, k0, min(k0 + m.zc, s.g.nz - HALO), L);
}

// ---------------------------------------------------------------------------
// batched sweeps: a grid-stride loop over a cross-shot work-list.  Item
// positions are arranged on the host so that position p runs on XCD p % 8
// and each XCD sees a contiguous band of one shot's tiles.
// ---------------------------------------------------------------------------
template <int NY>
__global__ __launch_bounds__(NTHR) void k_velocity_batch(const Shot<float, float> *__restrict__ shots,
                                                         const WorkItem *__restrict__ items, int nitems)
{
    __shared__ VelLds<float, NY> L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot < 0) continue;
        velocity_tile<float, float, NY>(shots[w.shot], w.bx, w.by, w.k0, w.k1, L);
        __syncthreads();
    }
}

template <int NY>
__global__ __launch_bounds__(NTHR) void k_stress_batch(const Shot<float, float> *__restrict__ shots,
                                                       const WorkItem *__restrict__ items, int nitems)
{
    __shared__ StrLds<float, NY> L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot < 0) continue;
        stress_tile<float, float, NY>(shots[w.shot], w.bx, w.by, w.k0, w.k1, L);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
inline TileMap make_map(const Geom &g, int ty, int zchunk)
{
    TileMap m;
    m.nbx = (g.ux + TX - 1) / TX;
    m.nby = (g.uy + ty - 1) / ty;
    m.zc = zchunk > 0 ? zchunk : g.uz;
    m.nbz = (g.uz + m.zc - 1) / m.zc;
    return m;
}

inline int grid_blocks(const TileMap &m)
{
    const int n = m.nbx * m.nby * m.nbz;
    return (n + 7) / 8 * 8;
}

// default z chunk: enough blocks for ~4 per CU on 256 CUs
inline int auto_zchunk(const Geom &g, int ty)
{
    const int cols = ((g.ux + TX - 1) / TX) * ((g.uy + ty - 1) / ty);
    int nbz = (1024 + cols - 1) / cols;
    if (nbz < 1) nbz = 1;
    int zc = (g.uz + nbz - 1) / nbz;
    if (zc < 16) zc = 16;
    return zc;
}

template <class T, class C>
hipError_t velocity_tiled(const Shot<T, C> &s, const Tile &t, hipStream_t st)
{
    const int ty = t.ty;
    const TileMap m = make_map(s.g, ty, t.zchunk > 0 ? t.zchunk : auto_zchunk(s.g, ty));
    const dim3 block(TX, 4);
    const dim3 grid(grid_blocks(m));
    if (ty == 4) hipLaunchKernelGGL((k_velocity_tiled<T, C, 1>), grid, block, 0, st, s, m);
    else if (ty == 8) hipLaunchKernelGGL((k_velocity_tiled<T, C, 2>), grid, block, 0, st, s, m);
    else hipLaunchKernelGGL((k_velocity_tiled<T, C, 4>), grid, block, 0, st, s, m);
    return hipGetLastError();
}

template <class T, class C>
hipError_t stress_tiled(const Shot<T, C> &s, const Tile &t, hipStream_t st)
{
    const int ty = t.ty;
    const TileMap m = make_map(s.g, ty, t.zchunk > 0 ? t.zchunk : auto_zchunk(s.g, ty));
    const dim3 block(TX, 4);
    const dim3 grid(grid_blocks(m));
    if (ty == 4) hipLaunchKernelGGL((k_stress_tiled<T, C, 1>), grid, block, 0, st, s, m);
    else if (ty == 8) hipLaunchKernelGGL((k_stress_tiled<T, C, 2>), grid, block, 0, st, s, m);
    else hipLaunchKernelGGL((k_stress_tiled<T, C, 4>), grid, block, 0, st, s, m);
    return hipGetLastError();
}

hipError_t launch_velocity_batch(const void *shots_dev, const WorkItem *items, int nitems, int grid, int ty,
                                 hipStream_t st)
{
    const Shot<float, float> *sh = static_cast<const Shot<float, float> *>(shots_dev);
    const dim3 block(TX, 4);
    if (ty == 4) hipLaunchKernelGGL((k_velocity_batch<1>), dim3(grid), block, 0, st, sh, items, nitems);
    else if (ty == 8) hipLaunchKernelGGL((k_velocity_batch<2>), dim3(grid), block, 0, st, sh, items, nitems);
    else hipLaunchKernelGGL((k_velocity_batch<4>), dim3(grid), block, 0, st, sh, items, nitems);
    return hipGetLastError();
}

hipError_t launch_stress_batch(const void *shots_dev, const WorkItem *items, int nitems, int grid, int ty,
                               hipStream_t st)
{
    const Shot<float, float> *sh = static_cast<const Shot<float, float> *>(shots_dev);
    const dim3 block(TX, 4);
    if (ty == 4) hipLaunchKernelGGL((k_stress_batch<1>), dim3(grid), block, 0, st, sh, items, nitems);
    else if (ty == 8) hipLaunchKernelGGL((k_stress_batch<2>), dim3(grid), block, 0, st, sh, items, nitems);
    else hipLaunchKernelGGL((k_stress_batch<4>), dim3(grid), block, 0, st, sh, items, nitems);
    return hipGetLastError();
}

}  // namespace sg
