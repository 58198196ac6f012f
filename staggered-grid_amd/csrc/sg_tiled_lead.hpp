// sg_tiled_lead.hpp -- the tiled sweeps with the x/y staging of the
// z-queued fields LD planes ahead of the plane being updated (included by
// sg_tiled.hpp; same boxes, work-lists and arithmetic as the sweeps there).
//
// A field with a z derivative sits in a register queue whose newest plane
// is k + 5 while plane k is updated.  Staging its plane k for the x/y
// derivatives would read that plane's halo rows again five plane-steps after
// the neighbouring boxes read them as their queue heads, by when the L2 has
// usually dropped them, so every halo row would come from the Infinity
// Cache or HBM a second time.  Here the x/y halos of such a field are read
// together with the queue head (plane k + 5, the "lead" plane), its x/y
// derivatives are taken there and kept in an LDS ring of LD planes (each
// thread its own entries) until plane k is updated.  Every plane is still
// staged once; the first LD planes of a box are staged before its first
// update (warm-up, loads two planes ahead).
// The derivatives are the same operations on the same values as in the
// plain sweeps, so results are bit-identical.
#pragma once

namespace sg {

constexpr int LD = 5;  // lead of the x/y staging over the updated plane

// LDS of the lead form: the plain sweep's staging plus the delay ring of
// the x/y derivatives (ND values per thread and plane, ring slot = plane %
// LD); each thread reads and writes only its own entries.  One row per
// thread, TY rows per box.
template <class C, int TY>
struct VelLdsL : VelLds<C, TY> {
    C dl[LD][2][TX * TY];
};
template <class C, int TY, int ND>
struct StrLdsL : StrLds<C, TY> {
    C dl[LD][ND][TX * TY];
};

// ---------------------------------------------------------------------------
// velocity sweep: sxz (x derivative) and syz (y derivative) at the lead
// plane; sxx, sxy, syy (no z derivative) and b at the updated plane
// ---------------------------------------------------------------------------
template <class T, class C, int TY, int NY, int DM, bool CP, class LT>
__device__ __forceinline__ void velocity_tile_ld(const Shot<T, C> &s, int i0, int i1, int j0, int j1, int k0,
                                                 int k1, LT &L)
{
    static_assert(NY == 1, "the lead form keeps one row per thread");
    constexpr int BY = TY / NY, NH = TileGeo<TY, NY>::NH;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long long sz = g.sz;
    TileGeo<TY, NY> t;
    tile_geo<TY, NY, (int)sizeof(T)>(g, i0, i1, j0, j1, t);
    using G = Stg<T, C>;
    using S = typename G::S;

    const T *Fxx = s.f[SXX], *Fyy = s.f[SYY], *Fzz = s.f[SZZ];
    const T *Fxy = s.f[SXY], *Fxz = s.f[SXZ], *Fyz = s.f[SYZ];
    const T *Fb = s.b;
    const float fi_xx = s.finv[SXX], fi_yy = s.finv[SYY], fi_zz = s.finv[SZZ];
    const float fi_xy = s.finv[SXY], fi_xz = s.finv[SXZ], fi_yz = s.finv[SYZ];
    const float ib = 1.0f / s.bscale;

    // register z queues: qxz/qyz[m] = plane k-5+m (up to the lead plane
    // k+5), qzz[m] = plane k-4+m, qb[m] = plane k+m
    C qxz[NY][11], qyz[NY][11], qzz[NY][10], qb[NY][2];
#pragma unroll
    for (int q = 0; q < 11; ++q) {
        const T *pxz = plane(Fxz, k0 - 5 + q, sz), *pyz = plane(Fyz, k0 - 5 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            qxz[r][q] = mld<T, C>(pxz, t.ro[r], t.rin[r], fi_xz);
            qyz[r][q] = mld<T, C>(pyz, t.ro[r], t.rin[r], fi_yz);
        }
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const T *pzz = plane(Fzz, k0 - 4 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) qzz[r][q] = mld<T, C>(pzz, t.ro[r], t.rin[r], fi_zz);
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        qb[r][0] = mld<T, C>(plane(Fb, k0, sz), t.ro[r], t.rin[r], ib);
        qb[r][1] = mld<T, C>(plane(Fb, k0 + 1, sz), t.ro[r], t.rin[r], ib);
    }

    // lead-plane staging: x halo of sxz, y halo of syz (staging registers,
    // converted when staged)
    struct Halo {
        S xz[NY], yz[NH];
    };
    auto fetch_lead = [&](int k, Halo &h) {
        const T *pxz = plane(Fxz, k, sz), *pyz = plane(Fyz, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) h.xz[r] = G::ld(pxz, t.hro[r], t.hok[r], fi_xz);
#pragma unroll
        for (int m = 0; m < NH; ++m) h.yz[m] = G::ld(pyz, t.yro[m], t.yok[m], fi_yz);
    };
    // centres of the lead plane come from the queues (index qi)
    auto stage_lead = [&](int qi, const Halo &h) {
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            L.sxz[lr][tx + 5] = qxz[r][qi];
            L.syz[lr + 5][tx] = qyz[r][qi];
        }
        if (tx < 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) L.sxz[ty * NY + r][t.hc] = G::use(h.xz[r], t.hok[r], fi_xz);
        }
#pragma unroll
        for (int m = 0; m < NH; ++m)
            if (ty + BY * m < 10) L.syz[t.ylr[m]][tx] = G::use(h.yz[m], t.yok[m], fi_yz);
    };
    // D-x sxz and D-y syz of the staged lead plane, into ring slot sl
    const int tid = ty * TX + tx;
    auto lead_derivs = [&](int sl) {
        const int lr = ty;
        L.dl[sl][0][tid] = lbwd<C>(&L.sxz[lr][0], tx + 5, 1, s.cx);
        L.dl[sl][1][tid] = lbwd<C>(&L.syz[0][tx], lr + 5, TX, s.cy);
    };

    // warm-up: x/y derivatives of the first LD planes into the ring, the
    // halos loaded two planes ahead (the first lead plane's into hn)
    Halo hw[2], hn;
    fetch_lead(k0, hw[0]);
    if (k0 + 1 < k1) fetch_lead(k0 + 1, hw[1]);
#pragma unroll
    for (int m = 0; m < LD; ++m) {
        if (k0 + m < k1) {
            stage_lead(5 + m, hw[m & 1]);
            if (k0 + m + 2 < k1) {
                if (m + 2 < LD) fetch_lead(k0 + m + 2, hw[m & 1]);
                else if (m + 2 == LD) fetch_lead(k0 + LD, hn);
            }
            __syncthreads();
            lead_derivs((k0 + m) % LD);
            __syncthreads();
        }
    }

    // updated-plane staging registers (converted when staged)
    S c_xx[NY], c_xy[NY], c_yy[NY], h_xx[NY], h_xy[NY], b_x[NY];
    S y_yy[NH], y_xy[NH], b_y;
    auto fetch = [&](int k) {
        const T *pxx = plane(Fxx, k, sz), *pxy = plane(Fxy, k, sz), *pyy = plane(Fyy, k, sz);
        const T *pb = plane(Fb, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            c_xx[r] = G::ld(pxx, t.ro[r], t.rin[r], fi_xx);
            c_xy[r] = G::ld(pxy, t.ro[r], t.rin[r], fi_xy);
            c_yy[r] = G::ld(pyy, t.ro[r], t.rin[r], fi_yy);
            h_xx[r] = G::ld(pxx, t.hro[r], t.hok[r], fi_xx);
            h_xy[r] = G::ld(pxy, t.hro[r], t.hok[r], fi_xy);
            b_x[r] = G::ld(pb, t.bxo[r], t.bxok && t.rin[r], ib);
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            y_yy[m] = G::ld(pyy, t.yro[m], t.yok[m], fi_yy);
            y_xy[m] = G::ld(pxy, t.yro[m], t.yok[m], fi_xy);
        }
        b_y = G::ld(pb, t.byo, t.byok, ib);
    };
    fetch(k0);

    // old velocities (stored values and residuals), converted where used
    auto load_rmw = [&](int kk, T (&out)[3][NY], T (&res)[3][NY]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T *p = plane(s.f[VX + c], kk, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) out[c][r] = bld_nt(p, t.ro[r]);
        }
        if (CP) {
#pragma unroll
            for (int c = 0; c < 3; ++c)
                if (s.comp[VX + c]) {
                    const T *pc = plane(s.comp[VX + c], kk, sz);
#pragma unroll
                    for (int r = 0; r < NY; ++r) res[c][r] = bld_nt(pc, t.ro[r]);
                }
        }
    };
    auto rmw_value = [&](const T (&raw)[3][NY], const T (&res)[3][NY], int c, int r) -> C {
        C v = Cvt<T, C>::ld(raw[c][r], s.finv[VX + c]);
        if (CP && s.comp[VX + c]) v += Cvt<T, C>::ld(res[c][r], s.cinv[VX + c]);
        return v;
    };
    constexpr bool RP = DM == 0 || sizeof(T) == 2;
    T vpre[3][NY], vpres[3][NY];
    if (RP) load_rmw(k0, vpre, vpres);

    constexpr bool PML = DM != 0, DX = (DM & 1) != 0, DY = (DM & 2) != 0, DZ = (DM & 4) != 0;
    const bool px = DX && in_pml(t.i, g.nx, g.pml);
    const bool py = DY && in_pml(j0, g.ny, g.pml);
    C A0 = (C)1, B0 = (C)1, A1[NY], B1[NY];
    const uint32_t step = (sizeof(T) == 2 && s.round_mode) ? (uint32_t)gld(s.d_step, 0u) : 0u;
    if (DX) {
        A0 = gld(s.dA[0], (unsigned)(t.ic));
        B0 = gld(s.dB[0], (unsigned)(t.ic));
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const unsigned jc = (unsigned)(t.jr[r] < g.ny ? t.jr[r] : g.ny - 1);
        A1[r] = DY ? gld(s.dA[1], jc) : (C)1;
        B1[r] = DY ? gld(s.dB[1], jc) : (C)1;
    }

    for (int k = k0; k < k1; ++k) {
        const bool lead = k + LD < k1;
        // ---- stage plane k (sxx, sxy, syy, b) and the lead plane k+LD
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            L.sxx[lr][tx + 5] = G::use(c_xx[r], t.rin[r], fi_xx);
            L.sxy[lr + 5][tx + 5] = G::use(c_xy[r], t.rin[r], fi_xy);
            L.syy[lr + 5][tx] = G::use(c_yy[r], t.rin[r], fi_yy);
            L.b[lr][tx] = qb[r][0];
        }
        if (tx < 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const int lr = ty * NY + r;
                L.sxx[lr][t.hc] = G::use(h_xx[r], t.hok[r], fi_xx);
                L.sxy[lr + 5][t.hc] = G::use(h_xy[r], t.hok[r], fi_xy);
            }
        }
        if (tx == 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) L.b[ty * NY + r][TX] = G::use(b_x[r], t.bxok && t.rin[r], ib);
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            if (ty + BY * m < 10) {
                L.syy[t.ylr[m]][tx] = G::use(y_yy[m], t.yok[m], fi_yy);
                L.sxy[t.ylr[m]][tx + 5] = G::use(y_xy[m], t.yok[m], fi_xy);
            }
        }
        if (ty == 0) L.b[TY][tx] = G::use(b_y, t.byok, ib);
        if (lead) stage_lead(10, hn);
        __syncthreads();

        // ---- read half of the read-modify-write of plane k
        T *pv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) pv[c] = plane(s.f[VX + c], k, sz);
        C vold[3][NY];
        if (RP) {
#pragma unroll
            for (int r = 0; r < NY; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) vold[c][r] = rmw_value(vpre, vpres, c, r);
        } else {
            T vr[3][NY], vrs[3][NY];
            load_rmw(k, vr, vrs);
#pragma unroll
            for (int r = 0; r < NY; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) vold[c][r] = rmw_value(vr, vrs, c, r);
        }
        bool pz = false, any[NY];
        C A2 = (C)1, B2 = (C)1, so[NY][3][3], vsr[NY];
        unsigned sb[NY];
        if (PML) {
            if (DZ) {
                pz = in_pml(k, g.nz, g.pml);
                A2 = gld(s.dA[2], (unsigned)(k));
                B2 = gld(s.dB[2], (unsigned)(k));
            }
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                any[r] = (px || py || pz) && t.rup[r];
                sb[r] = any[r] ? (unsigned)(split_slot(g, t.i, t.jr[r], k) * (long long)sizeof(C)) : 0u;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int a2 = 0; a2 < 3; ++a2) so[r][c][a2] = (C)0;
                vsr[r] = (s.vs && any[r]) ? bld(s.vs, sb[r]) : (C)0;
                if (any[r]) {
                    if (DX && px) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) so[r][c][0] = bld(s.sv[3 * c], sb[r]);
                    }
                    if (DY && py) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) so[r][c][1] = bld(s.sv[3 * c + 1], sb[r]);
                    }
                    if (DZ && pz) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) so[r][c][2] = bld(s.sv[3 * c + 2], sb[r]);
                    }
                }
            }
        }

        // ---- loads for later, in the order their values are needed: RMW
        // operands of plane k+1, queue heads (plane k+6, pushed after this
        // plane), then the staging values of plane k+1 and of the lead plane
        // k+LD+1 (staged after the next barrier).  The vector-memory counter
        // retires in issue order.  Past the end of the box the loads repeat
        // the box's last plane (cached) instead of being skipped, so no
        // branch makes the compiler's waits count them conservatively.
        const int kn = k + 1 < k1 ? k + 1 : k;
        if (RP) load_rmw(kn, vpre, vpres);
        S nxz[NY], nyz[NY], nzz[NY], nb[NY];
        const bool z6 = k + 6 < g.nz, z2 = k + 2 < g.nz;
        {
            const T *pxz = plane(Fxz, z6 ? k + 6 : k, sz), *pyz = plane(Fyz, z6 ? k + 6 : k, sz);
            const T *pzz = plane(Fzz, z6 ? k + 6 : k, sz), *pb = plane(Fb, z2 ? k + 2 : k, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                nxz[r] = G::ld(pxz, t.ro[r], t.rin[r] && z6, fi_xz);
                nyz[r] = G::ld(pyz, t.ro[r], t.rin[r] && z6, fi_yz);
                nzz[r] = G::ld(pzz, t.ro[r], t.rin[r] && z6, fi_zz);
                nb[r] = G::ld(pb, t.ro[r], t.rin[r] && z2, ib);
            }
        }
        fetch(kn);
        fetch_lead(k + LD + 1 < k1 ? k + LD + 1 : k1 - 1, hn);

        // ---- compute plane k (its delayed derivatives from ring slot sl,
        // which then takes those of the lead plane)
        const int sl = k % LD;
        const C dxz_x = L.dl[sl][0][tid], dyz_y = L.dl[sl][1][tid];
        if (lead) lead_derivs(sl);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            C d[3][3];
            d[0][0] = lfwd<C>(&L.sxx[lr][0], tx + 5, 1, s.cx);
            d[0][1] = lbwd<C>(&L.sxy[0][tx + 5], lr + 5, XW, s.cy);
            d[0][2] = lbwd<C>(qxz[r], 5, 1, s.cz);
            d[1][0] = lbwd<C>(&L.sxy[lr + 5][0], tx + 5, 1, s.cx);
            d[1][1] = lfwd<C>(&L.syy[0][tx], lr + 5, TX, s.cy);
            d[1][2] = lbwd<C>(qyz[r], 5, 1, s.cz);
            d[2][0] = dxz_x;
            d[2][1] = dyz_y;
            d[2][2] = lfwd<C>(qzz[r], 4, 1, s.cz);
            const C b0 = L.b[lr][tx];
            C db[3];
            db[0] = s.dt * ((b0 + L.b[lr][tx + 1]) * (C)0.5);
            db[1] = s.dt * ((b0 + L.b[lr + 1][tx]) * (C)0.5);
            db[2] = s.dt * ((b0 + qb[r][1]) * (C)0.5);
            C vn[3], sn[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) vn[c] = fmad(db[c], (d[c][0] + d[c][1]) + d[c][2], vold[c][r]);
            if (PML) {
                const bool ar = any[r];
                const bool dm[3] = {ar && px, ar && py, ar && pz};
                C A[3] = {A0, A1[r], A2};
                C B[3] = {B0, B1[r], B2};
                if (s.vs) vs_damping(vsr[r], A, B);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const C inc[3] = {db[c] * d[c][0], db[c] * d[c][1], db[c] * d[c][2]};
                    const C vp = pml_combine<C, 3>(vold[c][r], inc, dm, so[r][c], A, B, sn[c]);
                    vn[c] = ar ? vp : vn[c];
                }
            }
            if (t.rup[r]) {
                if (CP) {
                    store_comp<T, C>(s, VX, k, t.ro[r], vn[0], pv[0]);
                    store_comp<T, C>(s, VY, k, t.ro[r], vn[1], pv[1]);
                    store_comp<T, C>(s, VZ, k, t.ro[r], vn[2], pv[2]);
                } else if (sizeof(T) == 2 && s.round_mode) {
                    const uint32_t pt = t.ro[r] / (unsigned)sizeof(T) + (uint32_t)k * (uint32_t)sz;
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        bst_nt(pv[c], t.ro[r], CvtSR<T, C>::st(vn[c], s.fscale[VX + c], sr_bits(pt, step, VX + c)));
                } else {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst_nt(pv[c], t.ro[r], Cvt<T, C>::st(vn[c], s.fscale[VX + c]));
                }
            }
            if (PML && any[r]) {
                if (DX && px) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.sv[3 * c], sb[r], sn[c][0]);
                }
                if (DY && py) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.sv[3 * c + 1], sb[r], sn[c][1]);
                }
                if (DZ && pz) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.sv[3 * c + 2], sb[r], sn[c][2]);
                }
            }
        }
        __syncthreads();

        // ---- advance the z queues
#pragma unroll
        for (int r = 0; r < NY; ++r) {
#pragma unroll
            for (int q = 0; q < 10; ++q) {
                qxz[r][q] = qxz[r][q + 1];
                qyz[r][q] = qyz[r][q + 1];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) qzz[r][q] = qzz[r][q + 1];
            qxz[r][10] = G::use(nxz[r], t.rin[r] && z6, fi_xz);
            qyz[r][10] = G::use(nyz[r], t.rin[r] && z6, fi_yz);
            qzz[r][9] = G::use(nzz[r], t.rin[r] && z6, fi_zz);
            qb[r][0] = qb[r][1];
            qb[r][1] = G::use(nb[r], t.rin[r] && z2, ib);
        }
    }
}

// ---------------------------------------------------------------------------
// stress sweep: all three velocities staged at the lead plane; 1/mu ring,
// lambda, mu and the read-modify-write at the updated plane
// ---------------------------------------------------------------------------
template <class T, class C, int TY, int NY, int DM, bool CP, class LT>
__device__ __forceinline__ void stress_tile_ld(const Shot<T, C> &s, int i0, int i1, int j0, int j1, int k0, int k1,
                                               LT &L)
{
    static_assert(NY == 1, "the lead form keeps one row per thread");
    constexpr int BY = TY / NY, NH = TileGeo<TY, NY>::NH;
    // delayed x/y derivatives: D-x vx, D-y vy, D+x vy, D+y vx, D+x vz, D+y vz;
    // interior boxes keep D+x vy + D+y vx as one value (the update uses only
    // the sum)
    constexpr int ND = DM == 0 ? 5 : 6;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long long sz = g.sz;
    TileGeo<TY, NY> t;
    tile_geo<TY, NY, (int)sizeof(T)>(g, i0, i1, j0, j1, t);
    using G = Stg<T, C>;
    using S = typename G::S;

    const T *Fvx = s.f[VX], *Fvy = s.f[VY], *Fvz = s.f[VZ];
    const T *Fmu = s.mu, *Flam = s.lam;
    const float fi_vx = s.finv[VX], fi_vy = s.finv[VY], fi_vz = s.finv[VZ];
    const float il = 1.0f / s.lscale, im = 1.0f / s.mscale;

    // queues: qvx/qvy[m] = plane k-4+m, qvz[m] = plane k-5+m; the newest
    // entry of each is the lead plane k+5
    C qvx[NY][10], qvy[NY][10], qvz[NY][11];
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const T *px_ = plane(Fvx, k0 - 4 + q, sz), *py_ = plane(Fvy, k0 - 4 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            qvx[r][q] = mld<T, C>(px_, t.ro[r], t.rin[r], fi_vx);
            qvy[r][q] = mld<T, C>(py_, t.ro[r], t.rin[r], fi_vy);
        }
    }
#pragma unroll
    for (int q = 0; q < 11; ++q) {
        const T *pz_ = plane(Fvz, k0 - 5 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) qvz[r][q] = mld<T, C>(pz_, t.ro[r], t.rin[r], fi_vz);
    }

    // mu staging registers and the 1/mu ring of planes k and k+1; mc: mu
    // of the centre rows
    S m_c[NY], m_x[NY], m_y, m_xy;
    auto fetch_mu = [&](int k) {
        const T *pm = plane(Fmu, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            m_c[r] = G::ld(pm, t.ro[r], t.rin[r], im);
            m_x[r] = G::ld(pm, t.bxo[r], t.bxok && t.rin[r], im);
        }
        m_y = G::ld(pm, t.byo, t.byok, im);
        m_xy = G::ld(pm, t.bxyo, t.bxyok, im);
    };
    auto stage_mu = [&](int k, C (&mc)[NY]) {
        const int sl = k & 1;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            mc[r] = G::use(m_c[r], t.rin[r], im);
            L.rm[sl][ty * NY + r][tx] = t.rin[r] ? (C)1 / mc[r] : (C)0;
        }
        if (tx == 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const bool ok = t.bxok && t.rin[r];
                L.rm[sl][ty * NY + r][TX] = ok ? (C)1 / G::use(m_x[r], ok, im) : (C)0;
            }
        }
        if (ty == 0) {
            L.rm[sl][TY][tx] = t.byok ? (C)1 / G::use(m_y, t.byok, im) : (C)0;
            if (tx == 10) L.rm[sl][TY][TX] = t.bxyok ? (C)1 / G::use(m_xy, t.bxyok, im) : (C)0;
        }
    };
    C muk[NY], mun[NY];
    fetch_mu(k0);
    stage_mu(k0, muk);
    fetch_mu(k0 + 1);
    stage_mu(k0 + 1, mun);

    // lead-plane staging: x and y halos of the three velocities
    struct Halo {
        S hx[NY], hy[NY], hz[NY], yx[NH], yy[NH], yz[NH];
    };
    auto fetch = [&](int k, Halo &h) {
        const T *pvx = plane(Fvx, k, sz), *pvy = plane(Fvy, k, sz), *pvz = plane(Fvz, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            h.hx[r] = G::ld(pvx, t.hro[r], t.hok[r], fi_vx);
            h.hy[r] = G::ld(pvy, t.hro[r], t.hok[r], fi_vy);
            h.hz[r] = G::ld(pvz, t.hro[r], t.hok[r], fi_vz);
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            h.yx[m] = G::ld(pvx, t.yro[m], t.yok[m], fi_vx);
            h.yy[m] = G::ld(pvy, t.yro[m], t.yok[m], fi_vy);
            h.yz[m] = G::ld(pvz, t.yro[m], t.yok[m], fi_vz);
        }
    };
    // centres from the queues: vx, vy at index qa, vz at index qa + 1
    auto stage = [&](int qa, const Halo &h) {
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            L.vx[lr + 5][tx + 5] = qvx[r][qa];
            L.vy[lr + 5][tx + 5] = qvy[r][qa];
            L.vz[lr + 5][tx + 5] = qvz[r][qa + 1];
        }
        if (tx < 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const int lr = ty * NY + r;
                L.vx[lr + 5][t.hc] = G::use(h.hx[r], t.hok[r], fi_vx);
                L.vy[lr + 5][t.hc] = G::use(h.hy[r], t.hok[r], fi_vy);
                L.vz[lr + 5][t.hc] = G::use(h.hz[r], t.hok[r], fi_vz);
            }
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            if (ty + BY * m < 10) {
                L.vx[t.ylr[m]][tx + 5] = G::use(h.yx[m], t.yok[m], fi_vx);
                L.vy[t.ylr[m]][tx + 5] = G::use(h.yy[m], t.yok[m], fi_vy);
                L.vz[t.ylr[m]][tx + 5] = G::use(h.yz[m], t.yok[m], fi_vz);
            }
        }
    };
    // x/y derivatives of the staged lead plane into ring slot sl
    const int tid = ty * TX + tx;
    auto lead_derivs = [&](int sl) {
        const int lr = ty;
        L.dl[sl][0][tid] = lbwd<C>(&L.vx[lr + 5][0], tx + 5, 1, s.cx);   // D-x vx
        L.dl[sl][1][tid] = lbwd<C>(&L.vy[0][tx + 5], lr + 5, XW, s.cy);  // D-y vy
        const C xy_x = lfwd<C>(&L.vy[lr + 5][0], tx + 5, 1, s.cx);
        const C xy_y = lfwd<C>(&L.vx[0][tx + 5], lr + 5, XW, s.cy);
        L.dl[sl][2][tid] = lfwd<C>(&L.vz[lr + 5][0], tx + 5, 1, s.cx);   // D+x vz
        L.dl[sl][3][tid] = lfwd<C>(&L.vz[0][tx + 5], lr + 5, XW, s.cy);  // D+y vz
        if constexpr (ND == 5) {
            L.dl[sl][4][tid] = xy_x + xy_y;
        } else {
            L.dl[sl][4][tid] = xy_x;
            L.dl[sl][ND - 1][tid] = xy_y;
        }
    };

    // warm-up: x/y derivatives of the first LD planes into the ring, the
    // halos loaded two planes ahead (the first lead plane's into hn)
    Halo hw[2], hn;
    fetch(k0, hw[0]);
    if (k0 + 1 < k1) fetch(k0 + 1, hw[1]);
#pragma unroll
    for (int m = 0; m < LD; ++m) {
        if (k0 + m < k1) {
            stage(4 + m, hw[m & 1]);
            if (k0 + m + 2 < k1) {
                if (m + 2 < LD) fetch(k0 + m + 2, hw[m & 1]);
                else if (m + 2 == LD) fetch(k0 + LD, hn);
            }
            __syncthreads();
            lead_derivs((k0 + m) % LD);
            __syncthreads();
        }
    }

    // old stresses and lambda (stored values and residuals), converted where
    // used
    auto load_rmw = [&](int kk, T (&out)[6][NY], T (&res)[6][NY], T (&lo)[NY]) {
        const T *pl = plane(Flam, kk, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) lo[r] = bld_nt(pl, t.ro[r]);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const T *p = plane(s.f[SXX + c], kk, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) out[c][r] = bld_nt(p, t.ro[r]);
        }
        if (CP) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
                if (s.comp[SXX + c]) {
                    const T *pc = plane(s.comp[SXX + c], kk, sz);
#pragma unroll
                    for (int r = 0; r < NY; ++r) res[c][r] = bld_nt(pc, t.ro[r]);
                }
        }
    };
    auto rmw_value = [&](const T (&raw)[6][NY], const T (&res)[6][NY], int c, int r) -> C {
        C v = Cvt<T, C>::ld(raw[c][r], s.finv[SXX + c]);
        if (CP && s.comp[SXX + c]) v += Cvt<T, C>::ld(res[c][r], s.cinv[SXX + c]);
        return v;
    };
    constexpr bool RP = DM == 0 || sizeof(T) == 2;
    T spre[6][NY], spres[6][NY], lpre[NY];
    if (RP) load_rmw(k0, spre, spres, lpre);

    constexpr bool PML = DM != 0, DX = (DM & 1) != 0, DY = (DM & 2) != 0, DZ = (DM & 4) != 0;
    const bool px = DX && in_pml(t.i, g.nx, g.pml);
    const bool py = DY && in_pml(j0, g.ny, g.pml);
    C A0 = (C)1, B0 = (C)1, A1[NY], B1[NY];
    const uint32_t step = (sizeof(T) == 2 && s.round_mode) ? (uint32_t)gld(s.d_step, 0u) : 0u;
    if (DX) {
        A0 = gld(s.dA[0], (unsigned)(t.ic));
        B0 = gld(s.dB[0], (unsigned)(t.ic));
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const unsigned jc = (unsigned)(t.jr[r] < g.ny ? t.jr[r] : g.ny - 1);
        A1[r] = DY ? gld(s.dA[1], jc) : (C)1;
        B1[r] = DY ? gld(s.dB[1], jc) : (C)1;
    }
    const int src_k = s.src >= 0 ? (int)(s.src / sz) : -1;
    const unsigned src_o = s.src >= 0 ? (unsigned)(s.src - (long long)src_k * sz) * (unsigned)sizeof(T) : 0u;

    for (int k = k0; k < k1; ++k) {
        const bool lead = k + LD < k1;
        if (lead) stage(9, hn);
        __syncthreads();

        // ---- read half of the read-modify-write of plane k
        C lam[NY];
        T *ps[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) ps[c] = plane(s.f[SXX + c], k, sz);
        C old[6][NY];
        if (RP) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                lam[r] = Cvt<T, C>::ld(lpre[r], il);
#pragma unroll
                for (int c = 0; c < 6; ++c) old[c][r] = rmw_value(spre, spres, c, r);
            }
        } else {
            T sr[6][NY], srs[6][NY], lr_[NY];
            load_rmw(k, sr, srs, lr_);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                lam[r] = Cvt<T, C>::ld(lr_[r], il);
#pragma unroll
                for (int c = 0; c < 6; ++c) old[c][r] = rmw_value(sr, srs, c, r);
            }
        }
        if (k == src_k && s.src_saved) {
#pragma unroll
            for (int r = 0; r < NY; ++r)
                if (t.rup[r] && t.ro[r] == src_o)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        old[c][r] = Cvt<T, C>::ld(gld(s.src_saved, (unsigned)c), s.finv[SXX + c]);
                        if (CP && s.comp[SXX + c])
                            old[c][r] += Cvt<T, C>::ld(gld(s.src_saved, (unsigned)(3 + c)), s.cinv[SXX + c]);
                    }
        }
        bool pz = false, any[NY];
        C A2 = (C)1, B2 = (C)1;
        C sn3[NY][3][3], s2[NY][3][2];
        C vsr[NY];
        unsigned sb[NY];
        if (PML) {
            if (DZ) {
                pz = in_pml(k, g.nz, g.pml);
                A2 = gld(s.dA[2], (unsigned)(k));
                B2 = gld(s.dB[2], (unsigned)(k));
            }
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                any[r] = (px || py || pz) && t.rup[r];
                sb[r] = any[r] ? (unsigned)(split_slot(g, t.i, t.jr[r], k) * (long long)sizeof(C)) : 0u;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
#pragma unroll
                    for (int a2 = 0; a2 < 3; ++a2) sn3[r][c][a2] = (C)0;
#pragma unroll
                    for (int a2 = 0; a2 < 2; ++a2) s2[r][c][a2] = (C)0;
                }
                vsr[r] = (s.vs && any[r]) ? bld(s.vs, sb[r]) : (C)0;
                if (any[r]) {
                    const unsigned o = sb[r];
                    if (DX && px) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) sn3[r][c][0] = bld(s.ss[XX_X + 3 * c], o);
                        s2[r][0][0] = bld(s.ss[XY_X], o);
                        s2[r][1][0] = bld(s.ss[XZ_X], o);
                    }
                    if (DY && py) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) sn3[r][c][1] = bld(s.ss[XX_X + 3 * c + 1], o);
                        s2[r][0][1] = bld(s.ss[XY_Y], o);
                        s2[r][2][0] = bld(s.ss[YZ_Y], o);
                    }
                    if (DZ && pz) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) sn3[r][c][2] = bld(s.ss[XX_X + 3 * c + 2], o);
                        s2[r][1][1] = bld(s.ss[XZ_Z], o);
                        s2[r][2][1] = bld(s.ss[YZ_Z], o);
                    }
                }
            }
        }

        // ---- loads for later, in the order their values are needed: RMW
        // operands of plane k+1, queue heads (plane k+6) and mu of plane k+2
        // (used after this plane), the next lead plane's halos (staged after
        // the next barrier).  The vector-memory counter retires in issue
        // order.  Past the end of the box the loads repeat planes the box
        // has read (cached) instead of being skipped, so no branch makes the
        // compiler's waits count them conservatively.
        const bool more = k + 1 < k1;
        const int kn = more ? k + 1 : k;
        if (RP) load_rmw(kn, spre, spres, lpre);
        S nvx[NY], nvy[NY], nvz[NY];
        const bool z6 = k + 6 < g.nz;
        {
            const T *px_ = plane(Fvx, z6 ? k + 6 : k, sz), *py_ = plane(Fvy, z6 ? k + 6 : k, sz);
            const T *pz_ = plane(Fvz, z6 ? k + 6 : k, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                nvx[r] = G::ld(px_, t.ro[r], t.rin[r] && z6, fi_vx);
                nvy[r] = G::ld(py_, t.ro[r], t.rin[r] && z6, fi_vy);
                nvz[r] = G::ld(pz_, t.ro[r], t.rin[r] && z6, fi_vz);
            }
        }
        fetch_mu(kn + 1);
        fetch(k + LD + 1 < k1 ? k + LD + 1 : k1 - 1, hn);

        // the delayed derivatives of plane k from ring slot sl, which then
        // takes those of the lead plane
        const int sl = k % LD;
        C dk[ND];
#pragma unroll
        for (int e = 0; e < ND; ++e) dk[e] = L.dl[sl][e][tid];
        if (lead) lead_derivs(sl);

        const int s0 = k & 1, s1 = s0 ^ 1;
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            const C dX = dk[0], dY = dk[1];
            const C dxz_x = dk[2], dyz_y = dk[3];
            const C dZ = lbwd<C>(qvz[r], 5, 1, s.cz);
            const C dxz_z = lfwd<C>(qvx[r], 4, 1, s.cz);
            const C dyz_z = lfwd<C>(qvy[r], 4, 1, s.cz);
            const C r0 = L.rm[s0][lr][tx], rx = L.rm[s0][lr][tx + 1], ry = L.rm[s0][lr + 1][tx];
            const C rxy = L.rm[s0][lr + 1][tx + 1];
            const C rz = L.rm[s1][lr][tx], rxz = L.rm[s1][lr][tx + 1], ryz = L.rm[s1][lr + 1][tx];
            const C muxy = (C)4 / (((r0 + rx) + ry) + rxy);
            const C muxz = (C)4 / (((r0 + rx) + rz) + rxz);
            const C muyz = (C)4 / (((r0 + ry) + rz) + ryz);
            const C dl_ = s.dt * lam[r], dl2m = s.dt * (lam[r] + (C)2 * muk[r]);
            const C dmxy = s.dt * muxy, dmxz = s.dt * muxz, dmyz = s.dt * muyz;
            C nw[6];
            nw[0] = old[0][r] + fmad(dl2m, dX, dl_ * (dY + dZ));
            nw[1] = old[1][r] + fmad(dl2m, dY, dl_ * (dX + dZ));
            nw[2] = old[2][r] + fmad(dl2m, dZ, dl_ * (dX + dY));
            if constexpr (ND == 5) {
                nw[3] = fmad(dmxy, dk[4], old[3][r]);
            } else {
                nw[3] = fmad(dmxy, dk[4] + dk[ND - 1], old[3][r]);
            }
            nw[4] = fmad(dmxz, dxz_x + dxz_z, old[4][r]);
            nw[5] = fmad(dmyz, dyz_y + dyz_z, old[5][r]);
            C sn[3][3], t2[3][2];
            if (PML) {
                const C dxy_x = dk[4], dxy_y = dk[ND - 1];
                const bool ar = any[r];
                const bool mx = ar && px, my = ar && py, mz = ar && pz;
                const bool dm[3] = {mx, my, mz};
                C A[3] = {A0, A1[r], A2};
                C B[3] = {B0, B1[r], B2};
                if (s.vs) vs_damping(vsr[r], A, B);
                const C dnorm[3] = {dX, dY, dZ};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    C inc[3];
#pragma unroll
                    for (int a2 = 0; a2 < 3; ++a2) inc[a2] = (a2 == c ? dl2m : dl_) * dnorm[a2];
                    const C vp = pml_combine<C, 3>(old[c][r], inc, dm, sn3[r][c], A, B, sn[c]);
                    nw[c] = ar ? vp : nw[c];
                }
                {
                    const C inc[2] = {dmxy * dxy_x, dmxy * dxy_y};
                    const bool m2[2] = {mx, my};
                    const C A2_[2] = {A[0], A[1]}, B2_[2] = {B[0], B[1]};
                    const C vp = pml_combine<C, 2>(old[3][r], inc, m2, s2[r][0], A2_, B2_, t2[0]);
                    nw[3] = ar ? vp : nw[3];
                }
                {
                    const C inc[2] = {dmxz * dxz_x, dmxz * dxz_z};
                    const bool m2[2] = {mx, mz};
                    const C A2_[2] = {A[0], A[2]}, B2_[2] = {B[0], B[2]};
                    const C vp = pml_combine<C, 2>(old[4][r], inc, m2, s2[r][1], A2_, B2_, t2[1]);
                    nw[4] = ar ? vp : nw[4];
                }
                {
                    const C inc[2] = {dmyz * dyz_y, dmyz * dyz_z};
                    const bool m2[2] = {my, mz};
                    const C A2_[2] = {A[1], A[2]}, B2_[2] = {B[1], B[2]};
                    const C vp = pml_combine<C, 2>(old[5][r], inc, m2, s2[r][2], A2_, B2_, t2[2]);
                    nw[5] = ar ? vp : nw[5];
                }
            }
            if (t.rup[r]) {
                if (CP) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) store_comp<T, C>(s, SXX + c, k, t.ro[r], nw[c], ps[c]);
                } else if (sizeof(T) == 2 && s.round_mode) {
                    const uint32_t pt = t.ro[r] / (unsigned)sizeof(T) + (uint32_t)k * (uint32_t)sz;
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        bst_nt(ps[c], t.ro[r], CvtSR<T, C>::st(nw[c], s.fscale[SXX + c], sr_bits(pt, step, SXX + c)));
                } else {
#pragma unroll
                    for (int c = 0; c < 6; ++c) bst_nt(ps[c], t.ro[r], Cvt<T, C>::st(nw[c], s.fscale[SXX + c]));
                }
            }
            if (PML && any[r]) {
                const unsigned o = sb[r];
                if (DX && px) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.ss[XX_X + 3 * c], o, sn[c][0]);
                    bst(s.ss[XY_X], o, t2[0][0]);
                    bst(s.ss[XZ_X], o, t2[1][0]);
                }
                if (DY && py) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.ss[XX_X + 3 * c + 1], o, sn[c][1]);
                    bst(s.ss[XY_Y], o, t2[0][1]);
                    bst(s.ss[YZ_Y], o, t2[2][0]);
                }
                if (DZ && pz) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) bst(s.ss[XX_X + 3 * c + 2], o, sn[c][2]);
                    bst(s.ss[XZ_Z], o, t2[1][1]);
                    bst(s.ss[YZ_Z], o, t2[2][1]);
                }
            }
        }
        __syncthreads();

        // ---- ring, queues and delay line advance
        if (more) {
            C m2[NY];
            stage_mu(k + 2, m2);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                muk[r] = mun[r];
                mun[r] = m2[r];
            }
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                qvx[r][q] = qvx[r][q + 1];
                qvy[r][q] = qvy[r][q + 1];
            }
#pragma unroll
            for (int q = 0; q < 10; ++q) qvz[r][q] = qvz[r][q + 1];
            qvx[r][9] = G::use(nvx[r], t.rin[r] && z6, fi_vx);
            qvy[r][9] = G::use(nvy[r], t.rin[r] && z6, fi_vy);
            qvz[r][10] = G::use(nvz[r], t.rin[r] && z6, fi_vz);
        }
    }
}

}  // namespace sg
