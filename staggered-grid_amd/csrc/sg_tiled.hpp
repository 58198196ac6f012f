// sg_tiled.hpp -- the tiled velocity and stress sweeps (instantiated once per
// storage precision by sg_tiled_<prec>.hip).
//
// Mapping to gfx950:
//  * a block of 64 x TY/NY threads owns an xy box of up to 64 x TY points
//    (each thread computes NY consecutive y rows of one x column) and marches
//    through a chunk of z planes;
//  * the current z plane of every field that needs x or y derivatives is
//    staged in LDS with its radius-5 halo (x rows of 74, y columns of TY+10)
//    from x-contiguous, coalesced global reads;
//  * fields that need z derivatives sit in per-thread register queues of ten
//    planes, so each value is read from HBM once per sweep and z taps cost no
//    LDS traffic;
//  * the global loads of plane k+1 (staging values, queue heads) and the
//    read half of the read-modify-write of plane k are issued before plane k
//    is computed, so their latency overlaps the LDS arithmetic;
//  * loads are branch-free: out-of-grid addresses are clamped into the grid
//    and the value masked, so many loads stay in flight; addresses are a
//    per-plane scalar base plus a 32-bit per-thread offset;
//  * boxes never straddle a y or z PML plane, so y/z damping is uniform per
//    row and plane (per wave); x damping varies per lane only in the two
//    edge tiles of a row.  Each box carries the mask DM of the axes that can
//    damp inside it (1 = x, 2 = y, 4 = z): interior boxes (DM 0) run an
//    instantiation without any PML code, the three PML faces (DM 1, 2, 4)
//    instantiations that hold only that axis' splits, edges and corners the
//    general one (DM 7);
//  * work items are ordered on the host so each XCD (item position % 8)
//    walks a contiguous band of tiles, keeping halo re-reads in one L2.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "sg_internal.hpp"
#include "sg_precision.hpp"

#pragma clang fp contract(off)

namespace sg {

constexpr int TX = 64;       // x points per tile = threads in x
constexpr int XW = TX + 10;  // LDS row width with the x halo
constexpr int NTHR = 256;

template <class C, int TY>
struct VelLds {
    static constexpr int YH = TY + 10;
    C sxx[TY][XW];
    C sxz[TY][XW];
    C syy[YH][TX];
    C syz[YH][TX];
    C sxy[YH][XW];
    C b[TY + 1][TX + 1];
};

template <class C, int TY>
struct StrLds {
    static constexpr int YH = TY + 10;
    C vx[YH][XW];
    C vy[YH][XW];
    C vz[YH][XW];
    C rm[2][TY + 1][TX + 1];  // 1/mu of planes k and k+1 (ring)
};

// Double-buffered staging (SG_DB=1): plane k is staged into buffer k & 1 and
// the 1/mu ring has three slots, so a wave may stage plane k+1 while others
// still read plane k, and each plane needs one barrier instead of two.  Off
// by default: the same bits, 0.7-1.2 % slower at 512x512x384 and 500^3 on
// MI355X (twice the LDS per block; the waves do not wait at the barrier).
#ifndef SG_DB
#define SG_DB 0
#endif
template <class C, int TY>
struct VelLdsDB {
    VelLds<C, TY> b[2];
};
template <class C, int TY>
struct StrV {
    static constexpr int YH = TY + 10;
    C vx[YH][XW];
    C vy[YH][XW];
    C vz[YH][XW];
};
template <class C, int TY>
struct StrLdsDB {
    StrV<C, TY> v[2];
    C rm[3][TY + 1][TX + 1];
};
template <class C, int TY>
__device__ __forceinline__ VelLds<C, TY> &vplane(VelLds<C, TY> &L, int) { return L; }
template <class C, int TY>
__device__ __forceinline__ VelLds<C, TY> &vplane(VelLdsDB<C, TY> &L, int k) { return L.b[k & 1]; }
template <class C, int TY>
__device__ __forceinline__ StrLds<C, TY> &vplane(StrLds<C, TY> &L, int) { return L; }
template <class C, int TY>
__device__ __forceinline__ StrV<C, TY> &vplane(StrLdsDB<C, TY> &L, int k) { return L.v[k & 1]; }
template <class C, int TY>
__device__ __forceinline__ C (&rmslot(StrLds<C, TY> &L, int k))[TY + 1][TX + 1] { return L.rm[k & 1]; }
template <class C, int TY>
__device__ __forceinline__ C (&rmslot(StrLdsDB<C, TY> &L, int k))[TY + 1][TX + 1] { return L.rm[k % 3]; }
// the double-buffered staging of the interior kernels that load their
// staging rows straight into LDS (chosen at run time for grids without PML)
template <class C, int TY>
struct VelLdsDMA : VelLdsDB<C, TY> {};
template <class C, int TY>
struct StrLdsDMA : StrLdsDB<C, TY> {};
template <class LS>
struct IsDB : std::false_type {};
template <class C, int TY>
struct IsDB<VelLdsDB<C, TY>> : std::true_type {};
template <class C, int TY>
struct IsDB<StrLdsDB<C, TY>> : std::true_type {};
template <class C, int TY>
struct IsDB<VelLdsDMA<C, TY>> : std::true_type {};
template <class C, int TY>
struct IsDB<StrLdsDMA<C, TY>> : std::true_type {};
template <class LS>
struct IsDMA : std::false_type {};
template <class C, int TY>
struct IsDMA<VelLdsDMA<C, TY>> : std::true_type {};
template <class C, int TY>
struct IsDMA<StrLdsDMA<C, TY>> : std::true_type {};

// 5-pair staggered sums: forward at c uses f[c+n] - f[c+1-n], backward uses
// f[c+n-1] - f[c-n], n = 1..5, with stride st.
template <class C>
__device__ __forceinline__ C lfwd(const C *f, int c, int st, const C *k)
{
    C r = k[0] * (f[(c + 1) * st] - f[c * st]);
    r = fmad(k[1], f[(c + 2) * st] - f[(c - 1) * st], r);
    r = fmad(k[2], f[(c + 3) * st] - f[(c - 2) * st], r);
    r = fmad(k[3], f[(c + 4) * st] - f[(c - 3) * st], r);
    r = fmad(k[4], f[(c + 5) * st] - f[(c - 4) * st], r);
    return r;
}

template <class C>
__device__ __forceinline__ C lbwd(const C *f, int c, int st, const C *k)
{
    C r = k[0] * (f[c * st] - f[(c - 1) * st]);
    r = fmad(k[1], f[(c + 1) * st] - f[(c - 2) * st], r);
    r = fmad(k[2], f[(c + 2) * st] - f[(c - 3) * st], r);
    r = fmad(k[3], f[(c + 3) * st] - f[(c - 4) * st], r);
    r = fmad(k[4], f[(c + 4) * st] - f[(c - 5) * st], r);
    return r;
}

// Cross-lane x taps (SG_XLANE=1, FP32 arithmetic).  Lane l of a wave holds
// column x0+l of one row, and h the row's x halo: lanes 0-4 columns
// x0+64..x0+68, lanes 59-63 columns x0-5..x0-1.  The neighbours at x0+l+n
// come from n wave-wide shifts by one lane (DPP wave_shl:1); the lane that
// falls off the end (63) takes the next right-halo value, which a rotation
// of h by one lane (wave_rol:1) brings to lane 63 at each step.  Likewise
// to the left with wave_shr:1 / wave_ror:1 and lane 0.  The sums are the
// ones of lfwd/lbwd, term for term, so the results are the same bits.
#ifndef SG_XLANE
#define SG_XLANE 0
#endif
template <int CTRL>
__device__ __forceinline__ float dpp(float old, float src)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xf, 0xf, false));
}
constexpr int DPP_SHL1 = 0x130, DPP_ROL1 = 0x134, DPP_SHR1 = 0x138, DPP_ROR1 = 0x13C;

// p[n-1] = f[x+n] for n <= NP, m[n-1] = f[x-n] for n <= NM
template <int NP, int NM>
__device__ __forceinline__ void xnb(float c, float h, float (&p)[5], float (&m)[5])
{
    float t = h, v = c;
#pragma unroll
    for (int n = 0; n < NP; ++n) {
        t = dpp<DPP_ROL1>(t, t);
        v = dpp<DPP_SHL1>(t, v);
        p[n] = v;
    }
    float u = h, w = c;
#pragma unroll
    for (int n = 0; n < NM; ++n) {
        u = dpp<DPP_ROR1>(u, u);
        w = dpp<DPP_SHR1>(u, w);
        m[n] = w;
    }
}

// lfwd / lbwd along x with stride 1, from a lane's own value and the halo
__device__ __forceinline__ float xfwd(float c, float h, const float *k)
{
    float p[5], m[5];
    xnb<5, 4>(c, h, p, m);
    float r = k[0] * (p[0] - c);
    r = fmad(k[1], p[1] - m[0], r);
    r = fmad(k[2], p[2] - m[1], r);
    r = fmad(k[3], p[3] - m[2], r);
    r = fmad(k[4], p[4] - m[3], r);
    return r;
}
__device__ __forceinline__ float xbwd(float c, float h, const float *k)
{
    float p[5], m[5];
    xnb<4, 5>(c, h, p, m);
    float r = k[0] * (c - m[0]);
    r = fmad(k[1], p[0] - m[1], r);
    r = fmad(k[2], p[1] - m[2], r);
    r = fmad(k[3], p[2] - m[3], r);
    r = fmad(k[4], p[3] - m[4], r);
    return r;
}
// the double overloads are never called (FP64 keeps the LDS taps)
__device__ __forceinline__ double xfwd(double, double, const double *) { return 0.0; }
__device__ __forceinline__ double xbwd(double, double, const double *) { return 0.0; }

// Global-address-space access.  The shot descriptor lives in device memory,
// so the compiler cannot infer that its pointers are global and would emit
// FLAT instructions, which also count against the LDS wait counter and make
// every LDS read wait for outstanding HBM loads.
template <class T>
__device__ __forceinline__ T gld(const T *p, unsigned off)
{
    return ((const __attribute__((address_space(1))) T *)p)[off];
}

template <class T>
__device__ __forceinline__ void gst(T *p, unsigned off, T v)
{
    ((__attribute__((address_space(1))) T *)p)[off] = v;
}

template <>
__device__ __forceinline__ bf16 gld<bf16>(const bf16 *p, unsigned off)
{
    return bf16{((const __attribute__((address_space(1))) uint16_t *)p)[off]};
}

// Byte-offset access: a uniform (scalar) base plus a 32-bit per-lane byte
// offset maps onto the global_load/store "saddr + voffset" form, with no
// 64-bit per-lane address arithmetic and no address registers held per load.
template <class T>
__device__ __forceinline__ T bld(const T *p, unsigned boff)
{
    typedef __attribute__((address_space(1))) const char gcc;
    typedef __attribute__((address_space(1))) const T gT;
    return *(gT *)((gcc *)p + boff);
}

template <class T>
__device__ __forceinline__ void bst(T *p, unsigned boff, T v)
{
    typedef __attribute__((address_space(1))) char gc;
    typedef __attribute__((address_space(1))) T gT;
    *(gT *)((gc *)p + boff) = v;
}

// Streaming (nontemporal) variants for the read-modify-write operands of a
// sweep, which no other box reads: they bypass L2 retention so the planes
// shared through halos stay cached.  Enabled with SG_NT=1.
#ifndef SG_NT
#define SG_NT 0
#endif
template <class T>
__device__ __forceinline__ T bld_nt(const T *p, unsigned boff)
{
#if SG_NT
    typedef __attribute__((address_space(1))) const char gcc;
    typedef __attribute__((address_space(1))) const T gT;
    return __builtin_nontemporal_load((gT *)((gcc *)p + boff));
#else
    return bld(p, boff);
#endif
}

template <class T>
__device__ __forceinline__ void bst_nt(T *p, unsigned boff, T v)
{
#if SG_NT
    typedef __attribute__((address_space(1))) char gc;
    typedef __attribute__((address_space(1))) T gT;
    __builtin_nontemporal_store(v, (gT *)((gc *)p + boff));
#else
    bst(p, boff, v);
#endif
}

template <>
__device__ __forceinline__ bf16 bld<bf16>(const bf16 *p, unsigned boff)
{
    typedef __attribute__((address_space(1))) const char gcc;
    return bf16{*(__attribute__((address_space(1))) const uint16_t *)((gcc *)p + boff)};
}

template <>
__device__ __forceinline__ void bst<bf16>(bf16 *p, unsigned boff, bf16 v)
{
    typedef __attribute__((address_space(1))) char gc;
    *(__attribute__((address_space(1))) uint16_t *)((gc *)p + boff) = v.bits;
}

template <>
__device__ __forceinline__ bf16 bld_nt<bf16>(const bf16 *p, unsigned boff)
{
    return bld<bf16>(p, boff);
}

template <>
__device__ __forceinline__ void bst_nt<bf16>(bf16 *p, unsigned boff, bf16 v)
{
    bst<bf16>(p, boff, v);
}

template <>
__device__ __forceinline__ void gst<bf16>(bf16 *p, unsigned off, bf16 v)
{
    ((__attribute__((address_space(1))) uint16_t *)p)[off] = v.bits;
}

// Old PML splits of single-axis face boxes loaded one plane ahead: bit 0
// velocity sweep, bit 1 stress sweep.  Measured at 512x512x384 (300 steps):
// velocity sweep FP32 1.529 -> 1.505 ms, FP16 1.216 -> 1.164; stress sweep
// FP32 1.894 -> 1.910, FP16 1.546 -> 1.556 (five split operands per point
// held across the plane), so the velocity sweep only.
#ifndef SG_SPLIT_PREFETCH
#define SG_SPLIT_PREFETCH 1
#endif

// FP32 staging rows loaded straight into LDS (global_load_lds_dword) instead
// of through registers: the y-halo rows of both sweeps and the velocity
// sweep's centre rows of sxx, sxy, syy.  A bit per box list (1 interior, 2 x
// faces, 4 y faces, 8 z faces, 16 edges and corners); those lists use the
// double-buffered staging, so the rows of plane k+1 land in the other buffer
// while plane k is read.  Values outside the array are not zeroed, as the
// register path zeroes them; no updated point reads them (its taps stay
// inside [0, n)).  Measured on MI355X with the interior lists only
// (SG_LDSDMA=1): their kernels drop from 137-144 to 113-122 VGPRs (4 waves
// per SIMD instead of 3), 500^3 without PML 30.28 -> 31.31 Gpt/s, but
// 512x512x384 P20 and 256^3 P16 unchanged (29.45 / 29.42, 24.26 / 24.27)
// while the interior lists fetch 8-9 % more (more blocks share each L2);
// with the PML lists as well it is slower (512x512x384: 28.96 with the x
// faces, 28.82 with every list).  Off by default: the solver runs the
// interior kernels of this form (k_velocity_wl_lds, k_stress_wl_lds) for
// grids without PML, where the interior list runs alone.
#ifndef SG_LDSDMA
#define SG_LDSDMA 0
#endif
constexpr int list_bit(int dm) { return dm == 0 ? 0 : (dm == 1 ? 1 : (dm == 2 ? 2 : (dm == 4 ? 3 : 4))); }
template <class T, int DM>
struct DmaOn {
    static constexpr bool v = std::is_same<T, float>::value && ((SG_LDSDMA >> list_bit(DM)) & 1) != 0;
};

// lane l of the wave loads the 32-bit value at p + boff into dst[l] (dst is
// the same for every lane of the wave)
__device__ __forceinline__ void lds_dma(const float *p, unsigned boff, float *dst)
{
    typedef __attribute__((address_space(1))) const char gcc;
    __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) const void *)((gcc *)p + boff),
                                     (__attribute__((address_space(3))) void *)dst, 4, 0, 0);
}

// wait until every vector-memory load of the wave has completed (the LDS rows
// above), before the barrier that publishes them
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// masked load at a byte offset that is always inside the array
template <class T, class C>
__device__ __forceinline__ C mld(const T *__restrict__ p, unsigned boff, bool ok, float inv)
{
    const C v = Cvt<T, C>::ld(bld(p, boff), inv);
    return ok ? v : (C)0;
}

// The same in two halves: a load of the stored value, and its conversion and
// mask where the value is first used.  Any arithmetic on a loaded register
// makes the compiler wait for that load right there; with the conversion and
// the mask at the use (the next plane's staging, the update, the queue push)
// the loads issued for plane k+1 stay in flight while plane k is computed.
template <class T, class C>
__device__ __forceinline__ C cvm(T raw, bool ok, float inv)
{
    const C v = Cvt<T, C>::ld(raw, inv);
    return ok ? v : (C)0;
}

// Staging registers of the sweeps: the stored value, converted and masked at
// its use (FP32, FP16, BF16), or for FP64, the checking build, the converted
// and masked value itself, as loaded (its loads are waited for where they are
// issued).
template <class T, class C>
struct Stg {
    static constexpr bool late = sizeof(T) != 8;
    using S = typename std::conditional<late, T, C>::type;
    static __device__ __forceinline__ S ld(const T *__restrict__ p, unsigned boff, bool ok, float inv)
    {
        if constexpr (late) return bld(p, boff);
        else return mld<T, C>(p, boff, ok, inv);
    }
    static __device__ __forceinline__ C use(S v, bool ok, float inv)
    {
        if constexpr (late) return cvm<T, C>(v, ok, inv);
        else return v;
    }
};

// Per-cell damping (the CUDA reference's form): A holds dt*g/2 per axis on
// entry; with the point's local S velocity v, hd = A*v and
// A <- (1-hd)/(1+hd), B <- 1/(1+hd).
template <class C>
__device__ __forceinline__ void vs_damping(C v, C (&A)[3], C (&B)[3])
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const C hd = A[a] * v;
        B[a] = (C)1 / ((C)1 + hd);
        A[a] = ((C)1 - hd) * B[a];
    }
}

// Split-field PML update of one component from the old splits `so` of its
// contributing axes (loaded by the caller): damped axes f_a <- A_a f_a +
// B_a inc_a (3DCPUwithPML.../ElasticWaveEquation3D.cpp:153-158), new splits
// returned in `sn`.  Result: the sum of the new splits when every axis damps
// (the reference rebuilds the component from its splits, .cpp:159), else the
// undamped remainder (fold - old damped splits) plus undamped increments
// plus new splits.
template <class C, int NA>
__device__ __forceinline__ C pml_combine(C fold, const C (&inc)[NA], const bool (&dm)[NA], const C (&so)[NA],
                                         const C (&A)[NA], const C (&B)[NA], C (&sn)[NA])
{
    bool all = true;
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        all = all && dm[a];
        sn[a] = dm[a] ? fmad(A[a], so[a], B[a] * inc[a]) : (C)0;
    }
    if (all) {
        C r = sn[0];
#pragma unroll
        for (int a = 1; a < NA; ++a) r += sn[a];
        return r;
    }
    C r = fold;
#pragma unroll
    for (int a = 0; a < NA; ++a) r = dm[a] ? r - so[a] : r + inc[a];
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (dm[a]) r += sn[a];
    return r;
}

template <class T>
__device__ __forceinline__ T *plane(T *base, int k, long long sz)
{
    return base + (long long)k * sz;
}

// Per-thread geometry of a box [i0,i1) x [j0,j1): the block is 64 x BY
// threads (BY = TY/NY), thread (tx, ty) owns column i0+tx and rows
// j0 + ty*NY + r.  Offsets inside a plane are 32-bit BYTE offsets (a plane
// holds less than 4 GiB); addresses outside the array are clamped and masked.
template <int TY, int NY>
struct TileGeo {
    static constexpr int BY = TY / NY;
    static constexpr int NH = (10 + BY - 1) / BY;  // y-halo rows per thread
    int i, ic, jr[NY];
    unsigned ro[NY];    // clamped offset of (i, jr) in a plane
    bool rin[NY];       // (i, jr) inside the array
    bool rup[NY];       // (i, jr) updated by this box
    unsigned hro[NY];   // x-halo offsets (threads tx < 10; with XL lanes
    bool hok[NY];       // 0-4 right of the tile and 59-63 left of it)
    int hc;             // LDS column of this thread's x-halo value
    unsigned yro[NH];   // y-halo offsets: rows h = ty + m*BY (h < 10)
    bool yok[NH];
    int ylr[NH];        // their LDS rows
    unsigned bxo[NY];   // offsets of column i0+64 (thread tx == 10)
    bool bxok;
    unsigned byo, bxyo; // row j0+TY (ty == 0) and the corner (i0+64, j0+TY)
    bool byok, bxyok;
};

template <int TY, int NY, int ES, bool XL = false>
__device__ __forceinline__ void tile_geo(const Geom &g, int i0, int i1, int j0, int j1, TileGeo<TY, NY> &t)
{
    constexpr int BY = TY / NY;  // ES: bytes per stored element
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int sy = (int)g.sy;
    t.i = i0 + tx;
    const bool xin = t.i < g.nx, xup = t.i >= HALO && t.i < i1;
    t.ic = xin ? t.i : g.nx - 1;
    t.hc = tx < 5 ? tx : TX + tx;
    // negative left of the first tile
    const int hi = XL ? (tx < 5 ? i0 + TX + tx : i0 + tx - TX) : i0 - HALO + t.hc;
    const bool hx = (XL ? (tx < 5 || tx >= TX - 5) : tx < 10) && hi >= 0 && hi < g.nx;
    const int hic = hi < 0 ? 0 : (hi < g.nx ? hi : g.nx - 1);
    const int bxi = i0 + TX;
    const int bxc = bxi < g.nx ? bxi : g.nx - 1;
    t.bxok = tx == 10 && bxi < g.nx;
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        const int j = j0 + ty * NY + r;
        t.jr[r] = j;
        const bool yin = j < g.ny;
        const int jc = yin ? j : g.ny - 1;
        t.ro[r] = (unsigned)(t.ic + jc * sy) * ES;
        t.rin[r] = xin && yin;
        t.rup[r] = xup && j < j1;
        t.hok[r] = hx && yin;
        // a lane whose value is masked reads its own centre instead: a cache
        // line the wave reads anyway, so a halo load touches no other lines
        // than the halo's own
        t.hro[r] = t.hok[r] ? (unsigned)(hic + jc * sy) * ES : t.ro[r];
        t.bxo[r] = (t.bxok && t.rin[r]) ? (unsigned)(bxc + jc * sy) * ES : t.ro[r];
    }
#pragma unroll
    for (int m = 0; m < TileGeo<TY, NY>::NH; ++m) {
        const int h = ty + BY * m;
        const int lr = h < 5 ? h : TY + h;
        const int jh = j0 - HALO + lr;
        t.ylr[m] = lr;
        t.yok[m] = h < 10 && xin && jh < g.ny;
        t.yro[m] = t.yok[m] ? (unsigned)(t.ic + jh * sy) * ES : t.ro[0];
    }
    const int byj = j0 + TY;
    const int byc = byj < g.ny ? byj : g.ny - 1;
    t.byok = ty == 0 && xin && byj < g.ny;
    t.bxyok = ty == 0 && tx == 10 && bxi < g.nx && byj < g.ny;
    t.byo = t.byok ? (unsigned)(t.ic + byc * sy) * ES : t.ro[0];
    t.bxyo = t.bxyok ? (unsigned)(bxc + byc * sy) * ES : t.ro[0];
}

// Store of an updated value with compensation when field f has a
// compensation array: the 16-bit value and the residual of its rounding
// (which the next update adds back); plain rounding otherwise.
template <class T, class C>
__device__ __forceinline__ void store_comp(const Shot<T, C> &s, int f, int k, unsigned boff, C v, T *pf)
{
    const T q = Cvt<T, C>::st(v, s.fscale[f]);
    bst_nt(pf, boff, q);
    if (s.comp[f]) {
        const C res = v - Cvt<T, C>::ld(q, s.finv[f]);
        bst_nt(plane(s.comp[f], k, s.g.sz), boff, Cvt<T, C>::st(res, s.cscale[f]));
    }
}

// ---------------------------------------------------------------------------
// velocity sweep
// ---------------------------------------------------------------------------
template <class T, class C, int TY, int NY, int DM, bool CP, class LS>
__device__ __forceinline__ void velocity_tile(const Shot<T, C> &s, int i0, int i1, int j0, int j1, int k0, int k1,
                                              LS &Ls)
{
    constexpr int BY = TY / NY, NH = TileGeo<TY, NY>::NH;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long long sz = g.sz;
    // x taps across lanes (FP32 arithmetic only; 1 = interior boxes, 2 = all)
    constexpr bool XL = (SG_XLANE == 2 || (SG_XLANE == 1 && DM == 0)) && sizeof(C) == 4;
    TileGeo<TY, NY> t;
    tile_geo<TY, NY, (int)sizeof(T), XL>(g, i0, i1, j0, j1, t);

    const T *Fxx = s.f[SXX], *Fyy = s.f[SYY], *Fzz = s.f[SZZ];
    const T *Fxy = s.f[SXY], *Fxz = s.f[SXZ], *Fyz = s.f[SYZ];
    const T *Fb = s.b;
    const float fi_xx = s.finv[SXX], fi_yy = s.finv[SYY], fi_zz = s.finv[SZZ];
    const float fi_xy = s.finv[SXY], fi_xz = s.finv[SXZ], fi_yz = s.finv[SYZ];
    const float ib = 1.0f / s.bscale;

    // register z queues: qxz/qyz[m] = plane k-5+m, qzz[m] = plane k-4+m, qb[m] = plane k+m
    C qxz[NY][10], qyz[NY][10], qzz[NY][10], qb[NY][2];
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const T *pxz = plane(Fxz, k0 - 5 + q, sz), *pyz = plane(Fyz, k0 - 5 + q, sz);
        const T *pzz = plane(Fzz, k0 - 4 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            qxz[r][q] = mld<T, C>(pxz, t.ro[r], t.rin[r], fi_xz);
            qyz[r][q] = mld<T, C>(pyz, t.ro[r], t.rin[r], fi_yz);
            qzz[r][q] = mld<T, C>(pzz, t.ro[r], t.rin[r], fi_zz);
        }
    }
#pragma unroll
    for (int r = 0; r < NY; ++r) {
        qb[r][0] = mld<T, C>(plane(Fb, k0, sz), t.ro[r], t.rin[r], ib);
        qb[r][1] = mld<T, C>(plane(Fb, k0 + 1, sz), t.ro[r], t.rin[r], ib);
    }

    // staging registers of one plane: stored values, converted and masked
    // when they are staged
    using G = Stg<T, C>;
    using S = typename G::S;
    S c_xx[NY], c_xy[NY], c_yy[NY], h_xx[NY], h_xy[NY], h_xz[NY], b_x[NY];
    S y_yy[NH], y_yz[NH], y_xy[NH], b_y;
    // centre rows of sxx, sxy, syy and the y-halo rows straight into LDS
    constexpr bool YD = (DmaOn<T, DM>::v || IsDMA<LS>::value) && std::is_same<T, float>::value && IsDB<LS>::value && !XL;
    auto fetch = [&](int k) {
        const T *pxx = plane(Fxx, k, sz), *pxy = plane(Fxy, k, sz), *pyy = plane(Fyy, k, sz);
        const T *pxz = plane(Fxz, k, sz), *pyz = plane(Fyz, k, sz), *pb = plane(Fb, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            if constexpr (!YD) {
                c_xx[r] = G::ld(pxx, t.ro[r], t.rin[r], fi_xx);
                c_xy[r] = G::ld(pxy, t.ro[r], t.rin[r], fi_xy);
                c_yy[r] = G::ld(pyy, t.ro[r], t.rin[r], fi_yy);
            }
            h_xx[r] = G::ld(pxx, t.hro[r], t.hok[r], fi_xx);
            h_xy[r] = G::ld(pxy, t.hro[r], t.hok[r], fi_xy);
            h_xz[r] = G::ld(pxz, t.hro[r], t.hok[r], fi_xz);
            b_x[r] = G::ld(pb, t.bxo[r], t.bxok && t.rin[r], ib);
        }
        if constexpr (YD) {
            b_y = G::ld(pb, t.byo, t.byok, ib);
            auto &Ln = vplane(Ls, k);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const int lr = ty * NY + r;
                lds_dma((const float *)pxx, t.ro[r], &Ln.sxx[lr][5]);
                lds_dma((const float *)pxy, t.ro[r], &Ln.sxy[lr + 5][5]);
                lds_dma((const float *)pyy, t.ro[r], &Ln.syy[lr + 5][0]);
            }
            // a wave with no row in slot m (ty + BY*m >= 10) loads its row of
            // slot m-1 again, so every wave issues the same loads; with more
            // than ten waves per block (BY > 10) waves 10.. have no row at all
#pragma unroll
            for (int m = 0; m < NH; ++m) {
                const bool own = ty + BY * m < 10;
                if (BY > 10 && !own) continue;
                const int lr = own ? t.ylr[m] : t.ylr[m > 0 ? m - 1 : 0];
                const unsigned o = own ? t.yro[m] : t.yro[m > 0 ? m - 1 : 0];
                lds_dma((const float *)pyy, o, &Ln.syy[lr][0]);
                lds_dma((const float *)pyz, o, &Ln.syz[lr][0]);
                lds_dma((const float *)pxy, o, &Ln.sxy[lr][5]);
            }
        } else {
#pragma unroll
            for (int m = 0; m < NH; ++m) {
                y_yy[m] = G::ld(pyy, t.yro[m], t.yok[m], fi_yy);
                y_yz[m] = G::ld(pyz, t.yro[m], t.yok[m], fi_yz);
                y_xy[m] = G::ld(pxy, t.yro[m], t.yok[m], fi_xy);
            }
            b_y = G::ld(pb, t.byo, t.byok, ib);
        }
    };
    fetch(k0);

    // old velocities of plane kk (stored values; compensated kernels also the
    // residuals), converted by rmw_value where the update uses them
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
    // compensated update: the old value is stored value + residual
    auto rmw_value = [&](const T (&raw)[3][NY], const T (&res)[3][NY], int c, int r) -> C {
        C v = Cvt<T, C>::ld(raw[c][r], s.finv[VX + c]);
        if (CP && s.comp[VX + c]) v += Cvt<T, C>::ld(res[c][r], s.cinv[VX + c]);
        return v;
    };
    // the RMW operands are prefetched one plane ahead in interior boxes, and in
    // every box with 16-bit storage (measured: +4 % FP16, -4 % FP32 in PML boxes)
    constexpr bool RP = DM == 0 || sizeof(T) == 2;
    T vpre[3][NY], vpres[3][NY];
    if (RP) load_rmw(k0, vpre, vpres);

    // damping: x per lane (edge tiles), y per box (y cuts at the PML planes),
    // z per plane, each only if the box's mask DM allows that axis; every
    // split array is addressed from its own base with the point's slot offset
    constexpr bool PML = DM != 0, DX = (DM & 1) != 0, DY = (DM & 2) != 0, DZ = (DM & 4) != 0;
    const bool px = DX && in_pml(t.i, g.nx, g.pml);
    const bool py = DY && in_pml(j0, g.ny, g.pml);
    C A0 = (C)1, B0 = (C)1, A1[NY], B1[NY];
    // step number keying stochastic rounding (16-bit storage only)
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
    // boxes that damp along one axis load the old splits of the next plane
    // with the other operands of the next plane (every lane loads, at slot 0
    // where its point does not damp; the values are used only where it does)
    constexpr bool SPF = (SG_SPLIT_PREFETCH & 1) && (DM == 1 || DM == 2 || DM == 4);
    constexpr int AX = DM == 1 ? 0 : (DM == 2 ? 1 : 2);
    C spo[NY][3], spv[NY];
    auto load_splits = [&](int kk) {
        const bool pzk = DZ && in_pml(kk, g.nz, g.pml);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const bool an = (px || py || pzk) && t.rup[r];
            const unsigned o = an ? (unsigned)(split_slot(g, t.i, t.jr[r], kk) * (long long)sizeof(C)) : 0u;
#pragma unroll
            for (int c = 0; c < 3; ++c) spo[r][c] = bld(s.sv[3 * c + AX], o);
            spv[r] = s.vs ? bld(s.vs, o) : (C)0;
        }
    };
    if (SPF) load_splits(k0);

    for (int k = k0; k < k1; ++k) {
        auto &L = vplane(Ls, k);
        // staged values of plane k (loaded during plane k-1)
        C cxx[NY], cxy[NY], hxx[NY], hxy[NY], hxz[NY];
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            if constexpr (!YD) {
                cxx[r] = G::use(c_xx[r], t.rin[r], fi_xx);
                cxy[r] = G::use(c_xy[r], t.rin[r], fi_xy);
            }
            hxx[r] = G::use(h_xx[r], t.hok[r], fi_xx);
            hxy[r] = G::use(h_xy[r], t.hok[r], fi_xy);
            hxz[r] = G::use(h_xz[r], t.hok[r], fi_xz);
        }
        // x derivatives across lanes, before the staging registers are reused
        C dxl[3][NY];
        if (XL) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                dxl[0][r] = xfwd(cxx[r], hxx[r], s.cx);
                dxl[1][r] = xbwd(cxy[r], hxy[r], s.cx);
                dxl[2][r] = xbwd(qxz[r][5], hxz[r], s.cx);
            }
        }
        // ---- stage plane k
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            if constexpr (!YD) {
                if (!XL) L.sxx[lr][tx + 5] = cxx[r];
                L.sxy[lr + 5][tx + 5] = cxy[r];
                L.syy[lr + 5][tx] = G::use(c_yy[r], t.rin[r], fi_yy);
            }
            if (!XL) L.sxz[lr][tx + 5] = qxz[r][5];
            L.syz[lr + 5][tx] = qyz[r][5];
            L.b[lr][tx] = qb[r][0];
        }
        if (!XL && tx < 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const int lr = ty * NY + r;
                L.sxx[lr][t.hc] = hxx[r];
                L.sxy[lr + 5][t.hc] = hxy[r];
                L.sxz[lr][t.hc] = hxz[r];
            }
        }
        if (tx == 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) L.b[ty * NY + r][TX] = G::use(b_x[r], t.bxok && t.rin[r], ib);
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            if (!YD && ty + BY * m < 10) {
                L.syy[t.ylr[m]][tx] = G::use(y_yy[m], t.yok[m], fi_yy);
                L.syz[t.ylr[m]][tx] = G::use(y_yz[m], t.yok[m], fi_yz);
                L.sxy[t.ylr[m]][tx + 5] = G::use(y_xy[m], t.yok[m], fi_xy);
            }
        }
        if (ty == 0) L.b[TY][tx] = G::use(b_y, t.byok, ib);
        if constexpr (YD) wait_vm_all();
        __syncthreads();

        // ---- read half of the read-modify-write of plane k: loaded one plane
        // ahead (RP); else issued first, so waiting for it does not wait for
        // the loads below
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
                if (SPF) {
                    vsr[r] = spv[r];
#pragma unroll
                    for (int c = 0; c < 3; ++c) so[r][c][AX] = spo[r][c];
                    continue;
                }
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

        // ---- loads for later, in the order their values are needed: the RMW
        // operands of plane k+1 (their registers are copied when the loop
        // wraps) and the queue heads (pushed after this plane) before the
        // staging values of plane k+1 (first used after the next barrier).
        // The vector-memory counter retires in issue order, so the waits
        // for the first two groups do not include the third.  The last
        // plane of the box reloads itself (just read, cached): with no branch
        // around these loads the compiler's waits count them in every
        // iteration.
        const int kn = k + 1 < k1 ? k + 1 : k;
        if (RP) load_rmw(kn, vpre, vpres);
        if (SPF) load_splits(kn);
        S nxz[NY], nyz[NY], nzz[NY], nb[NY];
        const bool z5 = k + 5 < g.nz, z6 = k + 6 < g.nz, z2 = k + 2 < g.nz;
        {
            const T *pxz = plane(Fxz, z5 ? k + 5 : k, sz), *pyz = plane(Fyz, z5 ? k + 5 : k, sz);
            const T *pzz = plane(Fzz, z6 ? k + 6 : k, sz), *pb = plane(Fb, z2 ? k + 2 : k, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                nxz[r] = G::ld(pxz, t.ro[r], t.rin[r] && z5, fi_xz);
                nyz[r] = G::ld(pyz, t.ro[r], t.rin[r] && z5, fi_yz);
                nzz[r] = G::ld(pzz, t.ro[r], t.rin[r] && z6, fi_zz);
                nb[r] = G::ld(pb, t.ro[r], t.rin[r] && z2, ib);
            }
        }
        fetch(kn);

        // ---- compute plane k
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            C d[3][3];
            d[0][0] = XL ? dxl[0][r] : lfwd<C>(&L.sxx[lr][0], tx + 5, 1, s.cx);
            d[0][1] = lbwd<C>(&L.sxy[0][tx + 5], lr + 5, XW, s.cy);
            d[0][2] = lbwd<C>(qxz[r], 5, 1, s.cz);
            d[1][0] = XL ? dxl[1][r] : lbwd<C>(&L.sxy[lr + 5][0], tx + 5, 1, s.cx);
            d[1][1] = lfwd<C>(&L.syy[0][tx], lr + 5, TX, s.cy);
            d[1][2] = lbwd<C>(qyz[r], 5, 1, s.cz);
            d[2][0] = XL ? dxl[2][r] : lbwd<C>(&L.sxz[lr][0], tx + 5, 1, s.cx);
            d[2][1] = lbwd<C>(&L.syz[0][tx], lr + 5, TX, s.cy);
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
                // split update evaluated on every lane and selected where a
                // point damps: the arithmetic is one basic block, and the
                // split stores come after the field stores
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
        if constexpr (!IsDB<LS>::value) __syncthreads();

        // ---- advance the z queues
#pragma unroll
        for (int r = 0; r < NY; ++r) {
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                qxz[r][q] = qxz[r][q + 1];
                qyz[r][q] = qyz[r][q + 1];
                qzz[r][q] = qzz[r][q + 1];
            }
            qxz[r][9] = G::use(nxz[r], t.rin[r] && z5, fi_xz);
            qyz[r][9] = G::use(nyz[r], t.rin[r] && z5, fi_yz);
            qzz[r][9] = G::use(nzz[r], t.rin[r] && z6, fi_zz);
            qb[r][0] = qb[r][1];
            qb[r][1] = G::use(nb[r], t.rin[r] && z2, ib);
        }
    }
}

// ---------------------------------------------------------------------------
// stress sweep
// ---------------------------------------------------------------------------
template <class T, class C, int TY, int NY, int DM, bool CP, class LS>
__device__ __forceinline__ void stress_tile(const Shot<T, C> &s, int i0, int i1, int j0, int j1, int k0, int k1,
                                              LS &Ls)
{
    constexpr int BY = TY / NY, NH = TileGeo<TY, NY>::NH;
    const Geom &g = s.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const long long sz = g.sz;
    // x taps across lanes (FP32 arithmetic only; 1 = interior boxes, 2 = all)
    constexpr bool XL = (SG_XLANE == 2 || (SG_XLANE == 1 && DM == 0)) && sizeof(C) == 4;
    TileGeo<TY, NY> t;
    tile_geo<TY, NY, (int)sizeof(T), XL>(g, i0, i1, j0, j1, t);

    const T *Fvx = s.f[VX], *Fvy = s.f[VY], *Fvz = s.f[VZ];
    const T *Fmu = s.mu, *Flam = s.lam;
    const float fi_vx = s.finv[VX], fi_vy = s.finv[VY], fi_vz = s.finv[VZ];
    const float il = 1.0f / s.lscale, im = 1.0f / s.mscale;

    // queues: qvx/qvy[m] = plane k-4+m, qvz[m] = plane k-5+m
    C qvx[NY][10], qvy[NY][10], qvz[NY][10];
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const T *px_ = plane(Fvx, k0 - 4 + q, sz), *py_ = plane(Fvy, k0 - 4 + q, sz);
        const T *pz_ = plane(Fvz, k0 - 5 + q, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            qvx[r][q] = mld<T, C>(px_, t.ro[r], t.rin[r], fi_vx);
            qvy[r][q] = mld<T, C>(py_, t.ro[r], t.rin[r], fi_vy);
            qvz[r][q] = mld<T, C>(pz_, t.ro[r], t.rin[r], fi_vz);
        }
    }

    // mu staging registers (center rows, x halo, y halo, corner) of one plane
    // (stored values, converted when staged)
    using G = Stg<T, C>;
    using S = typename G::S;
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
    // reciprocals into the ring slot of plane k (0 outside the array: never
    // read by an updated point); mc: mu of the centre rows
    auto stage_mu = [&](int k, C (&mc)[NY]) {
        auto &R = rmslot(Ls, k);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            mc[r] = G::use(m_c[r], t.rin[r], im);
            R[ty * NY + r][tx] = t.rin[r] ? (C)1 / mc[r] : (C)0;
        }
        if (tx == 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const bool ok = t.bxok && t.rin[r];
                R[ty * NY + r][TX] = ok ? (C)1 / G::use(m_x[r], ok, im) : (C)0;
            }
        }
        if (ty == 0) {
            R[TY][tx] = t.byok ? (C)1 / G::use(m_y, t.byok, im) : (C)0;
            if (tx == 10) R[TY][TX] = t.bxyok ? (C)1 / G::use(m_xy, t.bxyok, im) : (C)0;
        }
    };
    C muk[NY], mun[NY];  // mu of planes k and k+1 (for lambda + 2 mu)
    fetch_mu(k0);
    stage_mu(k0, muk);
    fetch_mu(k0 + 1);
    stage_mu(k0 + 1, mun);

    // staging registers: x and y halos of the three velocities (stored
    // values, converted and masked when staged)
    S h_vx[NY], h_vy[NY], h_vz[NY], y_vx[NH], y_vy[NH], y_vz[NH];
    // y-halo rows straight into LDS
    constexpr bool YD = (DmaOn<T, DM>::v || IsDMA<LS>::value) && std::is_same<T, float>::value && IsDB<LS>::value;
    auto fetch = [&](int k) {
        const T *pvx = plane(Fvx, k, sz), *pvy = plane(Fvy, k, sz), *pvz = plane(Fvz, k, sz);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            h_vx[r] = G::ld(pvx, t.hro[r], t.hok[r], fi_vx);
            h_vy[r] = G::ld(pvy, t.hro[r], t.hok[r], fi_vy);
            h_vz[r] = G::ld(pvz, t.hro[r], t.hok[r], fi_vz);
        }
        if constexpr (YD) {
            auto &Ln = vplane(Ls, k);
            // a wave with no row in slot m (ty + BY*m >= 10) loads its row of
            // slot m-1 again, so every wave issues the same loads; with more
            // than ten waves per block (BY > 10) waves 10.. have no row at all
#pragma unroll
            for (int m = 0; m < NH; ++m) {
                const bool own = ty + BY * m < 10;
                if (BY > 10 && !own) continue;
                const int lr = own ? t.ylr[m] : t.ylr[m > 0 ? m - 1 : 0];
                const unsigned o = own ? t.yro[m] : t.yro[m > 0 ? m - 1 : 0];
                lds_dma((const float *)pvx, o, &Ln.vx[lr][5]);
                lds_dma((const float *)pvy, o, &Ln.vy[lr][5]);
                lds_dma((const float *)pvz, o, &Ln.vz[lr][5]);
            }
        } else {
#pragma unroll
            for (int m = 0; m < NH; ++m) {
                y_vx[m] = G::ld(pvx, t.yro[m], t.yok[m], fi_vx);
                y_vy[m] = G::ld(pvy, t.yro[m], t.yok[m], fi_vy);
                y_vz[m] = G::ld(pvz, t.yro[m], t.yok[m], fi_vz);
            }
        }
    };
    fetch(k0);

    // old stresses and lambda of plane kk (stored values; compensated kernels
    // also the residuals), converted by rmw_value where the update uses them
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
    // compensated update: the old value is stored value + residual
    auto rmw_value = [&](const T (&raw)[6][NY], const T (&res)[6][NY], int c, int r) -> C {
        C v = Cvt<T, C>::ld(raw[c][r], s.finv[SXX + c]);
        if (CP && s.comp[SXX + c]) v += Cvt<T, C>::ld(res[c][r], s.cinv[SXX + c]);
        return v;
    };
    // the RMW operands are prefetched one plane ahead in interior boxes, and in
    // every box with 16-bit storage (measured: +4 % FP16, -4 % FP32 in PML boxes)
    constexpr bool RP = DM == 0 || sizeof(T) == 2;
    T spre[6][NY], spres[6][NY], lpre[NY];
    if (RP) load_rmw(k0, spre, spres, lpre);

    // damping as in the velocity sweep; splits per component: XX, YY, ZZ:
    // x, y, z; XY: x, y; XZ: x, z; YZ: y, z
    constexpr bool PML = DM != 0, DX = (DM & 1) != 0, DY = (DM & 2) != 0, DZ = (DM & 4) != 0;
    const bool px = DX && in_pml(t.i, g.nx, g.pml);
    const bool py = DY && in_pml(j0, g.ny, g.pml);
    C A0 = (C)1, B0 = (C)1, A1[NY], B1[NY];
    // step number keying stochastic rounding (16-bit storage only)
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
    // single-axis face boxes: old splits of the next plane loaded with its
    // other operands, as in the velocity sweep (the normal components' splits
    // on the damped axis, then the two shear components that have one)
    constexpr bool SPF = (SG_SPLIT_PREFETCH & 2) && (DM == 1 || DM == 2 || DM == 4);
    constexpr int AX = DM == 1 ? 0 : (DM == 2 ? 1 : 2);
    constexpr int SH0 = DM == 1 ? XY_X : (DM == 2 ? XY_Y : XZ_Z);
    constexpr int SH1 = DM == 1 ? XZ_X : (DM == 2 ? YZ_Y : YZ_Z);
    C spo[NY][5], spv[NY];
    auto load_splits = [&](int kk) {
        const bool pzk = DZ && in_pml(kk, g.nz, g.pml);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const bool an = (px || py || pzk) && t.rup[r];
            const unsigned o = an ? (unsigned)(split_slot(g, t.i, t.jr[r], kk) * (long long)sizeof(C)) : 0u;
#pragma unroll
            for (int c = 0; c < 3; ++c) spo[r][c] = bld(s.ss[XX_X + 3 * c + AX], o);
            spo[r][3] = bld(s.ss[SH0], o);
            spo[r][4] = bld(s.ss[SH1], o);
            spv[r] = s.vs ? bld(s.vs, o) : (C)0;
        }
    };
    if (SPF) load_splits(k0);

    for (int k = k0; k < k1; ++k) {
        auto &L = vplane(Ls, k);
        // x derivatives across lanes, before the staging registers are reused
        C dxl[3][NY];
        if (XL) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                dxl[0][r] = xbwd(qvx[r][4], G::use(h_vx[r], t.hok[r], fi_vx), s.cx);
                dxl[1][r] = xfwd(qvy[r][4], G::use(h_vy[r], t.hok[r], fi_vy), s.cx);
                dxl[2][r] = xfwd(qvz[r][5], G::use(h_vz[r], t.hok[r], fi_vz), s.cx);
            }
        }
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            L.vx[lr + 5][tx + 5] = qvx[r][4];
            L.vy[lr + 5][tx + 5] = qvy[r][4];
            L.vz[lr + 5][tx + 5] = qvz[r][5];
        }
        if (!XL && tx < 10) {
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                const int lr = ty * NY + r;
                L.vx[lr + 5][t.hc] = G::use(h_vx[r], t.hok[r], fi_vx);
                L.vy[lr + 5][t.hc] = G::use(h_vy[r], t.hok[r], fi_vy);
                L.vz[lr + 5][t.hc] = G::use(h_vz[r], t.hok[r], fi_vz);
            }
        }
#pragma unroll
        for (int m = 0; m < NH; ++m) {
            if (!YD && ty + BY * m < 10) {
                L.vx[t.ylr[m]][tx + 5] = G::use(y_vx[m], t.yok[m], fi_vx);
                L.vy[t.ylr[m]][tx + 5] = G::use(y_vy[m], t.yok[m], fi_vy);
                L.vz[t.ylr[m]][tx + 5] = G::use(y_vz[m], t.yok[m], fi_vz);
            }
        }
        if constexpr (YD) wait_vm_all();
        __syncthreads();

        // ---- read half of the read-modify-write of plane k: loaded one plane
        // ahead (RP), else issued first
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
            // TRANSIENT source: the stress update starts from sigma before
            // injection (saved: three stored values, then their residuals)
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
        C sn3[NY][3][3], s2[NY][3][2];  // old splits: normal [comp][x,y,z]; shear xy, xz, yz
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
                if (SPF) {
                    // normal splits on axis AX; shear: xy and xz store x first,
                    // xy and yz y, xz and yz z (the slot of each axis in s2)
                    vsr[r] = spv[r];
#pragma unroll
                    for (int c = 0; c < 3; ++c) sn3[r][c][AX] = spo[r][c];
                    if (DM == 1) { s2[r][0][0] = spo[r][3]; s2[r][1][0] = spo[r][4]; }
                    if (DM == 2) { s2[r][0][1] = spo[r][3]; s2[r][2][0] = spo[r][4]; }
                    if (DM == 4) { s2[r][1][1] = spo[r][3]; s2[r][2][1] = spo[r][4]; }
                    continue;
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

        // ---- loads for later.  First those used after this plane (queue
        // heads, mu of plane k+2): the vector-memory counter retires in issue
        // order, so waiting for them does not wait for plane k+1's staging
        // and RMW operands issued after them (first used after the next
        // barrier).
        const bool more = k + 1 < k1;
        S nvx[NY], nvy[NY], nvz[NY];
        const bool z6 = k + 6 < g.nz, z5 = k + 5 < g.nz;
        {
            const T *px_ = plane(Fvx, z6 ? k + 6 : k, sz), *py_ = plane(Fvy, z6 ? k + 6 : k, sz);
            const T *pz_ = plane(Fvz, z5 ? k + 5 : k, sz);
#pragma unroll
            for (int r = 0; r < NY; ++r) {
                nvx[r] = G::ld(px_, t.ro[r], t.rin[r] && z6, fi_vx);
                nvy[r] = G::ld(py_, t.ro[r], t.rin[r] && z6, fi_vy);
                nvz[r] = G::ld(pz_, t.ro[r], t.rin[r] && z5, fi_vz);
            }
        }
        // the last plane of the box reloads planes just read (cached) instead
        // of skipping the loads: with no branch around them the compiler's
        // wait for the queue heads counts them in every iteration
        const int kn = more ? k + 1 : k;
        fetch_mu(kn + 1);  // staged into the ring after this plane
        fetch(kn);
        if (RP) load_rmw(kn, spre, spres, lpre);
        if (SPF) load_splits(kn);

        auto &R0 = rmslot(Ls, k);
        auto &R1 = rmslot(Ls, k + 1);
#pragma unroll
        for (int r = 0; r < NY; ++r) {
            const int lr = ty * NY + r;
            const C dX = XL ? dxl[0][r] : lbwd<C>(&L.vx[lr + 5][0], tx + 5, 1, s.cx);
            const C dY = lbwd<C>(&L.vy[0][tx + 5], lr + 5, XW, s.cy);
            const C dZ = lbwd<C>(qvz[r], 5, 1, s.cz);
            const C dxy_x = XL ? dxl[1][r] : lfwd<C>(&L.vy[lr + 5][0], tx + 5, 1, s.cx);
            const C dxy_y = lfwd<C>(&L.vx[0][tx + 5], lr + 5, XW, s.cy);
            const C dxz_x = XL ? dxl[2][r] : lfwd<C>(&L.vz[lr + 5][0], tx + 5, 1, s.cx);
            const C dxz_z = lfwd<C>(qvx[r], 4, 1, s.cz);
            const C dyz_y = lfwd<C>(&L.vz[0][tx + 5], lr + 5, XW, s.cy);
            const C dyz_z = lfwd<C>(qvy[r], 4, 1, s.cz);
            const C r0 = R0[lr][tx], rx = R0[lr][tx + 1], ry = R0[lr + 1][tx];
            const C rxy = R0[lr + 1][tx + 1];
            const C rz = R1[lr][tx], rxz = R1[lr][tx + 1], ryz = R1[lr + 1][tx];
            const C muxy = (C)4 / (((r0 + rx) + ry) + rxy);
            const C muxz = (C)4 / (((r0 + rx) + rz) + rxz);
            const C muyz = (C)4 / (((r0 + ry) + rz) + ryz);
            const C dl = s.dt * lam[r], dl2m = s.dt * (lam[r] + (C)2 * muk[r]);
            const C dmxy = s.dt * muxy, dmxz = s.dt * muxz, dmyz = s.dt * muyz;
            C nw[6];
            nw[0] = old[0][r] + fmad(dl2m, dX, dl * (dY + dZ));
            nw[1] = old[1][r] + fmad(dl2m, dY, dl * (dX + dZ));
            nw[2] = old[2][r] + fmad(dl2m, dZ, dl * (dX + dY));
            nw[3] = fmad(dmxy, dxy_x + dxy_y, old[3][r]);
            nw[4] = fmad(dmxz, dxz_x + dxz_z, old[4][r]);
            nw[5] = fmad(dmyz, dyz_y + dyz_z, old[5][r]);
            C sn[3][3], t2[3][2];  // new splits: normal [comp][x,y,z]; shear xy, xz, yz
            if (PML) {
                // split update on every lane, selected where a point damps;
                // split stores follow the field stores
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
                    for (int a2 = 0; a2 < 3; ++a2) inc[a2] = (a2 == c ? dl2m : dl) * dnorm[a2];
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
        if constexpr (!IsDB<LS>::value) __syncthreads();

        // ---- ring and queues advance
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
                qvz[r][q] = qvz[r][q + 1];
            }
            qvx[r][9] = G::use(nvx[r], t.rin[r] && z6, fi_vx);
            qvy[r][9] = G::use(nvy[r], t.rin[r] && z6, fi_vy);
            qvz[r][9] = G::use(nvz[r], t.rin[r] && z5, fi_vz);
        }
    }
}

}  // namespace sg

#include "sg_tiled_lead.hpp"

namespace sg {

// Lead-plane staging (sg_tiled_lead.hpp) or staging at the updated plane,
// per sweep, storage size and work-list: bit l of a mask selects the lead
// form for the list with damping mask LIST_DM[l] (0, 1, 2, 4, 7).  Off by
// default: measured at 512x512x384 FP32 (interior lists, 4 waves per SIMD,
// lists one after another) the lead form reads 24 % (stress) and 17 %
// (velocity) fewer bytes but takes 1.063 / 0.914 ms against 1.035 /
// 0.846 ms: the plain sweeps run at the copy rate on their real traffic and
// the lead sweeps hit a latency limit of about the same size.
#ifndef SG_LEAD_V32
#define SG_LEAD_V32 0
#endif
#ifndef SG_LEAD_S32
#define SG_LEAD_S32 0
#endif
#ifndef SG_LEAD_V16
#define SG_LEAD_V16 0
#endif
#ifndef SG_LEAD_S16
#define SG_LEAD_S16 0
#endif
// (boxes of 4 rows, one row per thread only: the LDS ring holds one entry
// per thread and plane)
template <class T, int DM, int W, int TY, int NY>
struct LeadOn {
    static constexpr int bit = DM == 0 ? 0 : (DM == 1 ? 1 : (DM == 2 ? 2 : (DM == 4 ? 3 : 4)));
    static constexpr int mask = sizeof(T) == 8 ? 0
                                : sizeof(T) == 2 ? (W == 0 ? SG_LEAD_V16 : SG_LEAD_S16)
                                                 : (W == 0 ? SG_LEAD_V32 : SG_LEAD_S32);
    static constexpr bool v = ((mask >> bit) & 1) != 0 && TY == 4 && NY == 1;
};

// the LDS of a velocity / stress box for the chosen form
template <class T, class C, int TY, int NY, int DM>
struct VelLdsOf {
    static constexpr bool db = SG_DB != 0 || DmaOn<T, DM>::v;
    using type = typename std::conditional<LeadOn<T, DM, 0, TY, NY>::v, VelLdsL<C, TY>,
                                           typename std::conditional<db, VelLdsDB<C, TY>, VelLds<C, TY>>::type>::type;
};
template <class T, class C, int TY, int NY, int DM>
struct StrLdsOf {
    static constexpr bool db = SG_DB != 0 || DmaOn<T, DM>::v;
    using type = typename std::conditional<LeadOn<T, DM, 1, TY, NY>::v, StrLdsL<C, TY, DM == 0 ? 5 : 6>,
                                           typename std::conditional<db, StrLdsDB<C, TY>, StrLds<C, TY>>::type>::type;
};

template <class T, class C, int TY, int NY, int DM, bool CP, class LT>
__device__ __forceinline__ void vel_tile(const Shot<T, C> &s, const WorkItem &w, LT &L)
{
    if constexpr (LeadOn<T, DM, 0, TY, NY>::v)
        velocity_tile_ld<T, C, TY, NY, DM, CP>(s, w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
    else
        velocity_tile<T, C, TY, NY, DM, CP>(s, w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
}

template <class T, class C, int TY, int NY, int DM, bool CP, class LT>
__device__ __forceinline__ void str_tile(const Shot<T, C> &s, const WorkItem &w, LT &L)
{
    if constexpr (LeadOn<T, DM, 1, TY, NY>::v)
        stress_tile_ld<T, C, TY, NY, DM, CP>(s, w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
    else
        stress_tile<T, C, TY, NY, DM, CP>(s, w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
}

// ---------------------------------------------------------------------------
// kernels: a grid-stride loop over a work-list (one or many shots)
// ---------------------------------------------------------------------------
// Register caps (minimum waves per SIMD) of the face instantiations (DM 1,
// 2, 4: SG_PML_WAVES; 16-bit storage: SG_PML_WAVES16, velocity faces
// SG_PML_WAVES16_V) and of the interior
// ones (SG_INT_WAVES; 16-bit storage: SG_INT_WAVES16); edges and corners
// are never capped
#ifndef SG_PML_WAVES
#define SG_PML_WAVES 1
#endif
// FP32 velocity faces: capped at 3 waves they gained before the use-site
// conversion (512x512x384 27.70 -> 27.79, 256^3 20.55 -> 20.91 Gpt/s); since
// it they run faster uncapped (velocity sweep 1.5069 -> 1.4986 ms, 256^3
// 24.22 -> 24.45 Gpt/s); the stress faces lose with a cap (1.895 -> 1.908 ms
// at 3 waves)
#ifndef SG_PML_WAVES_V
#define SG_PML_WAVES_V 1
#endif
// 16-bit storage, measured at 512x512x384 and 500^3 (FP16 31.3 -> 32.0 and
// 37.7 -> 39.4 Gpt/s, BF16 30.5 -> 31.5; FP32 unaffected): faces at 3 waves
// (the x-face kernels need 162-166 VGPRs then, 169-174 uncapped, no spill),
// the interior velocity sweep at 4 (128 VGPRs, a 12-byte spill)
#ifndef SG_PML_WAVES16
#define SG_PML_WAVES16 3
#endif
// 16-bit velocity faces uncapped since the use-site conversion (512x512x384
// FP16, 400 steps: velocity sweep 1.1645 -> 1.1356 ms uncapped, while the
// stress sweep wants its faces at 3 waves: 1.5455 against 1.5764 ms)
#ifndef SG_PML_WAVES16_V
#define SG_PML_WAVES16_V 1
#endif
#ifndef SG_INT_WAVES
#define SG_INT_WAVES 1
#endif
#ifndef SG_INT_WAVES16
#define SG_INT_WAVES16 1
#endif
#ifndef SG_INT_WAVES16_V
#define SG_INT_WAVES16_V 4
#endif
// W: 0 velocity, 1 stress (the 16-bit interior velocity sweep has its own
// cap); the compensated 16-bit kernels (CP) are never capped (they would
// spill: 92-120 bytes per lane at 4 waves)
template <class T, int DM, int NY, int W, bool CP>
struct WavesCap {
    static constexpr bool F32 = sizeof(T) == 4, H = sizeof(T) == 2 && !CP;
    static constexpr int i16 = W == 0 ? SG_INT_WAVES16_V : SG_INT_WAVES16;
    static constexpr int f32 = W == 0 ? SG_PML_WAVES_V : SG_PML_WAVES;
    static constexpr int h16 = W == 0 ? SG_PML_WAVES16_V : SG_PML_WAVES16;
    static constexpr int v = DM == 7 ? 1
                             : DM   ? (F32 ? f32 : (H ? h16 : 1))
                                    : (NY == 1 ? (F32 ? SG_INT_WAVES : (H ? i16 : 1)) : 1);
};

template <class T, class C, int TY, int NY, int DM, bool CP>
__global__ __launch_bounds__(TX *TY / NY) __attribute__((amdgpu_waves_per_eu(WavesCap<T, DM, NY, 0, CP>::v))) void k_velocity_wl(const Shot<T, C> *__restrict__ shots,
                                                            const WorkItem *__restrict__ items, int nitems)
{
    __shared__ typename VelLdsOf<T, C, TY, NY, DM>::type L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) vel_tile<T, C, TY, NY, DM, CP>(shots[w.shot], w, L);
        __syncthreads();
    }
}

template <class T, class C, int TY, int NY, int DM, bool CP>
__global__ __launch_bounds__(TX *TY / NY) __attribute__((amdgpu_waves_per_eu(WavesCap<T, DM, NY, 1, CP>::v))) void k_stress_wl(const Shot<T, C> *__restrict__ shots,
                                                          const WorkItem *__restrict__ items, int nitems)
{
    __shared__ typename StrLdsOf<T, C, TY, NY, DM>::type L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) str_tile<T, C, TY, NY, DM, CP>(shots[w.shot], w, L);
        __syncthreads();
    }
}

template <class T, class C, int TY, int NY, int DM, bool CP>
static hipError_t sweep_tnd(int which, const void *shots, const WorkItem *items, int nitems, int grid,
                            hipStream_t st)
{
    const Shot<T, C> *sh = static_cast<const Shot<T, C> *>(shots);
    const dim3 block(TX, TY / NY), g(grid);
    if (which == 0) hipLaunchKernelGGL((k_velocity_wl<T, C, TY, NY, DM, CP>), g, block, 0, st, sh, items, nitems);
    else hipLaunchKernelGGL((k_stress_wl<T, C, TY, NY, DM, CP>), g, block, 0, st, sh, items, nitems);
    return hipGetLastError();
}

// Interior boxes with the staging rows loaded straight into LDS (FP32): the
// same arithmetic as k_velocity_wl / k_stress_wl with damping mask 0.  The
// solver launches them for grids without PML, where the interior list runs
// alone (measured on MI355X, 500^3: 30.28 -> 31.31 Gpt/s; beside the PML
// lists they gain nothing).
template <class T, class C, int TY, int NY>
__global__ __launch_bounds__(TX *TY / NY) __attribute__((amdgpu_waves_per_eu(WavesCap<T, 0, NY, 0, false>::v))) void k_velocity_wl_lds(
    const Shot<T, C> *__restrict__ shots, const WorkItem *__restrict__ items, int nitems)
{
    __shared__ VelLdsDMA<C, TY> L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) velocity_tile<T, C, TY, NY, 0, false>(shots[w.shot], w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
        __syncthreads();
    }
}

template <class T, class C, int TY, int NY>
__global__ __launch_bounds__(TX *TY / NY) __attribute__((amdgpu_waves_per_eu(WavesCap<T, 0, NY, 1, false>::v))) void k_stress_wl_lds(
    const Shot<T, C> *__restrict__ shots, const WorkItem *__restrict__ items, int nitems)
{
    __shared__ StrLdsDMA<C, TY> L;
    for (int p = blockIdx.x; p < nitems; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) stress_tile<T, C, TY, NY, 0, false>(shots[w.shot], w.i0, w.i1, w.j0, w.j1, w.k0, w.k1, L);
        __syncthreads();
    }
}

template <class T, class C, int TY, int NY, bool CP>
static hipError_t sweep_tnc(int which, int dm, const void *shots, const WorkItem *items, int nitems, int grid,
                            hipStream_t st)
{
    switch (dm) {
    case DM_INTERIOR_LDS:
        if constexpr (std::is_same<T, float>::value && !CP) {
            const Shot<T, C> *sh = static_cast<const Shot<T, C> *>(shots);
            const dim3 block(TX, TY / NY), g(grid);
            if (which == 0) hipLaunchKernelGGL((k_velocity_wl_lds<T, C, TY, NY>), g, block, 0, st, sh, items, nitems);
            else hipLaunchKernelGGL((k_stress_wl_lds<T, C, TY, NY>), g, block, 0, st, sh, items, nitems);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    case 0: return sweep_tnd<T, C, TY, NY, 0, CP>(which, shots, items, nitems, grid, st);
    case 1: return sweep_tnd<T, C, TY, NY, 1, CP>(which, shots, items, nitems, grid, st);
    case 2: return sweep_tnd<T, C, TY, NY, 2, CP>(which, shots, items, nitems, grid, st);
    case 4: return sweep_tnd<T, C, TY, NY, 4, CP>(which, shots, items, nitems, grid, st);
    case 7: return sweep_tnd<T, C, TY, NY, 7, CP>(which, shots, items, nitems, grid, st);
    default: return hipErrorInvalidValue;
    }
}

// compensated instantiations exist for 16-bit storage only
template <class T, class C, int TY, int NY>
static hipError_t sweep_tn(int which, int dm, bool comp, const void *shots, const WorkItem *items, int nitems,
                           int grid, hipStream_t st)
{
    if constexpr (sizeof(T) == 2) {
        if (comp) return sweep_tnc<T, C, TY, NY, true>(which, dm, shots, items, nitems, grid, st);
    } else {
        if (comp) return hipErrorInvalidValue;
    }
    return sweep_tnc<T, C, TY, NY, false>(which, dm, shots, items, nitems, grid, st);
}

// supported (tile rows, rows per thread): (4,1) (8,1) (8,2) (16,1) (16,2);
// damping masks 0, 1, 2, 4, 7, and DM_INTERIOR_LDS (interior boxes with the
// staging rows loaded straight into LDS, FP32); comp: compensated time
// update (16-bit only)
template <class T, class C>
hipError_t sweep_t(int which, int ty, int ny, int dm, bool comp, const void *shots, const WorkItem *items,
                   int nitems, int grid, hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (grid <= 0 || grid > nitems) grid = nitems;
    if (ty == 4 && ny == 1) return sweep_tn<T, C, 4, 1>(which, dm, comp, shots, items, nitems, grid, st);
    if (ty == 8 && ny == 1) return sweep_tn<T, C, 8, 1>(which, dm, comp, shots, items, nitems, grid, st);
    if (ty == 8 && ny == 2) return sweep_tn<T, C, 8, 2>(which, dm, comp, shots, items, nitems, grid, st);
    if (ty == 16 && ny == 1) return sweep_tn<T, C, 16, 1>(which, dm, comp, shots, items, nitems, grid, st);
    if (ty == 16 && ny == 2) return sweep_tn<T, C, 16, 2>(which, dm, comp, shots, items, nitems, grid, st);
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// all five work-lists of a sweep in one launch (small, launch-bound grids):
// each item's list, found from the list offsets, selects the damping-mask
// instantiation; the per-item code is that of the separate launches, so
// results are bit-identical.  (4,1) boxes for interior and PML alike.
// ---------------------------------------------------------------------------
struct ListOffsets {
    int off[NLIST + 1];
};

__device__ __forceinline__ int list_of_item(const ListOffsets &lo, int p)
{
    int l = 0;
#pragma unroll
    for (int q = 1; q < NLIST; ++q) l += p >= lo.off[q] ? 1 : 0;
    return l;
}

// LDS of a fused launch: large enough for every list's form
template <class T, class C, int W>
struct FusedLds {
    static constexpr int mask = sizeof(T) == 8 ? 0
                                : sizeof(T) == 2 ? (W == 0 ? SG_LEAD_V16 : SG_LEAD_S16)
                                                 : (W == 0 ? SG_LEAD_V32 : SG_LEAD_S32);
    // (the fused launch keeps single-buffered staging, so its interior boxes
    // stage through registers: double-buffered, 100^3 P10 ran 0.5 % slower)
    static constexpr bool db = SG_DB != 0;
    using type = typename std::conditional<
        W == 0,
        typename std::conditional<(mask & 0x1F) != 0, VelLdsL<C, 4>,
                                  typename std::conditional<db, VelLdsDB<C, 4>, VelLds<C, 4>>::type>::type,
        typename std::conditional<(mask & 0x1F) != 0, StrLdsL<C, 4, 6>,
                                  typename std::conditional<db, StrLdsDB<C, 4>, StrLds<C, 4>>::type>::type>::type;
};

// WV: minimum waves per SIMD the register allocation must allow (0: no
// bound).  One launch carries the registers of its heaviest list (edges and
// corners), which otherwise caps the occupancy of every box.
template <class T, class C, bool CP, int WV>
__global__ __launch_bounds__(TX * 4) __attribute__((amdgpu_waves_per_eu(WV > 0 ? WV : 1))) void k_velocity_fused(
    const Shot<T, C> *__restrict__ shots, const WorkItem *__restrict__ items, ListOffsets lo)
{
    __shared__ typename FusedLds<T, C, 0>::type L;
    for (int p = blockIdx.x; p < lo.off[NLIST]; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) {
            const Shot<T, C> &s = shots[w.shot];
            switch (list_of_item(lo, p)) {
            case 0: vel_tile<T, C, 4, 1, 0, CP>(s, w, L); break;
            case 1: vel_tile<T, C, 4, 1, 1, CP>(s, w, L); break;
            case 2: vel_tile<T, C, 4, 1, 2, CP>(s, w, L); break;
            case 3: vel_tile<T, C, 4, 1, 4, CP>(s, w, L); break;
            default: vel_tile<T, C, 4, 1, 7, CP>(s, w, L); break;
            }
        }
        __syncthreads();
    }
}

template <class T, class C, bool CP, int WV>
__global__ __launch_bounds__(TX * 4) __attribute__((amdgpu_waves_per_eu(WV > 0 ? WV : 1))) void k_stress_fused(
    const Shot<T, C> *__restrict__ shots, const WorkItem *__restrict__ items, ListOffsets lo)
{
    __shared__ typename FusedLds<T, C, 1>::type L;
    for (int p = blockIdx.x; p < lo.off[NLIST]; p += gridDim.x) {
        const WorkItem w = items[p];
        if (w.shot >= 0) {
            const Shot<T, C> &s = shots[w.shot];
            switch (list_of_item(lo, p)) {
            case 0: str_tile<T, C, 4, 1, 0, CP>(s, w, L); break;
            case 1: str_tile<T, C, 4, 1, 1, CP>(s, w, L); break;
            case 2: str_tile<T, C, 4, 1, 2, CP>(s, w, L); break;
            case 3: str_tile<T, C, 4, 1, 4, CP>(s, w, L); break;
            default: str_tile<T, C, 4, 1, 7, CP>(s, w, L); break;
            }
        }
        __syncthreads();
    }
}

template <class T, class C, bool CP, int WV>
static void launch_fused(int which, const Shot<T, C> *sh, const WorkItem *items, const ListOffsets &lo, int grid,
                         hipStream_t st)
{
    const dim3 block(TX, 4), g(grid);
    if (which == 0) hipLaunchKernelGGL((k_velocity_fused<T, C, CP, WV>), g, block, 0, st, sh, items, lo);
    else hipLaunchKernelGGL((k_stress_fused<T, C, CP, WV>), g, block, 0, st, sh, items, lo);
}

// off: list offsets into items (off[NLIST] items in all); waves: the
// register bound of the launch (0 none; 2, 3 or 4 waves per SIMD, FP32
// storage only)
template <class T, class C>
hipError_t sweep_fused_t(int which, bool comp, int waves, const void *shots, const WorkItem *items, const int *off,
                         int grid, hipStream_t st)
{
    ListOffsets lo;
    for (int l = 0; l <= NLIST; ++l) lo.off[l] = off[l];
    const int n = lo.off[NLIST];
    if (n <= 0) return hipSuccess;
    if (grid <= 0 || grid > n) grid = n;
    const Shot<T, C> *sh = static_cast<const Shot<T, C> *>(shots);
    if constexpr (sizeof(T) == 2) {
        if (waves != 0) return hipErrorInvalidValue;
        if (comp) launch_fused<T, C, true, 0>(which, sh, items, lo, grid, st);
        else launch_fused<T, C, false, 0>(which, sh, items, lo, grid, st);
    } else if constexpr (sizeof(T) == 8) {
        if (comp || waves != 0) return hipErrorInvalidValue;
        launch_fused<T, C, false, 0>(which, sh, items, lo, grid, st);
    } else {
        if (comp) return hipErrorInvalidValue;
        switch (waves) {
        case 0: launch_fused<T, C, false, 0>(which, sh, items, lo, grid, st); break;
        case 2: launch_fused<T, C, false, 2>(which, sh, items, lo, grid, st); break;
        case 3: launch_fused<T, C, false, 3>(which, sh, items, lo, grid, st); break;
        case 4: launch_fused<T, C, false, 4>(which, sh, items, lo, grid, st); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

}  // namespace sg
