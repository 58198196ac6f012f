// sg_worklist.hpp -- host-side construction of the work-lists of the tiled
// sweeps (shared by the single-shot solver and the batch).
//
// The updated range [5, n-5) of a shot is cut into boxes:
//  * x: tiles of 64 columns at multiples of 64 (rows start on cache-line
//    boundaries; the halo columns i < 5 are never updated); a tile holding
//    x-PML columns damps per lane;
//  * y and z: cut at the two PML planes (5+P and n-5-P), each of the three
//    pieces then split into near-equal parts of at most TY rows / zc planes
//    (z PML slabs: max(zc, min(P, 32)) planes), so damping in y and z is
//    uniform per box and row/plane.
// Boxes are listed by the set of axes that can damp inside them: interior
// (none), x, y or z face, edges and corners (several); each list runs its own
// kernel instantiation.  Each list is interleaved so that item position p
// lands on XCD p % 8 (blocks are dispatched round-robin over the 8 XCDs)
// while every XCD walks a contiguous band of boxes.
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

#include "sg_internal.hpp"

namespace sg {

struct Seg {
    int a, b;
    bool pml;
};

// [lo, hi) in ceil(len / w) near-equal pieces
inline void cut_even(std::vector<Seg> &out, int lo, int hi, int w, bool pml)
{
    if (hi <= lo) return;
    const int len = hi - lo, n = (len + w - 1) / w;
    for (int p = 0; p < n; ++p) out.push_back(Seg{lo + (int)((long long)len * p / n), lo + (int)((long long)len * (p + 1) / n), pml});
}

// pieces of at most w (interior) / wp (PML slabs) along one axis
inline std::vector<Seg> axis_segments(int n, int pml, int w, int wp)
{
    std::vector<Seg> s;
    const int lo = HALO, p0 = std::min(HALO + pml, n - HALO), p1 = std::max(n - HALO - pml, p0), hi = n - HALO;
    cut_even(s, lo, p0, wp, true);
    cut_even(s, p0, p1, w, false);
    cut_even(s, p1, hi, wp, true);
    return s;
}

// x tiles [64m, min(64m+64, n-5)); the kernels never update the halo
// columns i < 5, so the first tile starts at 0 and every tile's rows start on
// a 256-byte boundary.  pml = the tile holds a damped column.
inline std::vector<Seg> x_tiles(int n, int pml)
{
    std::vector<Seg> s;
    const int hi = n - HALO, p0 = HALO + pml, p1 = n - HALO - pml;
    for (int a = 0; a < hi; a += 64) {
        const int b = std::min(a + 64, hi);
        s.push_back(Seg{a, b, std::max(a, HALO) < p0 || b > p1});
    }
    return s;
}

// z chunk length.  The z PML slabs stay whole (up to 32 planes), so with
// chunks as thick as the slabs every box of a list marches the same number
// of planes and no box trails the launch: P planes when 8 <= P <= 32
// (measured on MI355X: 100^3 P10 8.3 -> 11.1 Gpt/s against 8-plane chunks,
// 256^3 P16 20.9 -> 22.2 against 24, 512x512x384 P20 unchanged).  Else 24
// planes (ten planes of z-queue warm-up per chunk against load balance),
// shortened down to 8 while the shot would give fewer than `target` boxes.
inline int auto_zchunk(const Geom &g, int ty, int target = 2048)
{
    if (g.pml >= 8 && g.pml <= 32) return std::max(1, std::min(g.pml, g.uz));
    const long long cols = (long long)((g.ux + 63) / 64) * ((g.uy + ty - 1) / ty);
    int zc = 24;
    while (zc > 8 && cols * ((g.uz + zc - 1) / zc) < target) zc -= 4;
    return std::max(1, std::min(zc, g.uz));
}

// interleave for the XCD round-robin; pads with shot = -1 to a multiple of 8
inline std::vector<WorkItem> xcd_order(const std::vector<WorkItem> &seq)
{
    const int n = (int)seq.size();
    const int per = (n + 7) / 8;
    std::vector<WorkItem> tab((size_t)per * 8, WorkItem{-1, 0, 0, 0, 0, 0, 0});
    for (int s = 0; s < n; ++s) tab[(size_t)(s % per) * 8 + s / per] = seq[s];
    return tab;
}

// One list per damping mask (LIST_DM order: interior, x, y, z faces, edges
// and corners).
struct WorkLists {
    std::vector<WorkItem> l[NLIST];
};

inline int list_of_mask(int m)
{
    switch (m) {
    case 0: return 0;
    case 1: return 1;
    case 2: return 2;
    case 4: return 3;
    default: return 4;
    }
}

// Interior boxes use y pieces of at most `ty` rows, PML boxes of at most
// `pty`; both sets of y cuts include the PML planes, so the lists partition
// the updated range.
inline void build_worklists(const Geom &g, int shot, int ty, int pty, int zc, WorkLists &wl)
{
    const std::vector<Seg> xs = x_tiles(g.nx, g.pml);
    const std::vector<Seg> ysi = axis_segments(g.ny, g.pml, ty, ty);
    const std::vector<Seg> ysp = axis_segments(g.ny, g.pml, pty, pty);
    // a z piece costs ten extra planes of queue warm-up: PML slabs stay whole
    // up to 32 planes
    const std::vector<Seg> zs = axis_segments(g.nz, g.pml, zc, std::max(zc, std::min(g.pml, 32)));
    for (const Seg &z : zs) {
        for (const Seg &y : ysi)
            for (const Seg &x : xs)
                if (!(x.pml || y.pml || z.pml)) wl.l[0].push_back(WorkItem{shot, x.a, x.b, y.a, y.b, z.a, z.b});
        for (const Seg &y : ysp)
            for (const Seg &x : xs) {
                const int m = (x.pml ? 1 : 0) | (y.pml ? 2 : 0) | (z.pml ? 4 : 0);
                if (m) wl.l[list_of_mask(m)].push_back(WorkItem{shot, x.a, x.b, y.a, y.b, z.a, z.b});
            }
    }
}

// XCD-ordered lists back to back; list l occupies [off[l], off[l+1])
inline std::vector<WorkItem> pack_worklists(const WorkLists &wl, int off[NLIST + 1])
{
    std::vector<WorkItem> all;
    off[0] = 0;
    for (int l = 0; l < NLIST; ++l) {
        const std::vector<WorkItem> t = xcd_order(wl.l[l]);
        all.insert(all.end(), t.begin(), t.end());
        off[l + 1] = (int)all.size();
    }
    return all;
}

}  // namespace sg
