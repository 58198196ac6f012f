// sg_worklist.hpp -- host-side construction of the tile work-lists of the
// tiled sweeps (shared by the single-shot solver and the batch).
//
// A shot's updated box is cut into 64-wide x tiles, TY-high y tiles and z
// chunks whose boundaries include the two PML planes k = 5+P and n-5-P.  A
// work item is INTERIOR when its x, y and z ranges lie inside the interior
// (no PML point), else PML; the two kinds run different kernel
// instantiations.  Items of each kind are then interleaved so that item
// position p lands on XCD p % 8 (blocks are dispatched round-robin over the
// 8 XCDs) while every XCD walks a contiguous band of tiles.
#pragma once

#include <algorithm>
#include <vector>

#include "sg_internal.hpp"

namespace sg {

// z chunk length giving about `target` work items for the shot (>= 16 planes)
inline int auto_zchunk(const Geom &g, int ty, int target = 2048)
{
    const long long cols = (long long)((g.ux + 63) / 64) * ((g.uy + ty - 1) / ty);
    long long nbz = (target + cols - 1) / cols;
    if (nbz < 1) nbz = 1;
    int zc = (int)((g.uz + nbz - 1) / nbz);
    return std::max(16, std::min(zc, g.uz));
}

inline void build_worklists(const Geom &g, int shot, int ty, int zc, std::vector<WorkItem> &interior,
                            std::vector<WorkItem> &pml)
{
    const int P = g.pml;
    const int nbx = (g.ux + 63) / 64, nby = (g.uy + ty - 1) / ty;
    // z ranges: [5, 5+P), interior chunks, [nz-5-P, nz-5)
    std::vector<std::pair<int, int>> zr;
    const int zlo = HALO + P, zhi = g.nz - HALO - P;
    for (int k = HALO; k < zlo; k += zc) zr.push_back({k, std::min(k + zc, zlo)});
    for (int k = zlo; k < zhi; k += zc) zr.push_back({k, std::min(k + zc, zhi)});
    for (int k = std::max(zhi, zlo); k < g.nz - HALO; k += zc) zr.push_back({k, std::min(k + zc, g.nz - HALO)});
    for (const auto &z : zr) {
        const bool zin = z.first >= zlo && z.second <= zhi;
        for (int by = 0; by < nby; ++by) {
            const int j0 = HALO + by * ty, j1 = std::min(j0 + ty, g.ny - HALO);
            const bool yin = j0 >= HALO + P && j1 <= g.ny - HALO - P;
            for (int bx = 0; bx < nbx; ++bx) {
                const int i0 = HALO + bx * 64, i1 = std::min(i0 + 64, g.nx - HALO);
                const bool xin = i0 >= HALO + P && i1 <= g.nx - HALO - P;
                const WorkItem w{shot, bx, by, z.first, z.second};
                (zin && yin && xin ? interior : pml).push_back(w);
            }
        }
    }
}

// interleave for the XCD round-robin; pads with shot = -1 to a multiple of 8
inline std::vector<WorkItem> xcd_order(const std::vector<WorkItem> &seq)
{
    const int n = (int)seq.size();
    const int per = (n + 7) / 8;
    std::vector<WorkItem> tab((size_t)per * 8, WorkItem{-1, 0, 0, 0, 0});
    for (int s = 0; s < n; ++s) tab[(size_t)(s % per) * 8 + s / per] = seq[s];
    return tab;
}

}  // namespace sg
