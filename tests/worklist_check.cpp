// Host check of the tiled sweeps' work-lists: for a geometry and box shapes
// given on the command line, the five lists (interior, x/y/z face, edges and
// corners) together cover every updated point [5, n-5)^3 exactly once, no
// box exceeds 64 x ty (or pty) points in x/y or crosses a 64-column
// boundary, x tiles start at multiples of 64, y/z damping is uniform inside
// a box, and the axes that damp at some point of a box are exactly those of
// its list's mask (edges/corners: two or more); every damped point has its
// own PML split slot inside the split store.  Prints
// "ok <n_interior> <n_x> <n_y> <n_z> <n_edge>" or the first violation.
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "sg_worklist.hpp"

int main(int argc, char **argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: nx ny nz pml ty pty zc\n");
        return 2;
    }
    sg::Geom g{};
    g.nx = std::atoi(argv[1]);
    g.ny = std::atoi(argv[2]);
    g.nz = std::atoi(argv[3]);
    g.pml = std::atoi(argv[4]);
    g.ux = g.nx - 10;
    g.uy = g.ny - 10;
    g.uz = g.nz - 10;
    const int ty = std::atoi(argv[5]), pty = std::atoi(argv[6]), zc = std::atoi(argv[7]);
    sg::WorkLists wl;
    sg::build_worklists(g, 0, ty, pty, zc, wl);
    std::vector<unsigned char> hit((size_t)g.nx * g.ny * g.nz, 0);
    for (int list = 0; list < sg::NLIST; ++list) {
        const std::vector<sg::WorkItem> &v = wl.l[list];
        for (const sg::WorkItem &w : v) {
            const int h = list ? pty : ty;
            if (w.i0 % 64 || w.i1 - w.i0 > 64 || w.i1 <= w.i0 || w.j1 - w.j0 > h || w.j1 <= w.j0 ||
                w.k1 - w.k0 > std::max(zc, std::min(g.pml, 32)) || w.k1 <= w.k0 || w.i1 > g.nx - 5 || w.j0 < 5 || w.j1 > g.ny - 5 ||
                w.k0 < 5 || w.k1 > g.nz - 5) {
                std::printf("bad box %d %d %d %d %d %d\n", w.i0, w.i1, w.j0, w.j1, w.k0, w.k1);
                return 1;
            }
            int mask = 0;
            for (int k = w.k0; k < w.k1; ++k)
                for (int j = w.j0; j < w.j1; ++j)
                    for (int i = w.i0 < 5 ? 5 : w.i0; i < w.i1; ++i) {
                        mask |= (sg::in_pml(i, g.nx, g.pml) ? 1 : 0) | (sg::in_pml(j, g.ny, g.pml) ? 2 : 0) |
                                (sg::in_pml(k, g.nz, g.pml) ? 4 : 0);
                        // y/z damping must be uniform inside a box
                        if (sg::in_pml(j, g.ny, g.pml) != sg::in_pml(w.j0, g.ny, g.pml) ||
                            sg::in_pml(k, g.nz, g.pml) != sg::in_pml(w.k0, g.nz, g.pml)) {
                            std::printf("box straddles a y/z PML plane at %d %d %d\n", i, j, k);
                            return 1;
                        }
                        unsigned char &c = hit[((size_t)k * g.ny + j) * g.nx + i];
                        if (c) {
                            std::printf("point %d %d %d covered twice\n", i, j, k);
                            return 1;
                        }
                        c = 1;
                    }
            const bool ok = list < 4 ? mask == sg::LIST_DM[list] : (mask != 0 && (mask & (mask - 1)) != 0);
            if (!ok) {
                std::printf("box in list %d has damping mask %d\n", list, mask);
                return 1;
            }
        }
    }
    for (int k = 5; k < g.nz - 5; ++k)
        for (int j = 5; j < g.ny - 5; ++j)
            for (int i = 5; i < g.nx - 5; ++i)
                if (!hit[((size_t)k * g.ny + j) * g.nx + i]) {
                    std::printf("point %d %d %d not covered\n", i, j, k);
                    return 1;
                }
    // split store: every damped updated point has its own slot in [0, nslots)
    sg::split_layout(g);
    std::vector<unsigned char> used((size_t)g.nslots, 0);
    for (int k = 5; k < g.nz - 5; ++k)
        for (int j = 5; j < g.ny - 5; ++j)
            for (int i = 5; i < g.nx - 5; ++i) {
                if (!(sg::in_pml(i, g.nx, g.pml) || sg::in_pml(j, g.ny, g.pml) || sg::in_pml(k, g.nz, g.pml)))
                    continue;
                const long long sl = sg::split_slot(g, i, j, k);
                if (sl < 0 || sl >= g.nslots) {
                    std::printf("split slot %lld of %d %d %d outside [0, %lld)\n", sl, i, j, k, g.nslots);
                    return 1;
                }
                if (used[(size_t)sl]) {
                    std::printf("split slot %lld used twice (%d %d %d)\n", sl, i, j, k);
                    return 1;
                }
                used[(size_t)sl] = 1;
            }
    std::printf("ok %zu %zu %zu %zu %zu\n", wl.l[0].size(), wl.l[1].size(), wl.l[2].size(), wl.l[3].size(),
                wl.l[4].size());
    return 0;
}
