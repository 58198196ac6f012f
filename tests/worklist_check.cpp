// Host check of the tiled sweeps' work-lists: for a geometry and box shapes
// given on the command line, the interior and PML lists together cover every
// updated point [5, n-5)^3 exactly once, no box exceeds 64 x ty (or pty)
// points in x/y or crosses a 64-column boundary, interior boxes hold no
// damped point and PML boxes at least one, and x tiles start at multiples of
// 64.  Prints "ok <n_interior> <n_pml>" or the first violation.
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
    std::vector<sg::WorkItem> in, pml;
    sg::build_worklists(g, 0, ty, pty, zc, in, pml);
    std::vector<unsigned char> hit((size_t)g.nx * g.ny * g.nz, 0);
    auto damped = [&](int i, int j, int k) {
        return sg::in_pml(i, g.nx, g.pml) || sg::in_pml(j, g.ny, g.pml) || sg::in_pml(k, g.nz, g.pml);
    };
    for (int list = 0; list < 2; ++list) {
        const std::vector<sg::WorkItem> &v = list ? pml : in;
        for (const sg::WorkItem &w : v) {
            const int h = list ? pty : ty;
            if (w.i0 % 64 || w.i1 - w.i0 > 64 || w.i1 <= w.i0 || w.j1 - w.j0 > h || w.j1 <= w.j0 ||
                w.k1 - w.k0 > std::max(zc, std::min(g.pml, 32)) || w.k1 <= w.k0 || w.i1 > g.nx - 5 || w.j0 < 5 || w.j1 > g.ny - 5 ||
                w.k0 < 5 || w.k1 > g.nz - 5) {
                std::printf("bad box %d %d %d %d %d %d\n", w.i0, w.i1, w.j0, w.j1, w.k0, w.k1);
                return 1;
            }
            bool any = false;
            for (int k = w.k0; k < w.k1; ++k)
                for (int j = w.j0; j < w.j1; ++j)
                    for (int i = w.i0 < 5 ? 5 : w.i0; i < w.i1; ++i) {
                        const bool d = damped(i, j, k);
                        any = any || d;
                        if (!list && d) {
                            std::printf("damped point %d %d %d in an interior box\n", i, j, k);
                            return 1;
                        }
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
            if (list && !any) {
                std::printf("PML box without a damped point\n");
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
    std::printf("ok %zu %zu\n", in.size(), pml.size());
    return 0;
}
