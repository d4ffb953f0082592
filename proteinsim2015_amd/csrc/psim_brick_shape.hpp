// psim_brick_shape.hpp -- brick shapes of the sparse pair kernel (host code,
// no HIP dependency; shared by the kernels' launch code and the C ABI).
//
// A brick is bx x by x bz home cells of one system's S^3 grid; the kernel
// stages the brick's halo, (bx+2)(by+2)(bz+2) cells, in LDS.  The limits
// below are the kernel's LDS table sizes: every shape brick_shape() returns
// fits them, whatever the occupancy.
#pragma once

#include <algorithm>
#include <cmath>

namespace psim {

constexpr int kBrickHaloCap = 768;     // halo cells per brick
constexpr int kBrickRowCap = 128;      // home rows (by * bz) per brick
constexpr int kBrickMaxEdge = 16;      // home cells along one edge
#ifndef PSIM_BRICK_STAGE_CAP
#define PSIM_BRICK_STAGE_CAP 768
#endif
constexpr int kBrickStageCap = PSIM_BRICK_STAGE_CAP;   // staged particles per chunk

struct BrickGeo {
    int bx, by, bz;      // home cells per brick along x, y, z
    int nbx, nby, nbz;   // bricks along x, y, z (the last one may be thinner)
    int nbricks;
};

// A brick edge that splits S cells into equal parts (up to one cell).
inline bool brick_edge_balanced(int S, int b)
{
    const int parts = (S + b - 1) / b;
    return b == (S + parts - 1) / parts;
}

// The shape for a system of S^3 cells at `occupancy` particles per cell:
// about `target` home particles per brick, small halo-to-home ratio, and an
// expected halo population within half the staging capacity where possible.
inline BrickGeo brick_shape(int S, double occupancy, double target)
{
    const double occ = std::max(occupancy, 1e-6);
    const double want = std::max(1.0, target / occ);   // home cells wanted
    const int emax = std::min(S, kBrickMaxEdge);
    BrickGeo best{1, 1, 1, S, S, S, S * S * S};
    double best_score = 1e300;
    for (int bx = 1; bx <= emax; ++bx) {
        if (!brick_edge_balanced(S, bx)) continue;
        for (int by = 1; by <= bx; ++by) {
            if (!brick_edge_balanced(S, by)) continue;
            for (int bz = 1; bz <= by; ++bz) {
                if (!brick_edge_balanced(S, bz)) continue;
                const int halo = (bx + 2) * (by + 2) * (bz + 2);
                if (halo > kBrickHaloCap || by * bz > kBrickRowCap) continue;
                const double cells = (double)bx * by * bz;
                double score = std::fabs(std::log(cells / want)) +
                               0.5 * std::log((double)halo / cells);
                if (halo * occ > 0.5 * kBrickStageCap) score += 10.0;
                if (score < best_score) {
                    best_score = score;
                    best.bx = bx;
                    best.by = by;
                    best.bz = bz;
                }
            }
        }
    }
    best.nbx = (S + best.bx - 1) / best.bx;
    best.nby = (S + best.by - 1) / best.by;
    best.nbz = (S + best.bz - 1) / best.bz;
    best.nbricks = best.nbx * best.nby * best.nbz;
    return best;
}

}  // namespace psim
