// psim_pairs.hpp -- the 27-cell pair gather shared by every force path.
//
// One lane per target particle i.  The lane walks the 27 neighbour cells in
// the reference's order (x outer, z inner; LJPotential.cpp:154-177) and culls
// candidates with a cheap FP64 distance on the periodic image of the
// neighbour cell (shift applied once per cell).  Survivors go to a per-lane
// hit list in LDS; the list is drained (pair forces evaluated) whenever any
// lane of the wave fills it, so the expensive pair arithmetic runs with
// (nearly) all lanes busy instead of under a sparse divergent mask.
//
// Neighbour sets are bit-exact with the reference: the cull accepts with a
// guard band, and any survivor whose fast r^2 lies within the band of rc^2
// is re-tested with util::pbcDist's exact operation order
// (utilities.cpp:79-104) against the strict r^2 < rc^2 of
// LJPotential.cpp:82 / AOPotential.cpp:80 / Calibration.cpp:60.
#pragma once

#include "psim_device.hpp"

namespace psim {

// Positions in the engine's slot-ordered SoA ([3][N]) with one radius.
struct SoAPos {
    const double *__restrict__ p;
    int N;
    double radius;
    __device__ double x(int j) const { return p[j]; }
    __device__ double y(int j) const { return p[N + j]; }
    __device__ double z(int j) const { return p[2 * N + j]; }
    __device__ double rad(int) const { return radius; }
};

// The reference's sortedParticles layout (x,y,z,r per sorted index,
// systemHandling.cpp:79-83).
struct AoS4Pos {
    const double *__restrict__ p;
    __device__ double x(int j) const { return p[4 * (size_t)j]; }
    __device__ double y(int j) const { return p[4 * (size_t)j + 1]; }
    __device__ double z(int j) const { return p[4 * (size_t)j + 2]; }
    __device__ double rad(int j) const { return p[4 * (size_t)j + 3]; }
};

// Engine cell table: start offsets, cell c = [cs[c], cs[c+1]).
struct StartTab {
    const int *__restrict__ cs;
    __device__ void range(int h, int &b, int &e) const
    {
        b = cs[h];
        e = cs[h + 1];
    }
};

// The reference's cellStartEnd: (start, end) pairs, -1 for empty
// (system.h:166, systemHandling.cpp:60-75).
struct PairTab {
    const int *__restrict__ se;
    __device__ void range(int h, int &b, int &e) const
    {
        b = se[2 * h];
        e = se[2 * h + 1];
        if (b < 0 || e < b) {
            b = 0;
            e = 0;
        }
    }
};

// Original particle index of a slot (for error reports).
struct Orig {
    const int *__restrict__ p;
    int stride;
    int off;
    __device__ int operator()(int j) const { return p[stride * j + off]; }
};

__device__ inline double shifted(double xi, int k, double L)
{
    return k == 0 ? xi : (k == 1 ? xi + L : xi - L);
}

// Drain one lane's hit list: pair forces for entries [0, nh).  The list is a
// column of stride BLOCK in the block's LDS array.
template <int MASK, bool EXACT, int BLOCK = kBlock, class Src>
__device__ void drain_hits(const Geo &g, const Pot &p, const Src &src, int s,
                           double xi, double yi, double zi, double ri,
                           const unsigned *hits, int nh, double &fx,
                           double &fy, double &fz, DevErr *err,
                           const Orig &orig)
{
    for (int k = 0; k < nh; ++k) {
        unsigned e = hits[k * BLOCK];
        int j = (int)(e & kIdxMask);
        double xj = src.x(j), yj = src.y(j), zj = src.z(j);
        double size = ri + src.rad(j);
        double dx, dy, dz, r2, rinv, r;
        if (!EXACT) {
            unsigned code = e >> kIdxBits;
            int kx = code % 3, ky = (code / 3) % 3, kz = code / 9;
            dx = xj - shifted(xi, kx, g.L);
            dy = yj - shifted(yi, ky, g.L);
            dz = zj - shifted(zi, kz, g.L);
            r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 >= p.rc2_lo) {
                double r2x = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
                if (!(r2x < p.rc2)) continue;
            }
            rinv = rsqrt_nr(r2);
            r = r2 * rinv;
        } else {
            // the reference's own arithmetic throughout (S == 3 boxes)
            r2 = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
            r = sqrt(r2);
            rinv = 1.0 / r;
            dx = min_image_uv(xj - xi, g.L, g.half_uv);
            dy = min_image_uv(yj - yi, g.L, g.half_uv);
            dz = min_image_uv(zj - zi, g.L, g.half_uv);
        }
        double ov = p.ov_factor * size;
        if (r2 < ov * ov * (1.0 + p.ov_band)) {
            double r2x = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
            if (sqrt(r2x) < ov) report_error(err, 7704, orig(s), orig(j));
        }
        bool nan_ao = false;
        double f = pair_fnet<MASK>(p, r2, rinv, r, size, &nan_ao);
        if (nan_ao) report_error(err, 7705, orig(s), orig(j));
        double fr = f * rinv;
        fx = fma(fr, dx, fx);
        fy = fma(fr, dy, fy);
        fz = fma(fr, dz, fz);
    }
}

// Candidates of one neighbour cell h (image code `code`) appended to the
// lane's hit list.
template <int MASK, bool EXACT, int BLOCK, class Src>
__device__ __forceinline__ void cull_cell(const Geo &g, const Pot &p, const Src &src, int s,
                                          double xi, double yi, double zi, double xs,
                                          double ys, double zs, int b, int e, unsigned code,
                                          unsigned *hits, int &nh, double ri, double &fx,
                                          double &fy, double &fz, DevErr *err,
                                          const Orig &orig)
{
    for (int j = b; j < e; ++j) {
        bool hit;
        if (!EXACT) {
            double dx = src.x(j) - xs;
            double dy = src.y(j) - ys;
            double dz = src.z(j) - zs;
            double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            hit = (r2 < p.rc2_hi) && (j != s);
        } else {
            double r2 = pbc_dist_ref(xi, yi, zi, src.x(j), src.y(j), src.z(j), g.L,
                                     g.half_pbc);
            hit = (r2 < p.rc2) && (j != s);
        }
        if (hit) {
            hits[nh * BLOCK] = (unsigned)j | code;
            ++nh;
        }
        if (__any(nh == kHitCap)) {
            drain_hits<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, ri, hits, nh, fx, fy, fz,
                                           err, orig);
            nh = 0;
        }
    }
}

// Force on the particle in slot s (position xi,yi,zi, radius ri).
// hits points at this lane's column of the block's LDS hit list.  With
// SPLIT > 1, SPLIT consecutive lanes share the particle and lane `sub` walks
// the neighbour cells sub, sub + SPLIT, ... of the 27 (its partial force;
// the caller sums the SPLIT partials).
template <int MASK, bool EXACT, int BLOCK = kBlock, int SPLIT = 1, class Src, class Tab>
__device__ void gather_force(const Geo &g, const Pot &p, const Src &src,
                             const Tab &tab, int s, double xi, double yi,
                             double zi, double ri, unsigned *hits, double &fx,
                             double &fy, double &fz, DevErr *err,
                             const Orig &orig, int sub = 0)
{
    const int S = g.S;
    int cx = cell_1d(xi, g.cell, S);
    int cy = cell_1d(yi, g.cell, S);
    int cz = cell_1d(zi, g.cell, S);
    int nh = 0;
    if constexpr (SPLIT > 1) {
        // few particles: more lanes per particle, so more loads in flight
        for (int q = sub; q < 27; q += SPLIT) {
            int rx = cx + q / 9 - 1, ry = cy + (q / 3) % 3 - 1, rz = cz + q % 3 - 1;
            int kx = 0, ky = 0, kz = 0;
            if (rx < 0) { rx += S; kx = 1; } else if (rx >= S) { rx -= S; kx = 2; }
            if (ry < 0) { ry += S; ky = 1; } else if (ry >= S) { ry -= S; ky = 2; }
            if (rz < 0) { rz += S; kz = 1; } else if (rz >= S) { rz -= S; kz = 2; }
            int b, e;
            tab.range(g.c0 + rx + S * ry + g.S2 * rz, b, e);
            cull_cell<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, shifted(xi, kx, g.L),
                                          shifted(yi, ky, g.L), shifted(zi, kz, g.L), b, e,
                                          (unsigned)(kx + 3 * ky + 9 * kz) << kIdxBits, hits,
                                          nh, ri, fx, fy, fz, err, orig);
        }
        drain_hits<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, ri, hits, nh, fx, fy, fz,
                                       err, orig);
        return;
    }
    for (int ox = -1; ox <= 1; ++ox) {
        int rx = cx + ox, kx = 0;
        if (rx < 0) { rx += S; kx = 1; } else if (rx >= S) { rx -= S; kx = 2; }
        double xs = shifted(xi, kx, g.L);
        for (int oy = -1; oy <= 1; ++oy) {
            int ry = cy + oy, ky = 0;
            if (ry < 0) { ry += S; ky = 1; } else if (ry >= S) { ry -= S; ky = 2; }
            double ys = shifted(yi, ky, g.L);
            for (int oz = -1; oz <= 1; ++oz) {
                int rz = cz + oz, kz = 0;
                if (rz < 0) { rz += S; kz = 1; } else if (rz >= S) { rz -= S; kz = 2; }
                double zs = shifted(zi, kz, g.L);
                int h = g.c0 + rx + S * ry + g.S2 * rz;
                int b, e;
                tab.range(h, b, e);
                unsigned code = (unsigned)(kx + 3 * ky + 9 * kz) << kIdxBits;
                for (int j = b; j < e; ++j) {
                    bool hit;
                    if (!EXACT) {
                        double dx = src.x(j) - xs;
                        double dy = src.y(j) - ys;
                        double dz = src.z(j) - zs;
                        double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                        hit = (r2 < p.rc2_hi) && (j != s);
                    } else {
                        double r2 = pbc_dist_ref(xi, yi, zi, src.x(j), src.y(j),
                                                 src.z(j), g.L, g.half_pbc);
                        hit = (r2 < p.rc2) && (j != s);
                    }
                    if (hit) {
                        hits[nh * BLOCK] = (unsigned)j | code;
                        ++nh;
                    }
                    if (__any(nh == kHitCap)) {
                        drain_hits<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, ri, hits,
                                                       nh, fx, fy, fz, err, orig);
                        nh = 0;
                    }
                }
            }
        }
    }
    drain_hits<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, ri, hits, nh, fx, fy, fz,
                                   err, orig);
}

}  // namespace psim
