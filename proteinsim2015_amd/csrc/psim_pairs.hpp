// psim_pairs.hpp -- the 27-cell pair gather shared by every force path.
//
// One lane per target particle i.  The lane walks the 27 neighbour cells of
// the reference's stencil (LJPotential.cpp:154-177) as 9 runs along x: for
// each (dy, dz), y outer and z inner, the cells x-1, x, x+1 of that row hold
// consecutive slots, so the run is one slot range (two where the row wraps
// the periodic boundary, each with its own image shift), candidates in slot
// order.  It culls them with a cheap FP64 distance on the periodic image of
// the neighbour cell (shift applied once per segment).  Survivors go to a per-lane
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

// x-runs whose slot ranges are read together (1: one run at a time)
#ifndef PSIM_RUN_GROUP
#define PSIM_RUN_GROUP 3
#endif
constexpr int kRunGroup = PSIM_RUN_GROUP;

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
    // slots of the consecutive cells h_lo..h_hi of one row
    __device__ void row_range(int h_lo, int h_hi, int &b, int &e) const
    {
        b = cs[h_lo];
        e = cs[h_hi + 1];
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
    // cells are stored in cell order, so consecutive non-empty cells are
    // adjacent: the row's slots run from the first start to the last end
    __device__ void row_range(int h_lo, int h_hi, int &b, int &e) const
    {
        b = 0;
        e = 0;
        for (int h = h_lo; h <= h_hi; ++h) {
            int hb, he;
            range(h, hb, he);
            if (he > hb) {
                if (e == 0) b = hb;
                e = he;
            }
        }
    }
};

// The neighbour cells x-1, x, x+1 (cell x = cx) of row `row` as at most two
// slot segments in stencil order dx = -1, 0, +1: a wrapped end cell is a
// segment of its own with image code 1 (dx = -1 at cx = 0: image +L) or 2
// (dx = +1 at cx = S-1: image -L).  Segment 1 is empty unless the row wraps.
struct XRun {
    int b0, e0, k0, b1, e1, k1;
};

template <class Tab>
__device__ __forceinline__ XRun x_run(const Tab &tab, int row, int cx, int S)
{
    const bool lw = cx == 0, rw = cx == S - 1;
    XRun r;
    tab.row_range(row + (lw ? S - 1 : cx - 1), row + (lw || rw ? S - 1 : cx + 1), r.b0, r.e0);
    r.k0 = lw ? 1 : 0;
    if (lw || rw) {
        tab.row_range(row, row + (lw ? 1 : 0), r.b1, r.e1);
    } else {
        r.b1 = 0;   // an interior row is one segment (no table reads for the second)
        r.e1 = 0;
    }
    r.k1 = lw ? 0 : 2;
    return r;
}

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
// column of stride BLOCK in the block's LDS array.  Entries are taken two at
// a time: both partners' positions are loaded before either pair is
// evaluated (the evaluation order stays the list order).
template <int MASK, bool EXACT, int BLOCK = kBlock, class Src, class ST = NoStats>
__device__ void drain_hits(const Geo &g, const Pot &p, const Src &src, int s,
                           double xi, double yi, double zi, double ri,
                           const unsigned *hits, int nh, double &fx,
                           double &fy, double &fz, DevErr *err,
                           const Orig &orig, ST *st = nullptr)
{
    auto pair = [&](unsigned e, double xj, double yj, double zj) {
        const int j = (int)(e & kIdxMask);
        const double size = ri + src.rad(j);
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
                if (!(r2x < p.rc2)) return;
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
        if constexpr (ST::kOn) {
            st->add(orig(j));
            return;
        }
        double ov = p.ov_factor * size;
        if (r2 < ov * ov * (1.0 + p.ov_band)) {
            double r2x = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
            if (sqrt(r2x) < ov) report_error(err, 7704, orig(s), orig(j));
        }
        bool nan_ao = false;
        const double fr = pair_fr<MASK>(p, rinv, r, size, &nan_ao);
        if (nan_ao) report_error(err, 7705, orig(s), orig(j));
        fx = fma(fr, dx, fx);
        fy = fma(fr, dy, fy);
        fz = fma(fr, dz, fz);
    };
    // (one at a time with LJ and AO both enabled: the pair arithmetic then
    // leaves no registers for a second partner)
    constexpr bool kTwo = !((MASK & F_LJ) && (MASK & F_AO));
    if constexpr (kTwo) {
        for (int k = 0; k < nh; k += 2) {
            const bool two = k + 1 < nh;
            const unsigned e0 = hits[k * BLOCK];
            const unsigned e1 = two ? hits[(k + 1) * BLOCK] : e0;
            const int j0 = (int)(e0 & kIdxMask), j1 = (int)(e1 & kIdxMask);
            const double x0 = src.x(j0), y0 = src.y(j0), z0 = src.z(j0);
            const double x1 = src.x(j1), y1 = src.y(j1), z1 = src.z(j1);
            pair(e0, x0, y0, z0);
            if (two) pair(e1, x1, y1, z1);
        }
    } else {
        for (int k = 0; k < nh; ++k) {
            const unsigned e = hits[k * BLOCK];
            const int j = (int)(e & kIdxMask);
            pair(e, src.x(j), src.y(j), src.z(j));
        }
    }
}

// Candidates of one neighbour cell h (image code `code`) appended to the
// lane's hit list.
template <int MASK, bool EXACT, int BLOCK, class Src, class ST>
__device__ __forceinline__ void cull_cell(const Geo &g, const Pot &p, const Src &src, int s,
                                          double xi, double yi, double zi, double xs,
                                          double ys, double zs, int b, int e, unsigned code,
                                          unsigned *hits, int &nh, double ri, double &fx,
                                          double &fy, double &fz, DevErr *err,
                                          const Orig &orig, ST *st)
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
                                           err, orig, st);
            nh = 0;
        }
    }
}

// Force on the particle in slot s (position xi,yi,zi, radius ri).
// hits points at this lane's column of the block's LDS hit list.  With
// SPLIT > 1, SPLIT consecutive lanes share the particle and lane `sub` walks
// the x-runs sub, sub + SPLIT, ... of the 9 (its partial force; the caller
// sums the SPLIT partials).  With st (NbrStats), the accepted pairs are
// recorded there instead of evaluated.
template <int MASK, bool EXACT, int BLOCK = kBlock, int SPLIT = 1, class Src, class Tab,
          class ST = NoStats>
__device__ void gather_force(const Geo &g, const Pot &p, const Src &src,
                             const Tab &tab, int s, double xi, double yi,
                             double zi, double ri, unsigned *hits, double &fx,
                             double &fy, double &fz, DevErr *err,
                             const Orig &orig, int sub = 0, ST *st = nullptr)
{
    const int S = g.S;
    const int cx = cell_1d(xi, g.cell, S);
    const int cy = cell_1d(yi, g.cell, S);
    const int cz = cell_1d(zi, g.cell, S);
    int nh = 0;
    // SPLIT > 1 (few particles): lane `sub` takes runs sub, sub + SPLIT, ...
    // The slot ranges of up to three runs are read before the first of them
    // is culled (their table reads are in flight together).
    for (int q0 = sub; q0 < 9; q0 += kRunGroup * SPLIT) {
        XRun xr[kRunGroup];
        int kyz[kRunGroup];
#pragma unroll
        for (int u = 0; u < kRunGroup; ++u) {
            const int q = q0 + u * SPLIT;
            xr[u] = XRun{0, 0, 0, 0, 0, 0};
            kyz[u] = 0;
            if (q < 9) {
                int ry = cy + q / 3 - 1, rz = cz + q % 3 - 1;
                int ky = 0, kz = 0;
                if (ry < 0) { ry += S; ky = 1; } else if (ry >= S) { ry -= S; ky = 2; }
                if (rz < 0) { rz += S; kz = 1; } else if (rz >= S) { rz -= S; kz = 2; }
                xr[u] = x_run(tab, g.c0 + S * ry + g.S2 * rz, cx, S);
                kyz[u] = ky + 3 * kz;
            }
        }
#pragma unroll
        for (int u = 0; u < kRunGroup; ++u) {
            const int ky = kyz[u] % 3, kz = kyz[u] / 3;
            const double ys = shifted(yi, ky, g.L), zs = shifted(zi, kz, g.L);
#pragma unroll
            for (int sg = 0; sg < 2; ++sg) {
                const int kx = sg ? xr[u].k1 : xr[u].k0;
                cull_cell<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, shifted(xi, kx, g.L), ys,
                                              zs, sg ? xr[u].b1 : xr[u].b0,
                                              sg ? xr[u].e1 : xr[u].e0,
                                              (unsigned)(kx + 3 * ky + 9 * kz) << kIdxBits,
                                              hits, nh, ri, fx, fy, fz, err, orig, st);
            }
        }
    }
    drain_hits<MASK, EXACT, BLOCK>(g, p, src, s, xi, yi, zi, ri, hits, nh, fx, fy, fz, err,
                                   orig, st);
}

}  // namespace psim
