// psim_brick.hpp -- brick-per-workgroup pair kernel for sparse systems.
//
// Below two particles per cell, one lane per particle walking its 27 cells in
// global memory spends the step waiting: every cell costs a dependent
// cell-table load and a short candidate loop, and a wave runs each loop as
// long as its fullest lane.  Here one workgroup (kBrickThreads lanes) owns a
// brick of bx x by x bz home cells (BrickGeo).  It reads the start offsets
// of the brick's (bx+2)(by+2)(bz+2) halo cells once, scans them into a
// staged-offset table, and copies the halo's positions into LDS in halo
// order (x fastest), each load independent of the others.  Then G
// consecutive lanes take one home particle and run the thread-per-particle
// gather against LDS: lane gg walks the particle's x-runs gg, gg + G, ...
// (the 27 neighbour cells of LJPotential.cpp:154-177 in gather_force's
// order: (dy, dz) y outer, then dx = -1, 0, +1), candidates in slot order, with
// k_step's arithmetic (shifted home position per neighbour cell, guard-banded
// r^2, util::pbcDist's recheck, pair_fr), survivors through a per-lane LDS
// hit list, and the G partials summed by the same xor butterfly.  A brick
// step therefore equals a k_step step with SPLIT = G bit for bit; only where
// the operands come from changes.
//
// A halo holding more than kBrickStageCap particles (clusters) is processed
// in staged chunks, which changes only the order of the partial sums.
#pragma once

#include "psim_brick_shape.hpp"
#include "psim_pairs.hpp"

namespace psim {

constexpr int kBrickThreads = 256;
#ifndef PSIM_BRICK_HITCAP
#define PSIM_BRICK_HITCAP 8
#endif
constexpr int kBrickHitCap = PSIM_BRICK_HITCAP;   // per-lane hit-list entries
static_assert(kBrickHaloCap <= 4 * kBrickThreads, "halo table: up to 4 cells per lane");
static_assert(kBrickStageCap <= 1024, "a hit entry holds a 10-bit staged index");

struct BrickShared {
    double sx[kBrickStageCap], sy[kBrickStageCap], sz[kBrickStageCap];   // unshifted
    int sslot[kBrickStageCap];
    int hoff[kBrickHaloCap + 1];   // staged offset of each halo cell (whole halo)
    int hstart[kBrickHaloCap];     // slot of each halo cell's first particle
    int rowoff[kBrickRowCap + 1];  // home particles before each home row (y, z)
    int rowslot[kBrickRowCap];     // slot of each home row's first particle
    unsigned short hits[kBrickHitCap * kBrickThreads];   // staged index | image code << 10
    int wsum[kBrickThreads / 64];
};

// Exclusive scan of one value per lane over the workgroup; *total gets the
// sum.  Ends with every lane past a barrier that follows the wsum writes.
__device__ inline int brick_scan(int v, int *total, int *wsum)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kBrickThreads / 64; ++q) {
        const int s = wsum[q];
        before += q < w ? s : 0;
        tot += s;
    }
    *total = tot;
    return before + x - v;
}

// Pair forces of one lane's hit list, entries [0, nh) (drain_hits' arithmetic
// on staged operands).
template <int MASK, bool EXACT, class ST>
__device__ __forceinline__ void brick_drain(const Geo &g, const Pot &p, const BrickShared &sh,
                            const int *__restrict__ id, double xi, double yi, double zi,
                            double size, int slot_i, int nh, double &fx, double &fy,
                            double &fz, DevErr *err, ST &st)
{
    const int lane = threadIdx.x;
    for (int k = 0; k < nh; ++k) {
        const unsigned e = sh.hits[k * kBrickThreads + lane];
        const int j = (int)(e & 1023u);
        const double xj = sh.sx[j], yj = sh.sy[j], zj = sh.sz[j];
        double dx, dy, dz, r2, rinv, r;
        if (!EXACT) {
            const unsigned code = e >> 10;
            const int kx = code % 3, ky = (code / 3) % 3, kz = code / 9;
            dx = xj - shifted(xi, kx, g.L);
            dy = yj - shifted(yi, ky, g.L);
            dz = zj - shifted(zi, kz, g.L);
            r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            if (r2 >= p.rc2_lo) {
                const double r2x = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
                if (!(r2x < p.rc2)) continue;
            }
            rinv = rsqrt_nr(r2);
            r = r2 * rinv;
        } else {
            r2 = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
            r = sqrt(r2);
            rinv = 1.0 / r;
            dx = min_image_uv(xj - xi, g.L, g.half_uv);
            dy = min_image_uv(yj - yi, g.L, g.half_uv);
            dz = min_image_uv(zj - zi, g.L, g.half_uv);
        }
        if constexpr (ST::kOn) {
            st.add(id[sh.sslot[j]]);
            continue;
        }
        const double ov = p.ov_factor * size;
        if (r2 < ov * ov * (1.0 + p.ov_band)) {
            const double r2x = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
            if (sqrt(r2x) < ov) report_error(err, 7704, id[slot_i], id[sh.sslot[j]]);
        }
        bool nan_ao = false;
        const double fr = pair_fr<MASK>(p, rinv, r, size, &nan_ao);
        if (nan_ao) report_error(err, 7705, id[slot_i], id[sh.sslot[j]]);
        fx = fma(fr, dx, fx);
        fy = fma(fr, dy, fy);
        fz = fma(fr, dz, fz);
    }
}

// Forces on the home particles of brick b.  epi(slot, fx, fy, fz) is called
// once per home particle, by the first lane of its group.  With ST =
// NbrStats the accepted pairs are counted instead and written to *so by
// original index (epi is not called).
template <int MASK, bool EXACT, int G, class ST = NoStats, class Epi>
__device__ __forceinline__ void brick_forces(const Geo &g, const Pot &p, double radius,
                             const double *__restrict__ pos, const int *__restrict__ cstart,
                             const int *__restrict__ id, DevErr *err, const BrickGeo &bg,
                             int b, BrickShared &sh, Epi epi,
                             const NbrStatsOut *so = nullptr)
{
    static_assert(G >= 1 && 64 % G == 0 && (G & (G - 1)) == 0, "lane groups");
    constexpr int T = kBrickThreads;
    const int N = g.n, S = g.S;
    const int lane = threadIdx.x;
    const int ix = b % bg.nbx, iy = (b / bg.nbx) % bg.nby, iz = b / (bg.nbx * bg.nby);
    const int x0 = ix * bg.bx, y0 = iy * bg.by, z0 = iz * bg.bz;
    const int ex = min(bg.bx, S - x0), ey = min(bg.by, S - y0), ez = min(bg.bz, S - z0);
    const int HX = ex + 2, HY = ey + 2;
    const int nhalo = HX * HY * (ez + 2);
    // ---- halo table: lane owns halo cells 4 lane .. 4 lane + 3
    int cnt[4];
    int sum = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = 4 * lane + t;
        cnt[t] = 0;
        if (k < nhalo) {
            int rx = x0 - 1 + k % HX, ry = y0 - 1 + (k / HX) % HY, rz = z0 - 1 + k / (HX * HY);
            rx += rx < 0 ? S : (rx >= S ? -S : 0);
            ry += ry < 0 ? S : (ry >= S ? -S : 0);
            rz += rz < 0 ? S : (rz >= S ? -S : 0);
            const int h = g.c0 + rx + S * ry + g.S2 * rz;
            const int hb = cstart[h];
            cnt[t] = cstart[h + 1] - hb;
            sh.hstart[k] = hb;
            sum += cnt[t];
        }
    }
    int M;
    int pre = brick_scan(sum, &M, sh.wsum);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = 4 * lane + t;
        if (k < nhalo) sh.hoff[k] = pre;
        pre += cnt[t];
    }
    if (lane == 0) sh.hoff[nhalo] = M;
    __syncthreads();
    // ---- home rows: halo cells (1..ex, v, w) hold consecutive slots
    const int nrows = ey * ez;
    if (lane < 64) {
        int carry = 0;
        for (int r0 = 0; r0 < nrows; r0 += 64) {
            const int r = r0 + lane;
            int c = 0;
            if (r < nrows) {
                const int k1 = 1 + HX * ((r % ey + 1) + HY * (r / ey + 1));
                c = sh.hoff[k1 + ex] - sh.hoff[k1];
                sh.rowslot[r] = sh.hstart[k1];
            }
            int x = c;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (r < nrows) sh.rowoff[r] = carry + x - c;
            carry += __shfl(x, 63, 64);
        }
        if (lane == 0) sh.rowoff[nrows] = carry;
    }
    __syncthreads();
    const int nhome = sh.rowoff[nrows];
    constexpr int P = T / G;   // home particles per pass
    const int il = lane / G, gg = lane % G;
    const double size = radius + radius;
    int resident = -1;         // staged chunk (workgroup-uniform)
    for (int i0 = 0; i0 < nhome; i0 += P) {
        const int i = i0 + il;
        const bool act = i < nhome;
        int slot = 0, u = 0, v = 0, w = 0;
        double xi = 0.0, yi = 0.0, zi = 0.0;
        if (act) {
            int lo = 0, hi = nrows - 1;   // the row holding home particle i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sh.rowoff[mid] <= i) lo = mid; else hi = mid - 1;
            }
            slot = sh.rowslot[lo] + (i - sh.rowoff[lo]);
            xi = pos[slot];
            yi = pos[N + slot];
            zi = pos[2 * N + slot];
            // the cell the sort put it in (clamped: LDS indices stay in range)
            u = min(max(cell_1d(xi, g.cell, S) - x0 + 1, 1), ex);
            v = lo % ey + 1;
            w = lo / ey + 1;
        }
        double fx = 0.0, fy = 0.0, fz = 0.0;
        ST st;
        for (int s0 = 0; s0 < M; s0 += kBrickStageCap) {
            const int mc = min(kBrickStageCap, M - s0);
            if (s0 != resident) {
                __syncthreads();   // the previous chunk is no longer read
                for (int t = lane; t < mc; t += T) {
                    const int q = s0 + t;
                    int lo = 0, hi = nhalo - 1;   // the halo cell holding staged index q
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (sh.hoff[mid] <= q) lo = mid; else hi = mid - 1;
                    }
                    const int sl = sh.hstart[lo] + (q - sh.hoff[lo]);
                    sh.sx[t] = pos[sl];
                    sh.sy[t] = pos[N + sl];
                    sh.sz[t] = pos[2 * N + sl];
                    sh.sslot[t] = sl;
                }
                __syncthreads();
                resident = s0;
            }
            if (act) {
                int nh = 0;
                const int cx = x0 + u - 1;
                const bool lw = cx == 0, rw = cx == S - 1;
                for (int q = gg; q < 9; q += G) {
                    const int dv = q / 3 - 1, dw = q % 3 - 1;
                    const int ry = y0 + v - 1 + dv, rz = z0 + w - 1 + dw;
                    const int ky = ry < 0 ? 1 : (ry >= S ? 2 : 0);
                    const int kz = rz < 0 ? 1 : (rz >= S ? 2 : 0);
                    const double ys = shifted(yi, ky, g.L), zs = shifted(zi, kz, g.L);
                    // halo cells u-1, u, u+1 of this row are consecutive in the
                    // stage; split where the row wraps, as x_run does
                    const int krow = HX * ((v + dv) + HY * (w + dw));
                    const int a_hi = lw ? u - 1 : (rw ? u : u + 1);
#pragma unroll
                    for (int sg = 0; sg < 2; ++sg) {
                        if (sg == 1 && !(lw || rw)) break;
                        const int c_lo = sg ? a_hi + 1 : u - 1;
                        const int c_hi = sg ? u + 1 : a_hi;
                        const int kx = sg ? (lw ? 0 : 2) : (lw ? 1 : 0);
                        const double xs = shifted(xi, kx, g.L);
                        const unsigned code = (unsigned)(kx + 3 * ky + 9 * kz) << 10;
                        const int jb = max(sh.hoff[krow + c_lo], s0) - s0;
                        const int je = min(sh.hoff[krow + c_hi + 1], s0 + mc) - s0;
                        for (int j = jb; j < je; ++j) {
                            bool hit;
                            if (!EXACT) {
                                const double dx = sh.sx[j] - xs;
                                const double dy = sh.sy[j] - ys;
                                const double dz = sh.sz[j] - zs;
                                const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                                hit = (r2 < p.rc2_hi) && (sh.sslot[j] != slot);
                            } else {
                                const double r2 = pbc_dist_ref(xi, yi, zi, sh.sx[j], sh.sy[j],
                                                               sh.sz[j], g.L, g.half_pbc);
                                hit = (r2 < p.rc2) && (sh.sslot[j] != slot);
                            }
                            if (hit) {
                                sh.hits[nh * T + lane] = (unsigned short)((unsigned)j | code);
                                ++nh;
                            }
                            if (__any(nh == kBrickHitCap)) {
                                brick_drain<MASK, EXACT>(g, p, sh, id, xi, yi, zi, size, slot,
                                                         nh, fx, fy, fz, err, st);
                                nh = 0;
                            }
                        }
                    }
                }
                brick_drain<MASK, EXACT>(g, p, sh, id, xi, yi, zi, size, slot, nh, fx, fy, fz,
                                         err, st);
            }
        }
        if constexpr (ST::kOn) {
            split_stats<G>(st);
            if (act && gg == 0) {
                so->count[id[slot]] = st.count;
                so->hash[id[slot]] = st.hash;
            }
            continue;
        }
        fx = split_sum<G>(fx);
        fy = split_sum<G>(fy);
        fz = split_sum<G>(fz);
        if (act && gg == 0) epi(slot, fx, fy, fz);
    }
}

}  // namespace psim
