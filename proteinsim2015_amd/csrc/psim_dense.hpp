// psim_dense.hpp -- cell-per-workgroup pair kernel for dense systems.
//
// One workgroup (kDenseThreads lanes) owns one home cell.  It stages the 27
// neighbour cells' positions in LDS, in the reference's stencil order
// (x outer, z inner; LJPotential.cpp:154-177), with the periodic image shift
// of each neighbour cell applied once while staging.  The n_i home particles
// are spread over G = kDenseThreads / n_i lane groups; lane (i, g) culls
// candidates g, g+G, ... of the staged list in batches of up to 32, one bit
// per candidate in a register mask, then evaluates the pair forces of the
// batch's survivors (last candidate first).  Group partials are summed through
// LDS, then lanes 0..n_i-1 hand the home particles' forces to the epilogue.
//
// Neighbour sets are bit-exact with the reference: fast r^2 on the shifted
// image with a guard band, and util::pbcDist's exact arithmetic on the raw
// positions (utilities.cpp:79-104) for any survivor inside the band.
// Requires S >= 4 (the shifted image is then the minimum image).
#pragma once

#include "psim_device.hpp"

namespace psim {

// Tile shape (overridable at build time for launch-shape sweeps).
#ifndef PSIM_DENSE_THREADS
#define PSIM_DENSE_THREADS 256
#endif
#ifndef PSIM_STAGE_CAP
#define PSIM_STAGE_CAP 512
#endif
// Waves per SIMD the register allocation is held to: 7 leaves 72 VGPRs per
// lane, which the kernel needs for one module (or Calibration beside one),
// against 5 waves at the compiler's own choice of about 86; with both LJ and
// AO enabled it needs 80 (6 waves) to stay out of scratch memory.
#ifndef PSIM_DENSE_WAVES
#define PSIM_DENSE_WAVES 7
#endif
constexpr int kDenseThreads = PSIM_DENSE_THREADS;  // lanes per home cell (4 wave64s)
constexpr int kStageCap = PSIM_STAGE_CAP;          // staged candidates per chunk
static_assert(kDenseThreads >= 64 && kDenseThreads % 64 == 0, "whole wave64s per cell");

template <int T>
struct DenseSharedT {
    double sx[kStageCap], sy[kStageCap], sz[kStageCap];
    union {
        // the cull's copy: FP32, relative to the home cell's corner
        struct {
            float x[kStageCap], y[kStageCap], z[kStageCap];
        } cf;
        double red[3][T];   // per-lane partial forces (after the last drain)
    } u;
    int sslot[kStageCap];
    double etab[64];   // copy of kExp2Tab
    int coff[28];
    int cbase[27];
    double shx[27], shy[27], shz[27];
};
using DenseShared = DenseSharedT<kDenseThreads>;
// waves per SIMD for the kernels of force mask `mask` and compile-time LJ
// exponent `ljn` (0: ljNum read at run time): LJ with AO, or LJ with a
// run-time exponent, need more than 72 VGPRs (6 waves), and LJ with AO and a
// run-time exponent more than 80 (5 waves)
constexpr int dense_waves(int mask, int ljn = 0)
{
    const bool lj_ao = (mask & F_LJ) && (mask & F_AO);
    const int w = (lj_ao && ljn == 0) ? 5 : (lj_ao || ((mask & F_LJ) && ljn == 0)) ? 6 : 7;
    return PSIM_DENSE_WAVES < w ? PSIM_DENSE_WAVES : w;
}
// the workgroups that PSIM_DENSE_WAVES waves per SIMD make resident on one CU
// (4 SIMDs) must also fit its 160 KiB of LDS
static_assert(PSIM_DENSE_WAVES * 4 * 64 / kDenseThreads * sizeof(DenseShared) <= 160 * 1024,
              "LDS for the resident dense workgroups");

// Map a workgroup index to a cell so that each XCD (workgroups are handed
// out round-robin over 8 XCDs) walks one contiguous slab of cells and its
// private L2 keeps the slab's positions.
__device__ inline int xcd_cell(int b, int per_xcd)
{
    return (b & 7) * per_xcd + (b >> 3);
}

// global double at a 32-bit byte offset from base
__device__ __forceinline__ double ld_off(const double *__restrict__ base, unsigned off)
{
    return *(const double *)((const char *)base + off);
}

// element of a staged array at byte offset off (the LDS address is then the
// offset plus a constant, with no index scaling)
__device__ __forceinline__ double lds_at(const double *a, int off)
{
    return *(const double *)((const char *)a + off);
}

// Pair force of staged candidate j on the lane's home particle (FP64; the
// exact pbcDist recheck for survivors inside the guard band).
// j8 = 8 j, the byte offset of candidate j in the staged double arrays.
template <int MASK, int T, int LJN, class ST>
__device__ __forceinline__ void dense_pair(const Geo &g, const Pot &p,
                                           const DenseSharedT<T> &sh,
                                           const double *__restrict__ pos, double xi, double yi,
                                           double zi, double size, double ov2_hi, int j8,
                                           double &fx, double &fy, double &fz, DevErr *err,
                                           int id_i, const int *__restrict__ id, ST &st)
{
    const int N = g.n;
    const double dx = lds_at(sh.sx, j8) - xi, dy = lds_at(sh.sy, j8) - yi,
                 dz = lds_at(sh.sz, j8) - zi;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const int j = j8 >> 3;
    if (r2 >= p.rc2_lo || r2 < ov2_hi) {
        const int sj = sh.sslot[j];
        const double r2x = pbc_dist_ref(xi, yi, zi, pos[sj], pos[N + sj], pos[2 * N + sj],
                                        g.L, g.half_pbc);
        if (r2 >= p.rc2_lo && !(r2x < p.rc2)) return;
        if (!ST::kOn && sqrt(r2x) < p.ov_factor * size) report_error(err, 7704, id_i, id[sj]);
    }
    if constexpr (ST::kOn) {
        st.add(id[sh.sslot[j]]);
        return;
    }
    const double rinv = rsqrt_nr(r2);
    const double r = r2 * rinv;
    bool nan_ao = false;
    const double fr = pair_fr<MASK, true, LJN>(p, rinv, r, size, &nan_ao, sh.etab);
    if (nan_ao) report_error(err, 7705, id_i, id[sh.sslot[j]]);
    fx = fma(fr, dx, fx);
    fy = fma(fr, dy, fy);
    fz = fma(fr, dz, fz);
}

// a + b * c for |b|, |c| < 2^23 with c uniform over the wave (v_mad_i32_i24:
// one full-rate instruction; the compiler otherwise folds the index
// arithmetic into a 64-bit mad)
__device__ __forceinline__ int mad_i24(int b, int c_uniform, int a)
{
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(b), "s"(c_uniform), "v"(a));
    return r;
}

// Survivors of one batch of n <= 32 candidates j0, j0 + G, ...: bit
// (n - 1 - k) of m is set when candidate j0 + k G passed the cull.  The
// lowest set bit is taken first (m &= m - 1), so the batch is evaluated
// from its last candidate down: bit c is candidate j0 + (n - 1 - c) G, at
// byte offset 8 (j0 + (n - 1 - c) G) in the staged arrays.
template <int MASK, int T, int LJN, class ST>
__device__ void dense_drain_mask(const Geo &g, const Pot &p, const DenseSharedT<T> &sh,
                                 const double *__restrict__ pos, double xi, double yi,
                                 double zi, double size, double ov2_hi, int self, unsigned m,
                                 int n, int j0, int G, double &fx, double &fy, double &fz,
                                 DevErr *err, int id_i, const int *__restrict__ id, ST &st)
{
    const int jtop8 = 8 * (j0 + (n - 1) * G), self8 = 8 * self;
    while (m) {
        const int c = __builtin_ctz(m);
        m &= m - 1u;
        const int j8 = mad_i24(c, -8 * G, jtop8);
        if (j8 == self8) continue;
        dense_pair<MASK, T, LJN>(g, p, sh, pos, xi, yi, zi, size, ov2_hi, j8, fx, fy, fz, err,
                                 id_i, id, st);
    }
}

// Forces on the home particles of cell c.  On return, lanes 0..ni-1 hold the
// total force on home particle (i0 + lane) for each i-chunk via the callback
// epi(i, fx, fy, fz) (called by those lanes only).  With ST = NbrStats the
// accepted pairs are counted instead and written to *so by original index
// (epi is not called).
template <int MASK, int T, int LJN, class ST = NoStats, class Epi>
__device__ void dense_cell_forces(const Geo &g, const Pot &p, double radius,
                                  const double *__restrict__ pos,
                                  const int *__restrict__ cstart,
                                  const int *__restrict__ id, DevErr *err, int c,
                                  DenseSharedT<T> &sh, Epi epi,
                                  const NbrStatsOut *so = nullptr)
{
    const int N = g.n, S = g.S;
    const int lane = threadIdx.x;
    const int cl = c - g.c0;   // cell index within this system
    const int cx = cl % S, cy = (cl / S) % S, cz = cl / g.S2;
    const int home0 = cstart[c];
    const int ni = cstart[c + 1] - home0;
    if (ni == 0) return;   // uniform over the workgroup
    // origin of the cull's FP32 coordinates: the home cell's corner
    const double ox = cx * g.cell, oy = cy * g.cell, oz = cz * g.cell;
    // ---- neighbour-cell metadata (lanes 0..26, stencil order)
    int cnt = 0;
    if (lane < 27) {
        const int dxc = lane / 9 - 1, dyc = (lane / 3) % 3 - 1, dzc = lane % 3 - 1;
        int rx = cx + dxc, ry = cy + dyc, rz = cz + dzc;
        double sxv = 0.0, syv = 0.0, szv = 0.0;
        if (rx < 0) { rx += S; sxv = -g.L; } else if (rx >= S) { rx -= S; sxv = g.L; }
        if (ry < 0) { ry += S; syv = -g.L; } else if (ry >= S) { ry -= S; syv = g.L; }
        if (rz < 0) { rz += S; szv = -g.L; } else if (rz >= S) { rz -= S; szv = g.L; }
        const int h = g.c0 + rx + S * ry + g.S2 * rz;
        const int b = cstart[h];
        cnt = cstart[h + 1] - b;
        sh.cbase[lane] = b;
        sh.shx[lane] = sxv;
        sh.shy[lane] = syv;
        sh.shz[lane] = szv;
    }
    if (lane < 64) {
        sh.etab[lane] = kExp2Tab[lane];
        int incl = cnt;
#pragma unroll
        for (int d = 1; d < 32; d <<= 1) {
            int y = __shfl_up(incl, d, 64);
            if (lane >= d) incl += y;
        }
        if (lane < 27) sh.coff[lane] = incl - cnt;
        if (lane == 26) sh.coff[27] = incl;
    }
    __syncthreads();
    const int M = sh.coff[27];
    const int hb = sh.coff[13];   // staged index of the first home particle
    const double size = 2.0 * radius;
    const double ov = p.ov_factor * size;
    const double ov2_hi = ov * ov * (1.0 + p.ov_band);
    // FP32 cull: coordinates below 2 cells in magnitude carry errors below
    // 2.5e-7 cell each, so r^2 is within 2e-6 cell^2 of the FP64 value, and
    // r^2 - rc^2 evaluated as three FMAs adds at most 1.5 ulp of 12 cell^2
    // (2e-6 cell^2); the threshold is widened by 2e-5 cell^2
    const float nrc2_f = -(float)(p.rc2_hi + 2e-5 * g.cell * g.cell);

    for (int i0 = 0; i0 < ni; i0 += T) {
        const int nic = min(T, ni - i0);
        const int G = T / nic;
        const int gi = lane % nic, gg = lane / nic;
        const bool act = gg < G;
        const int si = home0 + i0 + gi;       // global slot of this lane's i
        const int hi = hb + i0 + gi;          // staged index of i
        const double xi = pos[si], yi = pos[N + si], zi = pos[2 * N + si];
        const float xf = (float)(xi - ox), yf = (float)(yi - oy), zf = (float)(zi - oz);
        const int id_i = id[si];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        ST st;
        for (int s0 = 0; s0 < M; s0 += kStageCap) {
            const int mc = min(kStageCap, M - s0);
            __syncthreads();   // the previous chunk is no longer read
            // kPer lanes per neighbour cell copy its particles that fall in
            // this chunk (no search for the cell of a staged index)
            constexpr int kPer = T / 27;
            if (lane < 27 * kPer) {
                const int q = lane / kPer, u = lane % kPer;
                const int t_lo = max(sh.coff[q], s0), t_hi = min(sh.coff[q + 1], s0 + mc);
                const int base = sh.cbase[q] - sh.coff[q];
                const double ex = sh.shx[q], ey = sh.shy[q], ez = sh.shz[q];
                for (int t = t_lo + u; t < t_hi; t += kPer) {
                    const int slot = base + t, k = t - s0;
                    // 32-bit byte offsets (3 N doubles stay below 4 GiB)
                    const unsigned o = 8u * (unsigned)slot, o3 = 8u * (unsigned)N;
                    const double x = ld_off(pos, o) + ex, y = ld_off(pos, o + o3) + ey,
                                 z = ld_off(pos, o + 2u * o3) + ez;
                    sh.sx[k] = x;
                    sh.sy[k] = y;
                    sh.sz[k] = z;
                    sh.u.cf.x[k] = (float)(x - ox);
                    sh.u.cf.y[k] = (float)(y - oy);
                    sh.u.cf.z[k] = (float)(z - oz);
                    sh.sslot[k] = slot;
                }
            }
            __syncthreads();
            if (act) {
                // self is accepted by the cull (r^2 = 0) and skipped in the drain.
                // Candidate k of this lane is gg + k G; every lane has at
                // least kmin = mc / G of them, so the unrolled part of each
                // batch runs a wave-uniform number of quads.
                const int kmin = __builtin_amdgcn_readfirstlane(mc) / G;
                int j = gg;
                for (int b = 0; j < mc; b += 32) {
                    const int j0 = j;
                    unsigned m = 0;
                    int n = 0;
                    const int nq = __builtin_amdgcn_readfirstlane(min(8, max(0, kmin - b) >> 2));
                    int jq = j;
                    for (int q = 0; q < nq; ++q, jq += 4 * G) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int jj = jq + u * G;
                            const float dx = sh.u.cf.x[jj] - xf, dy = sh.u.cf.y[jj] - yf,
                                        dz = sh.u.cf.z[jj] - zf;
                            // r^2 - rc^2: its sign bit is the survivor bit
                            const float d = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, nrc2_f)));
                            m = __builtin_amdgcn_alignbit(m, __float_as_uint(d), 31);
                        }
                    }
                    n = 4 * nq;
                    j += 4 * G * nq;
                    for (; n < 32 && j < mc; ++n, j += G) {
                        const float dx = sh.u.cf.x[j] - xf, dy = sh.u.cf.y[j] - yf,
                                    dz = sh.u.cf.z[j] - zf;
                        const float d = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, nrc2_f)));
                        m = __builtin_amdgcn_alignbit(m, __float_as_uint(d), 31);
                    }
                    dense_drain_mask<MASK, T, LJN>(g, p, sh, pos, xi, yi, zi, size, ov2_hi,
                                                   hi - s0, m, n, j0, G, fx, fy, fz, err, id_i,
                                                   id, st);
                }
            }
        }
        if constexpr (ST::kOn) {
            // ---- sum the G group counts of each home particle
            __syncthreads();   // the cull's copy is no longer read
            sh.u.red[0][lane] = (double)(act ? st.count : 0);
            sh.u.red[1][lane] = __longlong_as_double((long long)(act ? st.hash : 0ull));
            __syncthreads();
            if (lane < nic) {
                int pairs = 0;
                unsigned long long hs = 0;
                for (int q = 0; q < G; ++q) {
                    pairs += (int)sh.u.red[0][lane + q * nic];
                    hs += (unsigned long long)__double_as_longlong(sh.u.red[1][lane + q * nic]);
                }
                so->count[id_i] = pairs;
                so->hash[id_i] = hs;
            }
            __syncthreads();
            continue;
        }
        // ---- sum the G group partials of each home particle.  Up to 64
        // home particles: P = 64 / nic (at most 4) lanes of wave 0 per
        // particle each add every P-th partial, then the particle's lane
        // adds the P sums; beyond that, the particle's lane adds all G.
        __syncthreads();   // the cull's copy is no longer read
        sh.u.red[0][lane] = fx;
        sh.u.red[1][lane] = fy;
        sh.u.red[2][lane] = fz;
        __syncthreads();
        if (nic > 64) {
            if (lane < nic) {
                double Fx = 0.0, Fy = 0.0, Fz = 0.0;
                for (int q = 0; q < G; ++q) {
                    Fx += sh.u.red[0][lane + q * nic];
                    Fy += sh.u.red[1][lane + q * nic];
                    Fz += sh.u.red[2][lane + q * nic];
                }
                epi(i0 + lane, Fx, Fy, Fz);
            }
        } else if (lane < 64) {
            const int P = min(4, 64 / nic);
            double Fx = 0.0, Fy = 0.0, Fz = 0.0;
            if (lane < P * nic) {
                const int gi2 = lane % nic;
                for (int q = lane / nic; q < G; q += P) {
                    Fx += sh.u.red[0][gi2 + q * nic];
                    Fy += sh.u.red[1][gi2 + q * nic];
                    Fz += sh.u.red[2][gi2 + q * nic];
                }
            }
            for (int k = 1; k < P; ++k) {
                const int src = lane + k * nic;   // below 64 for lane < nic
                const double ax = __shfl(Fx, src, 64), ay = __shfl(Fy, src, 64),
                             az = __shfl(Fz, src, 64);
                if (lane < nic) {
                    Fx += ax;
                    Fy += ay;
                    Fz += az;
                }
            }
            if (lane < nic) epi(i0 + lane, Fx, Fy, Fz);
        }
        __syncthreads();
    }
}

}  // namespace psim
