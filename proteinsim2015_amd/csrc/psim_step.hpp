// psim_step.hpp -- the step kernels of one system: pair forces (both
// kernel paths), the G&B integrator epilogue and force-only evaluations.
#pragma once

#include "psim_brick.hpp"
#include "psim_dense.hpp"
#include "psim_gpu.h"
#include "psim_pairs.hpp"

namespace psim {
namespace {

// mean particles per cell above which the cell-per-workgroup kernel is used
constexpr double kDenseMinOccupancy = 2.0;
// below this many particles the thread-per-particle kernels use 64-lane
// blocks (256-lane blocks would leave CUs idle: 256 CUs x 4 blocks)
constexpr int kSmallSystem = 256 * 1024;
#ifndef PSIM_SMALL_SPLIT
#define PSIM_SMALL_SPLIT 4
#endif
constexpr int kSmallSplit = PSIM_SMALL_SPLIT;   // lanes per particle below kSmallSystem
// home particles per brick, as a fraction of the lane groups of one workgroup
#ifndef PSIM_BRICK_TARGET
#define PSIM_BRICK_TARGET 0.8
#endif
constexpr double kBrickTarget = PSIM_BRICK_TARGET;

// Checks of a psim_params shared by every entry point that takes one.
inline int check_params(const psim_params &P)
{
    if (P.n_particles <= 0 || P.n_particles >= (1 << kIdxBits)) return PSIM_E_ARG;
    // the 27-cell stencil repeats cells for S < 3 (double counting in the
    // reference, LJPotential.cpp:154-169)
    if (P.cells_per_dim < 3 || P.cells_per_dim > 1290) return PSIM_E_UNSUPPORTED;
    if (P.kernel_path < 0 || P.kernel_path > 3) return PSIM_E_ARG;
    if (!(P.box > 0) || !(P.cell_size > 0) || !(P.cutoff > 0) || !(P.dt > 0) ||
        !(P.mass > 0) || P.force_mask <= 0 || P.force_mask > 7)
        return PSIM_E_ARG;
    if ((P.force_mask & PSIM_FORCE_LJ) && !(P.debye_length > 0)) return PSIM_E_ARG;
    return 0;
}

// Pair-kernel path of a system: the cell-per-workgroup kernel needs S >= 4
// (shifted image = minimum image) and pays off once cells hold a few
// particles; below that, one thread per particle (bricks of cells per
// workgroup, path 3, give the same bits but measured slower at every sparse
// density).
struct PathChoice {
    bool dense;
    bool brick;
    int brick_g;   // lanes per particle of the brick step kernel
    BrickGeo bg;
};
inline PathChoice choose_path(const psim_params &P, const Geo &g)
{
    PathChoice c{};
    const double occupancy = (double)g.n / g.ncell;
    if (P.kernel_path == 1 || P.kernel_path == 3)
        c.dense = false;
    else if (P.kernel_path == 2)
        c.dense = g.S >= 4;
    else
        c.dense = g.S >= 4 && occupancy >= kDenseMinOccupancy;
    c.brick = P.kernel_path == 3;
    c.brick_g = g.n < kSmallSystem ? kSmallSplit : 1;
    c.bg = brick_shape(g.S, occupancy, kBrickTarget * kBrickThreads / c.brick_g);
    return c;
}


struct StepArgs {
    const double *pos, *pos0, *frc, *mem;
    const int *id, *cstart;
    const int *prev;     // slot of each particle in pos0 / frc / mem (null: the same slot)
    int x0_raw;          // pos0 holds the positions the last step started from (re-imaged
                         // here against the current ones); npos0 is then not written
    double *npos, *npos0, *nfrc, *nmem;
    int *nid;            // copy of id in the current order (null: not kept)
    double *vel;         // [3][N] by original index
    int *key, *arank, *count;
    DevErr *err;
    const double *noise; // [6][N] by original index, or null -> Philox
    unsigned long long seed, step;
    int first, do_vel;
    const StepCtl *ctl;  // when set, step / first / do_vel are read from it
    int *pbc;            // [3][N] periodic crossings by original index, or null
};

// G&B epilogue for the particle in slot s with total force F
// (brownianIntegrator.cpp:191-356, particle.cpp:61-99); writes the next state
// at slot s, the new cell key and an atomic per-cell rank.
// (o = id[s] and ps = prev[s] may be read by the caller ahead of the force
// loop, so that their latency overlaps it: pass them as o_pre / ps_pre.)
template <bool WRITE_F>
__device__ __forceinline__ void integrate_one(const Geo &g, const GB &gb, const StepArgs &a, int s,
                              double fx, double fy, double fz, int o_pre = -1, int ps_pre = -1,
                              const double *z_pre = nullptr)
{
    const int N = g.n;
    const int o = o_pre >= 0 ? o_pre : a.id[s];
    const int first = a.ctl ? a.ctl->first : a.first;
    const int do_vel = a.ctl ? a.ctl->do_vel : a.do_vel;
    double z[6];
    if (z_pre) {
#pragma unroll
        for (int k = 0; k < 6; ++k) z[k] = z_pre[k];
    } else if (a.noise) {
#pragma unroll
        for (int k = 0; k < 6; ++k) z[k] = a.noise[(size_t)k * N + o];
    } else {
        philox_normals6(a.seed, a.ctl ? a.ctl->step : a.step, (uint32_t)o, z);
    }
    const double xc[3] = {a.pos[s], a.pos[N + s], a.pos[2 * N + s]};
    const double F[3] = {fx, fy, fz};
    // the previous position, force and memory where the last step left them
    const int ps = ps_pre >= 0 ? ps_pre : (a.prev ? a.prev[s] : s);
    double x0c[3], xn[3], M[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x0c[d] = a.pos0[d * N + ps];
    if (a.x0_raw) {
        // what the last step would have stored: util::safeMod0 of its start
        // position against its end position, which is this step's start
#pragma unroll
        for (int d = 0; d < 3; ++d) x0c[d] = safe_mod0(x0c[d], xc[d], g.L);
    }
    if (first) {
        // firstStep :200-218
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            M[d] = z[3 + d];
            double v = a.vel[(size_t)d * N + o];
            xn[d] = xc[d] + (v * gb.coeff1 * gb.dt)
                  + (F[d] * gb.coeff3 * gb.dt * gb.dt * gb.m) + (gb.sig1 * M[d]);
        }
    } else {
        // normalStep :238-274
        double Mo[3], Fo[3], Cr[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Mo[d] = a.mem[d * N + ps];
            Fo[d] = a.frc[d * N + ps];
            Cr[d] = gb.sig2 * ((gb.corr * Mo[d]) + (gb.dev * z[d]));
            M[d] = z[3 + d];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double q = (gb.c0 * xc[d]);
            q -= (gb.coeff0 * x0c[d]);
            q += (gb.c1 * F[d]);
            q += (gb.c2 * (F[d] - Fo[d]));
            q += (gb.sig1 * M[d]) + (gb.coeff0 * Cr[d]);
            xn[d] = q;
        }
        if (do_vel) {
            // velocityStep :305-354
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double dx0 = xc[d] - x0c[d];
                double xN = safe_mod(xn[d], g.L);
                double x0 = safe_mod0(xc[d], xN, g.L);
                double dx = xN - x0;
                double vv = dx + dx0;
                vv += (gb.g2 * F[d]);
                vv -= (gb.g3 * (F[d] - Fo[d]));
                vv += (Cr[d] - gb.sig1 * M[d]);
                vv *= gb.g4;
                a.vel[(size_t)d * N + o] = vv;
            }
        }
    }
    bool bad = false;
    double xw[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        xw[d] = safe_mod(xn[d], g.L);
        double x0n = safe_mod0(xc[d], xw[d], g.L);
        bad |= (xw[d] < 0.0) || (xw[d] >= g.L);
        if (a.pbc) {
            // analysisManager::updateParticlePBC (analysisTracker.cpp:38-56):
            // the re-imaged previous position leaves the box exactly when the
            // particle crossed a face this step
            if (x0n < 0)
                a.pbc[(size_t)d * N + o] += 1;
            else if (x0n > g.L)
                a.pbc[(size_t)d * N + o] -= 1;
        }
        a.npos[d * N + s] = xw[d];
        if (!a.x0_raw) a.npos0[d * N + s] = x0n;
        if (WRITE_F) a.nfrc[d * N + s] = F[d];
        a.nmem[d * N + s] = M[d];
    }
    if (a.nid) a.nid[s] = o;
    if (bad) report_error(a.err, 7703, o, -1);
    int k = g.c0 + cell_1d(xw[0], g.cell, g.S) + g.S * cell_1d(xw[1], g.cell, g.S)
          + g.S2 * cell_1d(xw[2], g.cell, g.S);
    a.key[s] = k;
    a.arank[s] = wave_aggregated_rank(a.count, k);
}

// Thread-per-particle step (sparse systems and S < 4); with SPLIT > 1,
// SPLIT lanes per particle share its 27 neighbour cells (small systems,
// where one lane per particle leaves most SIMDs idle and each lane waits
// on 27 dependent cell lookups).
template <int MASK, int BLOCK, int SPLIT = 1>
__global__ __launch_bounds__(BLOCK) void k_step(Geo g, Pot p, GB gb, double radius,
                                                StepArgs a)
{
    static_assert(SPLIT >= 1 && 64 % SPLIT == 0 && (SPLIT & (SPLIT - 1)) == 0, "lane groups");
    __shared__ unsigned hits[kHitCap * BLOCK];
    const int t = xcd_block() * BLOCK + threadIdx.x;
    const int s = t / SPLIT, sub = t % SPLIT;
    if (s >= g.n) return;   // whole lane groups
    const int N = g.n;
    const double xi = a.pos[s], yi = a.pos[N + s], zi = a.pos[2 * N + s];
    // the integrator's index loads, in flight during the force loop
    const int o = a.id[s], ps = a.prev ? a.prev[s] : s;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{a.pos, N, radius};
    StartTab tab{a.cstart};
    Orig orig{a.id, 1, 0};
    if (g.exact_mi)
        gather_force<MASK, true, BLOCK, SPLIT>(g, p, src, tab, s, xi, yi, zi, radius,
                                               hits + threadIdx.x, fx, fy, fz, a.err, orig,
                                               sub);
    else
        gather_force<MASK, false, BLOCK, SPLIT>(g, p, src, tab, s, xi, yi, zi, radius,
                                                hits + threadIdx.x, fx, fy, fz, a.err, orig,
                                                sub);
    // (not with LJ: its pair arithmetic leaves no registers for the pairs)
    if constexpr (SPLIT >= 4 && !(MASK & F_LJ)) {
        fx = split_sum<SPLIT>(fx);
        fy = split_sum<SPLIT>(fy);
        fz = split_sum<SPLIT>(fz);
        if (!a.noise) {
            // in-kernel noise: lanes 0..2 of the group each make one
            // Box-Muller pair, gathered on lane 0
            double za = 0.0, zb = 0.0;
            if (sub < 3)
                philox_pair(a.seed, a.ctl ? a.ctl->step : a.step, (uint32_t)o, sub, za, zb);
            const int base = (int)(threadIdx.x & 63u) & ~(SPLIT - 1);
            double z[6];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                z[2 * q] = __shfl(za, base + q, 64);
                z[2 * q + 1] = __shfl(zb, base + q, 64);
            }
            if (sub != 0) return;
            integrate_one<true>(g, gb, a, s, fx, fy, fz, o, ps, z);
            return;
        }
        if (sub != 0) return;
    } else if constexpr (SPLIT > 1) {
        fx = split_sum<SPLIT>(fx);
        fy = split_sum<SPLIT>(fy);
        fz = split_sum<SPLIT>(fz);
        if (sub != 0) return;
    }
    integrate_one<true>(g, gb, a, s, fx, fy, fz, o, ps);
}

// Brick-per-workgroup step (sparse systems): G lanes per home particle, the
// G&B epilogue fused as in k_step (same result as k_step<MASK, *, G>).
template <int MASK, int G>
__global__ __launch_bounds__(kBrickThreads) void k_step_brick(Geo g, Pot p, GB gb,
                                                              double radius, BrickGeo bg,
                                                              int per_xcd, StepArgs a)
{
    __shared__ BrickShared sh;
    const int b = xcd_cell(blockIdx.x, per_xcd);
    if (b >= bg.nbricks) return;
    auto epi = [&](int s, double fx, double fy, double fz) {
        integrate_one<true>(g, gb, a, s, fx, fy, fz);
    };
    if (g.exact_mi)
        brick_forces<MASK, true, G>(g, p, radius, a.pos, a.cstart, a.id, a.err, bg, b, sh, epi);
    else
        brick_forces<MASK, false, G>(g, p, radius, a.pos, a.cstart, a.id, a.err, bg, b, sh, epi);
}

// Cell-per-workgroup forces for the step (dense systems, S >= 4): the total
// force on each particle goes to nfrc[slot], where k_integrate picks it up.
// (Fusing the G&B epilogue here would run it on n_i of the workgroup's lanes,
// latency-bound; the separate kernel runs it at full occupancy.)
template <int MASK, int LJN>
__global__ __launch_bounds__(kDenseThreads)
__attribute__((amdgpu_waves_per_eu(dense_waves(MASK, LJN)))) void k_step_dense(
    Geo g, Pot p, double radius, const double *pos, const int *id, const int *cstart,
    double *nfrc, DevErr *err, int per_xcd)
{
    __shared__ DenseShared sh;
    const int c = xcd_cell(blockIdx.x, per_xcd);
    if (c >= g.ncell) return;
    const int home0 = cstart[c];
    const int N = g.n;
    dense_cell_forces<MASK, kDenseThreads, LJN>(g, p, radius, pos, cstart, id, err, c, sh,
                            [=](int i, double Fx, double Fy, double Fz) {
                                const int s = home0 + i;
                                nfrc[s] = Fx;
                                nfrc[N + s] = Fy;
                                nfrc[2 * N + s] = Fz;
                            });
}

// G&B step for every slot with the force already in nfrc (dense path).
__global__ __launch_bounds__(256) void k_integrate(Geo g, GB gb, StepArgs a)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= g.n) return;
    const int N = g.n;
    integrate_one<false>(g, gb, a, s, a.nfrc[s], a.nfrc[N + s], a.nfrc[2 * N + s]);
}

// The force on the particle of original index o, by original index: [3][N]
// (SoA) or, with AOS, N x 3.
template <bool AOS>
__device__ __forceinline__ void store_force(double *fout, int N, int o, double fx, double fy,
                                            double fz)
{
    if (AOS) {
        fout[3 * (size_t)o] = fx;
        fout[3 * (size_t)o + 1] = fy;
        fout[3 * (size_t)o + 2] = fz;
    } else {
        fout[o] = fx;
        fout[N + o] = fy;
        fout[2 * N + o] = fz;
    }
}

// Forces only at the current state, written by original index.
template <int MASK, bool AOS = false>
__global__ __launch_bounds__(kBlock) void k_forces(Geo g, Pot p, double radius,
                                                   const double *pos, const int *id,
                                                   const int *cstart, double *fout,
                                                   DevErr *err)
{
    __shared__ unsigned hits[kHitCap * kBlock];
    const int s = xcd_block() * kBlock + threadIdx.x;
    if (s >= g.n) return;
    const int N = g.n;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{pos, N, radius};
    StartTab tab{cstart};
    Orig orig{id, 1, 0};
    if (g.exact_mi)
        gather_force<MASK, true>(g, p, src, tab, s, xi, yi, zi, radius,
                                 hits + threadIdx.x, fx, fy, fz, err, orig);
    else
        gather_force<MASK, false>(g, p, src, tab, s, xi, yi, zi, radius,
                                  hits + threadIdx.x, fx, fy, fz, err, orig);
    store_force<AOS>(fout, N, id[s], fx, fy, fz);
}

// Forces only, brick per workgroup (same result as k_forces).
template <int MASK, bool AOS = false>
__global__ __launch_bounds__(kBrickThreads) void k_forces_brick(Geo g, Pot p, double radius,
                                                                BrickGeo bg, int per_xcd,
                                                                const double *pos,
                                                                const int *id,
                                                                const int *cstart,
                                                                double *fout, DevErr *err)
{
    __shared__ BrickShared sh;
    const int b = xcd_cell(blockIdx.x, per_xcd);
    if (b >= bg.nbricks) return;
    const int N = g.n;
    auto epi = [&](int s, double fx, double fy, double fz) {
        store_force<AOS>(fout, N, id[s], fx, fy, fz);
    };
    if (g.exact_mi)
        brick_forces<MASK, true, 1>(g, p, radius, pos, cstart, id, err, bg, b, sh, epi);
    else
        brick_forces<MASK, false, 1>(g, p, radius, pos, cstart, id, err, bg, b, sh, epi);
}

template <int MASK, int LJN, bool AOS = false>
__global__ __launch_bounds__(kDenseThreads)
__attribute__((amdgpu_waves_per_eu(dense_waves(MASK, LJN)))) void k_forces_dense(
    Geo g, Pot p, double radius, const double *pos, const int *id, const int *cstart,
    double *fout, DevErr *err, int per_xcd)
{
    __shared__ DenseShared sh;
    const int c = xcd_cell(blockIdx.x, per_xcd);
    if (c >= g.ncell) return;
    const int home0 = cstart[c];
    const int N = g.n;
    dense_cell_forces<MASK, kDenseThreads, LJN>(g, p, radius, pos, cstart, id, err, c, sh,
                            [=](int i, double Fx, double Fy, double Fz) {
                                store_force<AOS>(fout, N, id[home0 + i], Fx, Fy, Fz);
                            });
}

// ---------------------------------------------------------------------------
// Neighbour-set statistics of the pair kernels: the launch shapes and the
// acceptance code of the step kernels (k_step with BLOCK / SPLIT,
// k_step_dense, k_step_brick with G), each accepted pair recorded in an
// NbrStats instead of evaluated; counts and partner hashes by original index.
// ---------------------------------------------------------------------------
template <int BLOCK, int SPLIT>
__global__ __launch_bounds__(BLOCK) void k_stats(Geo g, Pot p, double radius, const double *pos,
                                                 const int *id, const int *cstart,
                                                 NbrStatsOut so)
{
    __shared__ unsigned hits[kHitCap * BLOCK];
    const int t = xcd_block() * BLOCK + threadIdx.x;
    const int s = t / SPLIT, sub = t % SPLIT;
    if (s >= g.n) return;   // whole lane groups
    const int N = g.n;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{pos, N, radius};
    StartTab tab{cstart};
    Orig orig{id, 1, 0};
    NbrStats st;
    if (g.exact_mi)
        gather_force<F_CAL, true, BLOCK, SPLIT>(g, p, src, tab, s, xi, yi, zi, radius,
                                                hits + threadIdx.x, fx, fy, fz, nullptr, orig,
                                                sub, &st);
    else
        gather_force<F_CAL, false, BLOCK, SPLIT>(g, p, src, tab, s, xi, yi, zi, radius,
                                                 hits + threadIdx.x, fx, fy, fz, nullptr, orig,
                                                 sub, &st);
    split_stats<SPLIT>(st);
    if (sub == 0) {
        so.count[id[s]] = st.count;
        so.hash[id[s]] = st.hash;
    }
}

__global__ __launch_bounds__(kDenseThreads)
__attribute__((amdgpu_waves_per_eu(dense_waves(F_CAL)))) void k_stats_dense(
    Geo g, Pot p, double radius, const double *pos, const int *id, const int *cstart,
    NbrStatsOut so, int per_xcd)
{
    __shared__ DenseShared sh;
    const int c = xcd_cell(blockIdx.x, per_xcd);
    if (c >= g.ncell) return;
    dense_cell_forces<F_CAL, kDenseThreads, 0, NbrStats>(
        g, p, radius, pos, cstart, id, nullptr, c, sh, [](int, double, double, double) {}, &so);
}

template <int G>
__global__ __launch_bounds__(kBrickThreads) void k_stats_brick(Geo g, Pot p, double radius,
                                                               BrickGeo bg, int per_xcd,
                                                               const double *pos, const int *id,
                                                               const int *cstart, NbrStatsOut so)
{
    __shared__ BrickShared sh;
    const int b = xcd_cell(blockIdx.x, per_xcd);
    if (b >= bg.nbricks) return;
    auto epi = [](int, double, double, double) {};
    if (g.exact_mi)
        brick_forces<F_CAL, true, G, NbrStats>(g, p, radius, pos, cstart, id, nullptr, bg, b, sh,
                                               epi, &so);
    else
        brick_forces<F_CAL, false, G, NbrStats>(g, p, radius, pos, cstart, id, nullptr, bg, b,
                                                sh, epi, &so);
}

// Forces from the reference's host arrays (IForce::getAcceleration over the
// whole system): sorted4, particleHashIndex (hash, orig), cellStartEnd.
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_forces_table(Geo g, Pot p,
                                                         const double *sorted4,
                                                         const int *hash_index,
                                                         const int *cse, double *fout,
                                                         DevErr *err)
{
    __shared__ unsigned hits[kHitCap * kBlock];
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= g.n) return;
    AoS4Pos src{sorted4};
    PairTab tab{cse};
    Orig orig{hash_index, 2, 1};
    const double xi = src.x(s), yi = src.y(s), zi = src.z(s), ri = src.rad(s);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (g.exact_mi)
        gather_force<MASK, true>(g, p, src, tab, s, xi, yi, zi, ri,
                                 hits + threadIdx.x, fx, fy, fz, err, orig);
    else
        gather_force<MASK, false>(g, p, src, tab, s, xi, yi, zi, ri,
                                  hits + threadIdx.x, fx, fy, fz, err, orig);
    const int o = orig(s);
    fout[3 * (size_t)o] = fx;   // particleForce[3*orig + k] (LJPotential.cpp:179-181)
    fout[3 * (size_t)o + 1] = fy;
    fout[3 * (size_t)o + 2] = fz;
}


}  // namespace
}  // namespace psim
