// psim_engine.hip -- gfx950 kernels and the C ABI (include/psim_gpu.h).
//
// Data layout in HBM (one context = one system):
//   slot-ordered state, SoA [3][N] FP64: pos, pos0, frc (the force of the
//   previous step, i.e. frc0 of the next integrator call), mem (memX);
//   id[N] = original index of each slot.  Slots are kept sorted by cell
//   hash, so the cell table is a start-offset array cstart[S^3+1] and the
//   particles of one cell are contiguous in every state array.
//   vel[3][N] is kept by ORIGINAL index (read on the first step, written on
//   velocity steps only).
//
// One step (the order of system::run, system.cpp:350-374):
//   k_step      fused: 27-cell forces F(x_n) (IForce::getAcceleration),
//               push (frc0 <- frc), G&B integration (firstStep/normalStep/
//               velocityStep), wrap (particle::setPos), new cell hash and
//               an atomic per-cell count; writes the next state in the
//               current slot order to the "n*" buffers.
//   k_scan_*    counts -> cstart (exclusive scan), counts cleared.
//   k_scatter   slot of each particle in its new cell (arbitrary order).
//   k_move      canonical within-cell order (ascending previous slot, i.e.
//               a stable counting sort) and the state gather into the
//               sorted arrays.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "psim_gpu.h"
#include "psim_dense.hpp"
#include "psim_pairs.hpp"

using namespace psim;

namespace {

constexpr int kScanThreads = 1024;
// mean particles per cell above which the cell-per-workgroup kernel is used
constexpr double kDenseMinOccupancy = 2.0;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
struct StepArgs {
    const double *pos, *pos0, *frc, *mem;
    const int *id, *cstart;
    double *npos, *npos0, *nfrc, *nmem;
    int *nid;
    double *vel;         // [3][N] by original index
    int *key, *arank, *count;
    DevErr *err;
    const double *noise; // [6][N] by original index, or null -> Philox
    unsigned long long seed, step;
    int first, do_vel;
};

// G&B epilogue for the particle in slot s with total force F
// (brownianIntegrator.cpp:191-356, particle.cpp:61-99); writes the next state
// at slot s, the new cell key and an atomic per-cell rank.
template <bool WRITE_F>
__device__ void integrate_one(const Geo &g, const GB &gb, const StepArgs &a, int s,
                              double fx, double fy, double fz)
{
    const int N = g.n;
    const int o = a.id[s];
    double z[6];
    if (a.noise) {
#pragma unroll
        for (int k = 0; k < 6; ++k) z[k] = a.noise[(size_t)k * N + o];
    } else {
        philox_normals6(a.seed, a.step, (uint32_t)o, z);
    }
    const double xc[3] = {a.pos[s], a.pos[N + s], a.pos[2 * N + s]};
    const double F[3] = {fx, fy, fz};
    double x0c[3], xn[3], M[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x0c[d] = a.pos0[d * N + s];
    if (a.first) {
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
            Mo[d] = a.mem[d * N + s];
            Fo[d] = a.frc[d * N + s];
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
        if (a.do_vel) {
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
        a.npos[d * N + s] = xw[d];
        a.npos0[d * N + s] = x0n;
        if (WRITE_F) a.nfrc[d * N + s] = F[d];
        a.nmem[d * N + s] = M[d];
    }
    a.nid[s] = o;
    if (bad) report_error(a.err, 7703, o, -1);
    int k = cell_1d(xw[0], g.cell, g.S) + g.S * cell_1d(xw[1], g.cell, g.S)
          + g.S2 * cell_1d(xw[2], g.cell, g.S);
    a.key[s] = k;
    a.arank[s] = wave_aggregated_rank(a.count, k);
}

// Thread-per-particle step (sparse systems and S < 4).
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_step(Geo g, Pot p, GB gb, double radius,
                                                 StepArgs a)
{
    __shared__ unsigned hits[kHitCap * kBlock];
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= g.n) return;
    const int N = g.n;
    const double xi = a.pos[s], yi = a.pos[N + s], zi = a.pos[2 * N + s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{a.pos, N, radius};
    StartTab tab{a.cstart};
    Orig orig{a.id, 1, 0};
    if (g.exact_mi)
        gather_force<MASK, true>(g, p, src, tab, s, xi, yi, zi, radius,
                                 hits + threadIdx.x, fx, fy, fz, a.err, orig);
    else
        gather_force<MASK, false>(g, p, src, tab, s, xi, yi, zi, radius,
                                  hits + threadIdx.x, fx, fy, fz, a.err, orig);
    integrate_one<true>(g, gb, a, s, fx, fy, fz);
}

// Cell-per-workgroup forces for the step (dense systems, S >= 4): the total
// force on each particle goes to nfrc[slot], where k_integrate picks it up.
// (Fusing the G&B epilogue here would run it on n_i of the workgroup's lanes,
// latency-bound; the separate kernel runs it at full occupancy.)
template <int MASK>
__global__ __launch_bounds__(kDenseThreads) void k_step_dense(Geo g, Pot p, double radius,
                                                              const double *pos,
                                                              const int *id,
                                                              const int *cstart,
                                                              double *nfrc, DevErr *err,
                                                              int per_xcd)
{
    __shared__ DenseShared sh;
    const int c = xcd_cell(blockIdx.x, per_xcd);
    if (c >= g.ncell) return;
    const int home0 = cstart[c];
    const int N = g.n;
    dense_cell_forces<MASK>(g, p, radius, pos, cstart, id, err, c, sh,
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

// Forces only at the current state, written by original index ([3][N]).
template <int MASK>
__global__ __launch_bounds__(kBlock) void k_forces(Geo g, Pot p, double radius,
                                                   const double *pos, const int *id,
                                                   const int *cstart, double *fout,
                                                   DevErr *err)
{
    __shared__ unsigned hits[kHitCap * kBlock];
    const int s = blockIdx.x * kBlock + threadIdx.x;
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
    const int o = id[s];
    fout[o] = fx;
    fout[N + o] = fy;
    fout[2 * N + o] = fz;
}

template <int MASK>
__global__ __launch_bounds__(kDenseThreads) void k_forces_dense(Geo g, Pot p, double radius,
                                                                const double *pos,
                                                                const int *id,
                                                                const int *cstart,
                                                                double *fout, DevErr *err,
                                                                int per_xcd)
{
    __shared__ DenseShared sh;
    const int c = xcd_cell(blockIdx.x, per_xcd);
    if (c >= g.ncell) return;
    const int home0 = cstart[c];
    const int N = g.n;
    dense_cell_forces<MASK>(g, p, radius, pos, cstart, id, err, c, sh,
                            [=](int i, double Fx, double Fy, double Fz) {
                                const int o = id[home0 + i];
                                fout[o] = Fx;
                                fout[N + o] = Fy;
                                fout[2 * N + o] = Fz;
                            });
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

// Cell hash + atomic count for an arbitrary order (initial build).
__global__ void k_hash_count(Geo g, const double *pos, int *key, int *arank, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.n) return;
    const int N = g.n;
    int k = cell_1d(pos[s], g.cell, g.S) + g.S * cell_1d(pos[N + s], g.cell, g.S)
          + g.S2 * cell_1d(pos[2 * N + s], g.cell, g.S);
    key[s] = k;
    arank[s] = wave_aggregated_rank(count, k);
}

// Exclusive scan over a 1024-thread block of one value per thread.
__device__ int block_exclusive_scan(int v, int *total)
{
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    if (w == 0) {
        int t = lane < kScanThreads / 64 ? wsum[lane] : 0;
#pragma unroll
        for (int d = 1; d < kScanThreads / 64; d <<= 1) {
            int y = __shfl_up(t, d, 64);
            if (lane >= d) t += y;
        }
        if (lane < kScanThreads / 64) wsum[lane] = t;
    }
    __syncthreads();
    int pre = (w > 0 ? wsum[w - 1] : 0) + x - v;
    *total = wsum[kScanThreads / 64 - 1];
    __syncthreads();
    return pre;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int *count, int ncell,
                                                              int *bsum)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < ncell) v += count[base + k];
    int total;
    block_exclusive_scan(v, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_bsums(int *bsum, int nb)
{
    int carry = 0;
    for (int base = 0; base < nb; base += kScanThreads) {
        int i = base + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        int total;
        int pre = block_exclusive_scan(v, &total);
        if (i < nb) bsum[i] = carry + pre;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(int *count, int ncell,
                                                             const int *bsum,
                                                             int *cstart, int n)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int c[kScanItems];
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        c[k] = (base + k < ncell) ? count[base + k] : 0;
        v += c[k];
    }
    int total;
    int pre = block_exclusive_scan(v, &total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < ncell) {
            cstart[base + k] = pre;
            count[base + k] = 0;
        }
        pre += c[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cstart[ncell] = n;
}

__global__ void k_scatter(int n, const int *key, const int *arank, const int *cstart,
                          int *perm)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    perm[cstart[key[s]] + arank[s]] = s;
}

// Canonical order within a cell = ascending source slot (stable counting
// sort); gathers the state into the sorted arrays.
__global__ void k_move(int n, const int *perm, const int *key, const int *cstart,
                       const double *npos, const double *npos0, const double *nfrc,
                       const double *nmem, const int *nid, double *pos, double *pos0,
                       double *frc, double *mem, int *id)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm[t];
    const int c = key[s];
    const int b = cstart[c], e = cstart[c + 1];
    int rank = 0;
    for (int u = b; u < e; ++u) rank += (perm[u] < s);
    const int dst = b + rank;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        pos[d * n + dst] = npos[d * n + s];
        pos0[d * n + dst] = npos0[d * n + s];
        frc[d * n + dst] = nfrc[d * n + s];
        mem[d * n + dst] = nmem[d * n + s];
    }
    id[dst] = nid[s];
}

}  // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct psim_ctx {
    psim_params prm;
    Geo geo;
    Pot pot;
    GB gb;
    int n = 0, ncell = 0, nscan = 0;
    bool dense = false;   // cell-per-workgroup pair kernel (else thread per particle)
    hipStream_t stream = nullptr;
    double *pos = nullptr, *pos0 = nullptr, *frc = nullptr, *mem = nullptr;
    double *npos = nullptr, *npos0 = nullptr, *nfrc = nullptr, *nmem = nullptr;
    double *vel = nullptr, *fout = nullptr;
    int *id = nullptr, *nid = nullptr, *key = nullptr, *arank = nullptr, *perm = nullptr;
    int *count = nullptr, *cstart = nullptr, *bsum = nullptr;
    DevErr *err = nullptr;
    double *noise = nullptr;
    size_t noise_cap = 0;
    double *tsorted = nullptr;
    int *thi = nullptr, *tcse = nullptr;
    double time = 0.0;
    int vel_counter = 0;
    unsigned long long step_index = 0;
    bool have_state = false;
    // profiling
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct Mark { size_t e0, e1, e1b, e2; };
    std::vector<Mark> marks;
    psim_kernel_times times{};
};

#define CK(x)                                                                    \
    do {                                                                         \
        hipError_t _e = (x);                                                     \
        if (_e != hipSuccess) {                                                  \
            std::fprintf(stderr, "psim_gpu: %s failed: %s (%s:%d)\n", #x,        \
                         hipGetErrorString(_e), __FILE__, __LINE__);             \
            return PSIM_E_HIP;                                                   \
        }                                                                        \
    } while (0)

namespace {

void fill_constants(psim_ctx *c)
{
    const psim_params &P = c->prm;
    Geo &g = c->geo;
    g.n = P.n_particles;
    g.S = P.cells_per_dim;
    g.S2 = g.S * g.S;
    g.ncell = g.S2 * g.S;
    g.L = P.box;
    g.half_pbc = P.box / 2.0;
    if (P.half_uv > 0)
        g.half_uv = P.half_uv;
    else if (std::floor(P.box) == P.box)
        g.half_uv = (double)((long long)P.box / 2);   // int L / 2 (utilities.cpp:124)
    else
        g.half_uv = P.box / 2.0;
    g.cell = P.cell_size;
    g.exact_mi = (g.S < 4) ? 1 : 0;
    g.pad = 0;

    Pot &p = c->pot;
    p.mask = P.force_mask;
    p.lj_num = P.lj_num;
    const double rc = P.cutoff;
    p.rc2 = rc * rc;
    // guard band: the fast image distance differs from pbcDist's by a few
    // ulps of L on each component
    double band = 1e-12 * std::fmax(1.0, P.box / rc);
    p.rc2_hi = p.rc2 * (1.0 + band);
    p.rc2_lo = p.rc2 * (1.0 - band);
    p.ov_factor = 0.8;
    p.ov_band = band;
    p.lj_neg_kt_eps = -P.kT * P.lj_well_depth;
    p.lj_four_n = 4.0 * P.lj_num;
    p.debye_len = P.debye_length;
    p.debye_inv = 1.0 / P.debye_length;
    p.yuk_a = P.yukawa_strength;
    // AOPotential.cpp:50-55
    double q = (rc / (rc - 1.0));
    double a1 = -P.kT * P.ao_well_depth * q * q * q;
    double a2 = -3.0 / (2.0 * rc);
    double a3 = 1.0 / (2.0 * rc * rc * rc);
    p.ao_c1 = -a1 * a2;
    p.ao_c2 = -3.0 * a1 * a3;

    // brownianIntegrator.cpp:53-81, 111-175 (host FP64, glibc libm)
    const double kT = P.kT, gamma = P.gamma, dt = P.dt;
    double y = gamma * dt;
    double coeff0, coeff1, coeff2, coeff3, sig1, sig2, gn, goy2, goy3, hn;
    {
        double ty = 2.0 * y;
        coeff0 = std::exp(-y);
        double aa1 = 1.0 - coeff0;
        double aa2 = 0.5 * y * (1.0 + coeff0) - aa1;
        double aa3 = y - aa1;
        coeff1 = aa1 / y;
        coeff2 = aa2 / (y * y);
        coeff3 = aa3 / (y * y);
        sig1 = 2.0 * y - 3.0 + 4.0 * std::exp(-y) - std::exp(-ty);
        sig2 = -2.0 * y - 3.0 + 4.0 * std::exp(y) - std::exp(ty);
        gn = std::exp(y) - ty - std::exp(-y);
        goy2 = gn / (y * y);
        goy3 = gn / (y * y * y);
        hn = y / (std::exp(y) - std::exp(-y));
    }
    if (gamma < 0.05) {
        double y1 = y, y2 = y1 * y1, y3 = y2 * y1, y4 = y3 * y1, y5 = y4 * y1;
        double y6 = y5 * y1, y7 = y6 * y1, y8 = y7 * y1, y9 = y8 * y1;
        coeff1 = 1.0 - 0.5 * y1 + (1.0 / 6.0) * y2 - (1.0 / 24.0) * y3 + (1.0 / 120.0) * y4;
        coeff2 = (1.0 / 12.0) * y1 - (1.0 / 24.0) * y2 + (1.0 / 80.0) * y3 - (1.0 / 360.0) * y4;
        coeff3 = 0.5 - (1.0 / 6.0) * y1 + (1.0 / 24.0) * y2 - (1.0 / 120.0) * y3;
        sig1 = +(2.0 / 3.0) * y3 - 0.5 * y4 + (7.0 / 30.0) * y5 - (1.0 / 12.0) * y6
             + (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 + (127.0 / 90720.0) * y9;
        sig2 = -(2.0 / 3.0) * y3 - 0.5 * y4 - (7.0 / 30.) * y5 - (1.0 / 12.0) * y6
             - (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 - (127.0 / 90720.0) * y9;
        goy2 = (1.0 / 3.0) * y1 + (1.0 / 60.0) * y3;
        goy3 = 1.0 / 3.0 + (1.0 / 60.0) * y2;
        hn = 0.5 - (1.00 / 12.0) * y2 + (7.0 / 720.0) * y4;
        gn = (1.0 / 3.0) * y3 + (1.0 / 60.0) * y5;
    }
    if (gamma == 0) {
        coeff0 = 1.0;
        coeff1 = 1.0;
        coeff2 = 0.0;
        coeff3 = 0.5;
    }
    double gamma2 = gamma * gamma;
    sig1 = std::sqrt(+kT * sig1 / gamma2);
    sig2 = std::sqrt(-kT * sig2 / gamma2);
    double corr = (kT / (gamma2)) * (gn / (sig1 * sig2));
    double dev = std::sqrt(1.0 - (corr * corr));
    GB &b = c->gb;
    b.coeff0 = coeff0;
    b.coeff1 = coeff1;
    b.coeff3 = coeff3;
    b.c0 = 1.0 + coeff0;
    double m = 1.0 / P.mass;
    double dt2 = dt * dt;
    b.c1 = m * dt2 * coeff1;
    b.c2 = m * dt2 * coeff2;
    b.sig1 = sig1;
    b.sig2 = sig2;
    b.corr = corr;
    b.dev = dev;
    b.m = m;
    b.dt = dt;
    double dt3 = dt * dt2;
    b.g2 = m * dt2 * goy2;
    b.g3 = m * dt3 * goy3;
    b.g4 = hn * (1.0 / dt);
}

template <class T>
int dalloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        std::fprintf(stderr, "psim_gpu: hipMalloc(%zu B) failed: %s\n", count * sizeof(T),
                     hipGetErrorString(e));
        *p = nullptr;
        return PSIM_E_NOMEM;
    }
    return 0;
}

inline unsigned blocks_for(long long n, int threads)
{
    return (unsigned)((n + threads - 1) / threads);
}

hipEvent_t take_event(psim_ctx *c, size_t *idx)
{
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    *idx = c->ev_used;
    return c->ev_pool[c->ev_used++];
}

int collect_times(psim_ctx *c)
{
    if (c->marks.empty()) return 0;
    CK(hipStreamSynchronize(c->stream));
    for (const auto &mk : c->marks) {
        float f = 0.f, it = 0.f, b = 0.f;
        CK(hipEventElapsedTime(&f, c->ev_pool[mk.e0], c->ev_pool[mk.e1]));
        CK(hipEventElapsedTime(&it, c->ev_pool[mk.e1], c->ev_pool[mk.e1b]));
        CK(hipEventElapsedTime(&b, c->ev_pool[mk.e1b], c->ev_pool[mk.e2]));
        c->times.force_ms += f;
        c->times.force_launches += 1;
        c->times.integrate_ms += it;
        c->times.integrate_launches += c->dense ? 1 : 0;
        c->times.build_ms += b;
        c->times.build_launches += 1;
        c->times.total_ms += f + it + b;
        c->times.steps += 1;
    }
    c->marks.clear();
    c->ev_used = 0;
    return 0;
}

// counts (already accumulated) -> cstart, scatter, canonical move.
int launch_build(psim_ctx *c)
{
    const int n = c->n, nc = c->ncell;
    hipStream_t st = c->stream;
    k_scan_reduce<<<c->nscan, kScanThreads, 0, st>>>(c->count, nc, c->bsum);
    k_scan_bsums<<<1, kScanThreads, 0, st>>>(c->bsum, c->nscan);
    k_scan_apply<<<c->nscan, kScanThreads, 0, st>>>(c->count, nc, c->bsum, c->cstart, n);
    k_scatter<<<blocks_for(n, 256), 256, 0, st>>>(n, c->key, c->arank, c->cstart, c->perm);
    k_move<<<blocks_for(n, 256), 256, 0, st>>>(n, c->perm, c->key, c->cstart, c->npos,
                                               c->npos0, c->nfrc, c->nmem, c->nid, c->pos,
                                               c->pos0, c->frc, c->mem, c->id);
    CK(hipGetLastError());
    return 0;
}

template <int MASK>
void launch_step_mask(psim_ctx *c, const StepArgs &a)
{
    if (c->dense) {
        const int per = (c->ncell + 7) / 8;
        k_step_dense<MASK><<<8 * per, kDenseThreads, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, a.pos, a.id, a.cstart, a.nfrc, a.err, per);
    } else {
        k_step<MASK><<<blocks_for(c->n, kBlock), kBlock, 0, c->stream>>>(c->geo, c->pot, c->gb,
                                                                        c->prm.radius, a);
    }
}

int launch_step(psim_ctx *c, const StepArgs &a)
{
    switch (c->pot.mask) {
    case 1: launch_step_mask<1>(c, a); break;
    case 2: launch_step_mask<2>(c, a); break;
    case 3: launch_step_mask<3>(c, a); break;
    case 4: launch_step_mask<4>(c, a); break;
    case 5: launch_step_mask<5>(c, a); break;
    case 6: launch_step_mask<6>(c, a); break;
    case 7: launch_step_mask<7>(c, a); break;
    default: return PSIM_E_ARG;
    }
    CK(hipGetLastError());
    return 0;
}

template <int MASK>
void launch_forces_mask(psim_ctx *c)
{
    if (c->dense) {
        const int per = (c->ncell + 7) / 8;
        k_forces_dense<MASK><<<8 * per, kDenseThreads, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, c->pos, c->id, c->cstart, c->fout, c->err, per);
    } else {
        k_forces<MASK><<<blocks_for(c->n, kBlock), kBlock, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, c->pos, c->id, c->cstart, c->fout, c->err);
    }
}

template <int MASK>
void launch_table_mask(psim_ctx *c)
{
    k_forces_table<MASK><<<blocks_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->geo, c->pot, c->tsorted, c->thi, c->tcse, c->fout, c->err);
}

int run_steps(psim_ctx *c, int nsteps, const double *d_noise)
{
    if (!c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    for (int s = 0; s < nsteps; ++s) {
        StepArgs a;
        a.pos = c->pos; a.pos0 = c->pos0; a.frc = c->frc; a.mem = c->mem;
        a.id = c->id; a.cstart = c->cstart;
        a.npos = c->npos; a.npos0 = c->npos0; a.nfrc = c->nfrc; a.nmem = c->nmem;
        a.nid = c->nid; a.vel = c->vel; a.key = c->key; a.arank = c->arank;
        a.count = c->count; a.err = c->err;
        a.noise = d_noise ? d_noise + (size_t)s * 6 * n : nullptr;
        a.seed = c->prm.seed;
        a.step = c->step_index;
        a.first = (c->time == 0) ? 1 : 0;   // brownianIntegrator.cpp:183
        a.do_vel = 0;
        if (!a.first) {
            a.do_vel = (c->prm.vel_freq == 0 || c->vel_counter == c->prm.vel_freq) ? 1 : 0;
            // :292
            if (c->vel_counter == c->prm.vel_freq) c->vel_counter = 0; else c->vel_counter++;
        }
        psim_ctx::Mark mk{};
        if (c->profiling) {
            hipEvent_t e = take_event(c, &mk.e0);
            if (e) CK(hipEventRecord(e, c->stream));
        }
        int rc = launch_step(c, a);
        if (rc) return rc;
        if (c->profiling) {
            hipEvent_t e = take_event(c, &mk.e1);
            if (e) CK(hipEventRecord(e, c->stream));
        }
        if (c->dense) {
            k_integrate<<<blocks_for(n, 256), 256, 0, c->stream>>>(c->geo, c->gb, a);
            CK(hipGetLastError());
        }
        if (c->profiling) {
            hipEvent_t e = take_event(c, &mk.e1b);
            if (e) CK(hipEventRecord(e, c->stream));
        }
        rc = launch_build(c);
        if (rc) return rc;
        if (c->profiling) {
            hipEvent_t e = take_event(c, &mk.e2);
            if (e) CK(hipEventRecord(e, c->stream));
            c->marks.push_back(mk);
            if (c->ev_used > 3000) {
                rc = collect_times(c);
                if (rc) return rc;
            }
        }
        c->step_index += 1;
        c->time += c->prm.dt;   // system.cpp:373
    }
    return 0;
}

int read_error(psim_ctx *c, int *code, int *i, int *j)
{
    DevErr h{};
    CK(hipMemcpyAsync(&h, c->err, sizeof(DevErr), hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    if (code) *code = h.code;
    if (i) *i = h.i;
    if (j) *j = h.j;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int psim_gpu_abi_version(void) { return PSIM_GPU_ABI_VERSION; }

const char *psim_gpu_status_string(int s)
{
    switch (s) {
    case PSIM_OK: return "ok";
    case PSIM_ERR_INIT: return "7701: initial configuration could not be created";
    case PSIM_ERR_BOUNDS: return "7703: particle out of bounds";
    case PSIM_ERR_OVERLAP: return "7704: significant particle overlap";
    case PSIM_ERR_NAN: return "7705: numerically unstable system (NaN force)";
    case PSIM_E_ARG: return "invalid argument";
    case PSIM_E_HIP: return "HIP runtime error";
    case PSIM_E_NOMEM: return "device allocation failed";
    case PSIM_E_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

int psim_gpu_create(const psim_params *params, psim_ctx **out)
{
    if (!params || !out) return PSIM_E_ARG;
    *out = nullptr;
    const psim_params &P = *params;
    if (P.n_particles <= 0 || P.n_particles >= (1 << kIdxBits)) return PSIM_E_ARG;
    // the 27-cell stencil repeats cells for S < 3 (double counting in the
    // reference, LJPotential.cpp:154-169)
    if (P.cells_per_dim < 3 || P.cells_per_dim > 1290) return PSIM_E_UNSUPPORTED;
    if (P.kernel_path < 0 || P.kernel_path > 2) return PSIM_E_ARG;
    if (!(P.box > 0) || !(P.cell_size > 0) || !(P.cutoff > 0) || !(P.dt > 0) ||
        !(P.mass > 0) || P.force_mask <= 0 || P.force_mask > 7)
        return PSIM_E_ARG;
    if ((P.force_mask & PSIM_FORCE_LJ) && !(P.debye_length > 0)) return PSIM_E_ARG;
    psim_ctx *c = new (std::nothrow) psim_ctx();
    if (!c) return PSIM_E_NOMEM;
    c->prm = P;
    fill_constants(c);
    c->n = P.n_particles;
    c->ncell = c->geo.ncell;
    c->nscan = (c->ncell + kScanTile - 1) / kScanTile;
    // kernel path: the cell-per-workgroup kernel needs S >= 4 (shifted image
    // = minimum image) and pays off once cells hold a few particles
    if (P.kernel_path == 1)
        c->dense = false;
    else if (P.kernel_path == 2)
        c->dense = c->geo.S >= 4;
    else
        c->dense = c->geo.S >= 4 && (double)c->n / c->ncell >= kDenseMinOccupancy;
    const size_t n3 = 3 * (size_t)c->n;
    int rc = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = PSIM_E_HIP;
    if (!rc) rc = dalloc(&c->pos, n3);
    if (!rc) rc = dalloc(&c->pos0, n3);
    if (!rc) rc = dalloc(&c->frc, n3);
    if (!rc) rc = dalloc(&c->mem, n3);
    if (!rc) rc = dalloc(&c->npos, n3);
    if (!rc) rc = dalloc(&c->npos0, n3);
    if (!rc) rc = dalloc(&c->nfrc, n3);
    if (!rc) rc = dalloc(&c->nmem, n3);
    if (!rc) rc = dalloc(&c->vel, n3);
    if (!rc) rc = dalloc(&c->fout, n3);
    if (!rc) rc = dalloc(&c->id, c->n);
    if (!rc) rc = dalloc(&c->nid, c->n);
    if (!rc) rc = dalloc(&c->key, c->n);
    if (!rc) rc = dalloc(&c->arank, c->n);
    if (!rc) rc = dalloc(&c->perm, c->n);
    if (!rc) rc = dalloc(&c->count, c->ncell);
    if (!rc) rc = dalloc(&c->cstart, (size_t)c->ncell + 1);
    if (!rc) rc = dalloc(&c->bsum, (size_t)c->nscan);
    if (!rc) rc = dalloc(&c->err, 1);
    if (!rc && hipMemsetAsync(c->count, 0, sizeof(int) * c->ncell, c->stream) != hipSuccess)
        rc = PSIM_E_HIP;
    if (!rc && hipMemsetAsync(c->err, 0, sizeof(DevErr), c->stream) != hipSuccess)
        rc = PSIM_E_HIP;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PSIM_E_HIP;
    if (rc) {
        psim_gpu_destroy(c);
        return rc;
    }
    *out = c;
    return 0;
}

void psim_gpu_destroy(psim_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->pos, c->pos0, c->frc, c->mem, c->npos, c->npos0, c->nfrc,
                    c->nmem, c->vel, c->fout, c->id, c->nid, c->key, c->arank,
                    c->perm, c->count, c->cstart, c->bsum, c->err, c->noise,
                    c->tsorted, c->thi, c->tcse};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int psim_gpu_set_state(psim_ctx *c, const double *pos, const double *pos0,
                       const double *vel, const double *frc, const double *frc0,
                       const double *memx, double current_time, int vel_counter,
                       unsigned long long step_index)
{
    if (!c || !pos) return PSIM_E_ARG;
    const int n = c->n;
    const double L = c->prm.box;
    // host AoS3 (original order) -> SoA into the n* staging arrays
    std::vector<double> h(3 * (size_t)n);
    auto soa = [&](const double *src, double *dst, bool is_pos0) -> int {
        for (int i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                double v;
                if (src) {
                    v = src[3 * (size_t)i + d];
                } else if (is_pos0) {
                    // a fresh particle: pos0 = safeMod0(0, x) (particle.cpp:31-66)
                    double x = pos[3 * (size_t)i + d];
                    double dx = x - 0.0;
                    v = (std::fabs(dx) > (L / 2)) ? ((dx < 0) ? 0.0 - L : 0.0 + L) : 0.0;
                } else {
                    v = 0.0;
                }
                h[(size_t)d * n + i] = v;
            }
        CK(hipMemcpyAsync(dst, h.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          c->stream));
        CK(hipStreamSynchronize(c->stream));
        return 0;
    };
    int rc;
    if ((rc = soa(pos, c->npos, false))) return rc;
    if ((rc = soa(pos0, c->npos0, true))) return rc;
    if ((rc = soa(vel, c->vel, false))) return rc;
    // frc holds the force of the previous step (the next push makes it frc0);
    // frc0 of the caller is not needed by the integrator after the push
    (void)frc0;
    if ((rc = soa(frc, c->nfrc, false))) return rc;
    if ((rc = soa(memx, c->nmem, false))) return rc;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = i;
    CK(hipMemcpyAsync(c->nid, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    CK(hipMemsetAsync(c->count, 0, sizeof(int) * c->ncell, c->stream));
    k_hash_count<<<blocks_for(n, 256), 256, 0, c->stream>>>(c->geo, c->npos, c->key,
                                                            c->arank, c->count);
    CK(hipGetLastError());
    rc = launch_build(c);
    if (rc) return rc;
    CK(hipStreamSynchronize(c->stream));
    c->time = current_time;
    c->vel_counter = vel_counter;
    c->step_index = step_index;
    c->have_state = true;
    return 0;
}

int psim_gpu_get_state(psim_ctx *c, double *pos, double *pos0, double *vel, double *frc,
                       double *frc0, double *memx, double *current_time,
                       int *vel_counter, unsigned long long *step_index)
{
    if (!c || !c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    std::vector<int> ids(n);
    std::vector<double> h(3 * (size_t)n);
    CK(hipMemcpyAsync(ids.data(), c->id, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    auto aos = [&](const double *src, double *dst, bool by_slot) -> int {
        if (!dst) return 0;
        CK(hipMemcpyAsync(h.data(), src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                          c->stream));
        CK(hipStreamSynchronize(c->stream));
        for (int s = 0; s < n; ++s) {
            int o = by_slot ? ids[s] : s;
            for (int d = 0; d < 3; ++d) dst[3 * (size_t)o + d] = h[(size_t)d * n + s];
        }
        return 0;
    };
    int rc;
    if ((rc = aos(c->pos, pos, true))) return rc;
    if ((rc = aos(c->pos0, pos0, true))) return rc;
    if ((rc = aos(c->vel, vel, false))) return rc;
    if ((rc = aos(c->frc, frc, true))) return rc;
    if ((rc = aos(c->mem, memx, true))) return rc;
    if (frc0) {
        // frc0 of the last integrator call is not retained on the device
        for (size_t k = 0; k < 3 * (size_t)n; ++k) frc0[k] = std::nan("");
    }
    if (current_time) *current_time = c->time;
    if (vel_counter) *vel_counter = c->vel_counter;
    if (step_index) *step_index = c->step_index;
    return 0;
}

int psim_gpu_step(psim_ctx *c, int nsteps, const double *noise)
{
    if (!c || nsteps < 0) return PSIM_E_ARG;
    if (nsteps == 0) return 0;
    const double *d_noise = nullptr;
    if (noise) {
        size_t need = (size_t)nsteps * 6 * c->n;
        if (need > c->noise_cap) {
            if (c->noise) CK(hipFree(c->noise));
            c->noise = nullptr;
            c->noise_cap = 0;
            int rc = dalloc(&c->noise, need);
            if (rc) return rc;
            c->noise_cap = need;
        }
        CK(hipMemcpyAsync(c->noise, noise, need * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
        d_noise = c->noise;
    }
    int rc = run_steps(c, nsteps, d_noise);
    if (rc) return rc;
    int code = 0;
    rc = read_error(c, &code, nullptr, nullptr);
    if (rc) return rc;
    return code;
}

int psim_gpu_step_device_noise(psim_ctx *c, int nsteps, const double *d_noise)
{
    if (!c || nsteps < 0) return PSIM_E_ARG;
    return run_steps(c, nsteps, d_noise);
}

int psim_gpu_forces(psim_ctx *c, double *force_out)
{
    if (!c || !force_out || !c->have_state) return PSIM_E_ARG;
    switch (c->pot.mask) {
    case 1: launch_forces_mask<1>(c); break;
    case 2: launch_forces_mask<2>(c); break;
    case 3: launch_forces_mask<3>(c); break;
    case 4: launch_forces_mask<4>(c); break;
    case 5: launch_forces_mask<5>(c); break;
    case 6: launch_forces_mask<6>(c); break;
    case 7: launch_forces_mask<7>(c); break;
    default: return PSIM_E_ARG;
    }
    CK(hipGetLastError());
    const int n = c->n;
    std::vector<double> h(3 * (size_t)n);
    CK(hipMemcpyAsync(h.data(), c->fout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) force_out[3 * (size_t)i + d] = h[(size_t)d * n + i];
    int code = 0;
    int rc = read_error(c, &code, nullptr, nullptr);
    return rc ? rc : code;
}

int psim_gpu_force_table(psim_ctx *c, const double *sorted4, const int *hash_index,
                         const int *cell_start_end, double *force_out)
{
    if (!c || !sorted4 || !hash_index || !cell_start_end || !force_out) return PSIM_E_ARG;
    const int n = c->n;
    // the kernel indexes sorted4 through these tables: reject anything that
    // would leave [0, n)
    for (int h = 0; h < c->ncell; ++h) {
        int b = cell_start_end[2 * h], e = cell_start_end[2 * h + 1];
        if (b == -1) continue;
        if (b < 0 || e < b || e > n) return PSIM_E_ARG;
    }
    for (int i = 0; i < n; ++i) {
        int o = hash_index[2 * i + 1];
        if (o < 0 || o >= n) return PSIM_E_ARG;
    }
    int rc = 0;
    if (!c->tsorted) {
        if (!rc) rc = dalloc(&c->tsorted, 4 * (size_t)n);
        if (!rc) rc = dalloc(&c->thi, 2 * (size_t)n);
        if (!rc) rc = dalloc(&c->tcse, 2 * (size_t)c->ncell);
        if (rc) return rc;
    }
    CK(hipMemcpyAsync(c->tsorted, sorted4, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync(c->thi, hash_index, sizeof(int) * 2 * n, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync(c->tcse, cell_start_end, sizeof(int) * 2 * c->ncell,
                      hipMemcpyHostToDevice, c->stream));
    switch (c->pot.mask) {
    case 1: launch_table_mask<1>(c); break;
    case 2: launch_table_mask<2>(c); break;
    case 3: launch_table_mask<3>(c); break;
    case 4: launch_table_mask<4>(c); break;
    case 5: launch_table_mask<5>(c); break;
    case 6: launch_table_mask<6>(c); break;
    case 7: launch_table_mask<7>(c); break;
    default: return PSIM_E_ARG;
    }
    CK(hipGetLastError());
    CK(hipMemcpyAsync(force_out, c->fout, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    int code = 0;
    rc = read_error(c, &code, nullptr, nullptr);
    return rc ? rc : code;
}

int psim_gpu_get_cells(psim_ctx *c, int *key_by_orig, int *cell_start, int *orig_by_slot)
{
    if (!c || !c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    std::vector<int> ids(n), cs((size_t)c->ncell + 1);
    CK(hipMemcpyAsync(ids.data(), c->id, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    CK(hipMemcpyAsync(cs.data(), c->cstart, sizeof(int) * (c->ncell + 1),
                      hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    if (orig_by_slot) std::memcpy(orig_by_slot, ids.data(), sizeof(int) * n);
    if (cell_start) std::memcpy(cell_start, cs.data(), sizeof(int) * (c->ncell + 1));
    if (key_by_orig) {
        for (int cc = 0; cc < c->ncell; ++cc)
            for (int s = cs[cc]; s < cs[cc + 1]; ++s) key_by_orig[ids[s]] = cc;
    }
    return 0;
}

int psim_gpu_last_error(psim_ctx *c, int *code, int *i, int *j)
{
    if (!c) return PSIM_E_ARG;
    return read_error(c, code, i, j);
}

int psim_gpu_clear_error(psim_ctx *c)
{
    if (!c) return PSIM_E_ARG;
    CK(hipMemsetAsync(c->err, 0, sizeof(DevErr), c->stream));
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

int psim_gpu_sync(psim_ctx *c)
{
    if (!c) return PSIM_E_ARG;
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

int psim_gpu_set_profiling(psim_ctx *c, int enabled)
{
    if (!c) return PSIM_E_ARG;
    int rc = collect_times(c);
    c->profiling = enabled != 0;
    return rc;
}

int psim_gpu_kernel_times(psim_ctx *c, psim_kernel_times *out, int reset)
{
    if (!c) return PSIM_E_ARG;
    int rc = collect_times(c);
    if (rc) return rc;
    if (out) *out = c->times;
    if (reset) c->times = psim_kernel_times{};
    return 0;
}

void *psim_gpu_stream(psim_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"
