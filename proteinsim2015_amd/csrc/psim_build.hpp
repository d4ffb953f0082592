// psim_build.hpp -- the per-step cell rebuild (hash, counting sort,
// cell-start table, canonical reorder of the state), kernels shared by the
// single-system engine and the replica ensemble.
//
// Replaces system::hashParticles / sortParticles / clearCells /
// reorderParticles (systemHandling.cpp:33-85): instead of an O(N log N)
// comparison sort of (hash, index) tuples every step, each particle's new
// cell is counted with a wave-aggregated atomic during the step, the counts
// are scanned into a start table, and the state is gathered into cell order.
// Within a cell the order is the previous slot order (a stable counting
// sort), which makes every run reproducible bit for bit.
#pragma once

#include "psim_device.hpp"

namespace psim {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Cell hash + atomic count for an arbitrary order (initial build).
__global__ void k_hash_count(Geo g, const double *pos, int *key, int *arank, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.n) return;
    const int N = g.n;
    int k = g.c0 + cell_1d(pos[s], g.cell, g.S) + g.S * cell_1d(pos[N + s], g.cell, g.S)
          + g.S2 * cell_1d(pos[2 * N + s], g.cell, g.S);
    key[s] = k;
    arank[s] = wave_aggregated_rank(count, k);
}

// Exclusive scan over a 1024-thread block of one value per thread; *total
// receives the block's sum.
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

// The counts of one tile (kScanItems consecutive cells per thread, one
// 16-byte load where the tile is whole).
__device__ __forceinline__ void tile_counts(const int *count, int ncell, int base,
                                            int (&c)[kScanItems])
{
    static_assert(kScanItems == 4, "one int4 per thread");
    if (base + kScanItems <= ncell) {
        const int4 q = *reinterpret_cast<const int4 *>(count + base);
        c[0] = q.x;
        c[1] = q.y;
        c[2] = q.z;
        c[3] = q.w;
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) c[k] = (base + k < ncell) ? count[base + k] : 0;
    }
}

// Tile totals: bsum[b] = sum of the counts of tile b.
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int *count, int ncell,
                                                              int *bsum)
{
    int c[kScanItems];
    tile_counts(count, ncell, blockIdx.x * kScanTile + threadIdx.x * kScanItems, c);
    int v = c[0] + c[1] + c[2] + c[3];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    __shared__ int wsum[kScanThreads / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += wsum[w];
        bsum[blockIdx.x] = t;
    }
}

// The tile totals replaced by their exclusive prefix sums (one block).
__global__ __launch_bounds__(kScanThreads) void k_scan_bsum(int *bsum, int nb)
{
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int i = b0 + (int)threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int total;
        const int pre = block_exclusive_scan(v, &total);
        if (i < nb) bsum[i] = carry + pre;
        carry += total;
    }
}

// Cell start table from the counts: block b scans its tile on top of the
// prefix of the tile totals before it, writes the starts (also to cstart2
// when set) and clears the counts for the next step's counting.  With
// RESUM, the block sums the raw totals bsum[0..b) itself (few tiles);
// otherwise bsum already holds the prefixes (k_scan_bsum).
template <bool RESUM>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(int *count, int ncell,
                                                             const int *bsum,
                                                             int *cstart, int n,
                                                             int *cstart2)
{
    int before;
    if constexpr (RESUM) {
        int v = 0;
        for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads) v += bsum[b];
        block_exclusive_scan(v, &before);
    } else {
        before = bsum[blockIdx.x];
    }
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int c[kScanItems];
    tile_counts(count, ncell, base, c);
    int total;
    int pre = block_exclusive_scan(c[0] + c[1] + c[2] + c[3], &total) + before;
    if (base + kScanItems <= ncell) {
        const int4 st = make_int4(pre, pre + c[0], pre + c[0] + c[1], pre + c[0] + c[1] + c[2]);
        *reinterpret_cast<int4 *>(cstart + base) = st;
        if (cstart2) *reinterpret_cast<int4 *>(cstart2 + base) = st;
        *reinterpret_cast<int4 *>(count + base) = make_int4(0, 0, 0, 0);
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < ncell) {
                cstart[base + k] = pre;
                if (cstart2) cstart2[base + k] = pre;
                count[base + k] = 0;
            }
            pre += c[k];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cstart[ncell] = n;
        if (cstart2) cstart2[ncell] = n;
    }
}

// The scan of the cell counts into the start table for grids of more than
// kScanDirectTiles tiles: tile totals, then each tile's scan on top of the
// totals before it, summed by each block itself up to kScanResumTiles tiles
// (two launches), beyond that first turned into prefixes by one block
// (three launches; the re-summing grows with the square of the tile count).
constexpr int kScanResumTiles = 256;
inline void launch_scan3(hipStream_t st, int *count, int ncell, int nscan, int *bsum,
                         int *cstart, int n, int *cstart2 = nullptr)
{
    k_scan_reduce<<<nscan, kScanThreads, 0, st>>>(count, ncell, bsum);
    if (nscan <= kScanResumTiles) {
        k_scan_apply<true><<<nscan, kScanThreads, 0, st>>>(count, ncell, bsum, cstart, n,
                                                           cstart2);
    } else {
        k_scan_bsum<<<1, kScanThreads, 0, st>>>(bsum, nscan);
        k_scan_apply<false><<<nscan, kScanThreads, 0, st>>>(count, ncell, bsum, cstart, n,
                                                            cstart2);
    }
}

// Cell start table in one launch for grids of at most kScanDirectTiles
// tiles: each block sums the counts of the tiles before it itself instead of
// reading k_scan_reduce's tile sums.  The counts stay in place (other blocks
// are still reading them); k_scatter clears them.
// (alone it beats k_scan_reduce + k_scan_apply up to about 32 tiles, 7.0 vs
// 9.6 us at 20; inside a step it does not: at 16 tiles, N = 32768 AO with
// 64000 cells, the two launches make the whole step 3.5 us faster, so the
// direct form is kept for the smallest grids only)
constexpr int kScanDirectTiles = 4;
__global__ __launch_bounds__(kScanThreads) void k_scan_direct(const int *count, int ncell,
                                                              int *cstart, int n,
                                                              int *cstart2 = nullptr)
{
    const int lim = blockIdx.x * kScanTile;   // a multiple of 4, at most ncell
    int before = 0;
    // eight independent 16-byte loads in flight per lane (the sum is
    // latency-bound otherwise)
    constexpr int kStride = kScanThreads * 4;
    for (int i0 = threadIdx.x * 4; i0 < lim; i0 += 8 * kStride) {
        int4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = i0 + u * kStride;
            v[u] = i < lim ? *reinterpret_cast<const int4 *>(count + i) : make_int4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) before += v[u].x + v[u].y + v[u].z + v[u].w;
    }
    int before_total;
    block_exclusive_scan(before, &before_total);
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int c[kScanItems];
    int v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        c[k] = (base + k < ncell) ? count[base + k] : 0;
        v += c[k];
    }
    int total;
    int pre = block_exclusive_scan(v, &total) + before_total;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < ncell) {
            cstart[base + k] = pre;
            if (cstart2) cstart2[base + k] = pre;
        }
        pre += c[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cstart[ncell] = n;
        if (cstart2) cstart2[ncell] = n;
    }
}

// With clear_count set (after k_scan_direct), also zeroes the ncell counts
// for the next step's counting.  Each lane places U slots (s, s + blockDim.x,
// ... of its block's U * blockDim.x), their loads issued together.
template <int U = 1>
__global__ __launch_bounds__(256) void k_scatter(int n, const int *key, const int *arank,
                                                 const int *cstart, int *perm,
                                                 int *clear_count, int ncell)
{
    if (clear_count)
        for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
             c += gridDim.x * blockDim.x)
            clear_count[c] = 0;
    const int s0 = blockIdx.x * (blockDim.x * U) + threadIdx.x;
    int k[U], r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int s = s0 + u * (int)blockDim.x;
        k[u] = s < n ? key[s] : 0;
        r[u] = s < n ? arank[s] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int s = s0 + u * (int)blockDim.x;
        if (s < n) perm[cstart[k[u]] + r[u]] = s;
    }
}

// Canonical order within a cell = ascending source slot (stable counting
// sort); gathers the state into the sorted arrays.
__global__ void k_move(int n, const int *__restrict__ perm, const int *__restrict__ key,
                       const int *__restrict__ cstart, const double *__restrict__ npos,
                       const double *__restrict__ npos0, const double *__restrict__ nfrc,
                       const double *__restrict__ nmem, const int *__restrict__ nid,
                       double *__restrict__ pos, double *__restrict__ pos0,
                       double *__restrict__ frc, double *__restrict__ mem,
                       int *__restrict__ id)
{
    // the source and destination arrays are distinct buffers (restrict lets
    // all twelve gathers issue before the first store)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm[t];
    const int c = key[s];
    const int b = cstart[c], e = cstart[c + 1];
    // four independent loads per round (the count is latency-bound otherwise)
    int rank = 0;
    int u = b;
    for (; u + 4 <= e; u += 4) {
        const int p0 = perm[u], p1 = perm[u + 1], p2 = perm[u + 2], p3 = perm[u + 3];
        rank += (p0 < s) + (p1 < s) + (p2 < s) + (p3 < s);
    }
    for (; u < e; ++u) rank += (perm[u] < s);
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


// As k_move, but only the positions and the original ids are gathered into
// the new order; prev[dst] records the slot each particle came from, so the
// integrator reads the previous positions, forces and memory of the last
// step where it wrote them (at prev[s]) instead of every array moving every
// step.
// With ctl set (a step sequence whose control lives in device memory), the
// kernel also ends the step: it advances the clock and the next step's
// flags, which no kernel of this step reads after the integrator.  With
// pos2 / id2 set, the new positions and ids are also written there (a
// record of the state, e.g. the adjoint's recomputed trajectory), and with
// inv2 the slot of each original index.
//
// Each lane moves U particles (slots t, t + blockDim.x, ... of its block's
// U * blockDim.x), their loads interleaved: the kernel is bound by the
// latency of its chain of dependent gathers (slot -> particle -> cell ->
// the cell's members), not by its bytes.
#ifndef PSIM_MOVE_ILP
#define PSIM_MOVE_ILP 4
#endif
constexpr int kMoveIlp = PSIM_MOVE_ILP;
// below this many particles the grid is too small to give up lanes for
// interleaving (one particle per lane)
constexpr int kMoveIlpMin = 1 << 20;
template <int U>
__global__ __launch_bounds__(256) void k_move_lagged(
    int n, const int *__restrict__ perm, const int *__restrict__ key,
    const int *__restrict__ cstart, const double *__restrict__ npos,
    const int *__restrict__ id_in, double *__restrict__ pos, int *__restrict__ id_out,
    int *__restrict__ prev, StepCtl *ctl, double *__restrict__ pos2 = nullptr,
    int *__restrict__ id2 = nullptr, int *__restrict__ inv2 = nullptr)
{
    const int t0 = blockIdx.x * (blockDim.x * U) + threadIdx.x;
    if (ctl && t0 == 0) ctl_advance(ctl);
    if (t0 >= n) return;
    int s[U], b[U], len[U], rank[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = t0 + u * (int)blockDim.x;
        ok[u] = t < n;
        s[u] = perm[ok[u] ? t : t0];   // a missing slot repeats the first
    }
    int most = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int c = key[s[u]];
        b[u] = cstart[c];
        len[u] = cstart[c + 1] - b[u];
        rank[u] = 0;
        most = max(most, len[u]);
    }
    // rank among the cell's members: four independent loads per particle
    // per round
    for (int k = 0; k < most; k += 4) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (k + q < len[u]) rank[u] += (perm[b[u] + k + q] < s[u]);
        }
    }
    double x[U][3];
    int o[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
#pragma unroll
        for (int d = 0; d < 3; ++d) x[u][d] = npos[d * n + s[u]];
        o[u] = id_in[s[u]];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const int dst = b[u] + rank[u];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            pos[d * n + dst] = x[u][d];
            if (pos2) pos2[d * n + dst] = x[u][d];
        }
        id_out[dst] = o[u];
        if (id2) id2[dst] = o[u];
        if (inv2) inv2[o[u]] = dst;
        prev[dst] = s[u];
    }
}

// The lagged arrays (previous positions, last forces, memory, kept in the
// order of the step that wrote them) gathered into the current slot order.
// With pos set, pos0 holds the positions the last step started from and is
// re-imaged against the current positions (util::safeMod0, as the step would
// have stored it).
__global__ void k_settle(int n, const int *__restrict__ prev, const double *__restrict__ pos0,
                         const double *__restrict__ frc, const double *__restrict__ mem,
                         double *__restrict__ pos0_out, double *__restrict__ frc_out,
                         double *__restrict__ mem_out, const double *__restrict__ pos = nullptr,
                         double L = 0.0)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = prev[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double x0 = pos0[d * n + p];
        pos0_out[d * n + s] = pos ? safe_mod0(x0, pos[d * n + s], L) : x0;
        frc_out[d * n + s] = frc[d * n + p];
        mem_out[d * n + s] = mem[d * n + p];
    }
}

}  // namespace
}  // namespace psim
