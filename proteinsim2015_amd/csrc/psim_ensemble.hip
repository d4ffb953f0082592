// psim_ensemble.hip -- many independent systems advanced together.
//
// A replica ensemble stacks R systems of uneven N, box, cell grid and force
// strength in one set of arrays:
//   particles: replica r owns original ids [p0_r, p0_r + n_r); the state is
//     kept in cell order over the whole ensemble (SoA [3][Ntot]);
//   cells: replica r owns global cells [c0_r, c0_r + S_r^3), so one counting
//     sort over the global cell keys rebuilds every replica's table at once
//     and the replicas stay contiguous in slot order.
// Each step is one launch per kernel stage for all replicas (pair forces,
// one per occupancy class, then G&B, scan, scatter, reorder), so a step of
// 512 replicas costs a handful of launches, and the sequence can be captured
// once in a hipGraph and replayed.  The step counter, time and velocity-step
// flags live in device memory (StepCtl) so replays advance them.
//
// These systems hold ~0.1-1.5 particles per cell.  Two pair kernels give the
// same bits: the thread-per-particle gather (k_ens_step, the default) and
// bricks of cells per workgroup (k_ens_step_brick: the launch's workgroups
// are the bricks of every replica, each replica cut by its own density).
// The replicas fall into occupancy classes (particles per cell), and the
// autotune times both families over their tiles (kEnsBlocks workgroup sizes,
// kEnsTargets brick sizes) for each class's replicas and keeps the fastest
// for that class: the pair stage is one launch per class, the integrator one
// launch for all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "psim_build.hpp"
#include "psim_gpu.h"
#include "psim_params.hpp"
#include "psim_step.hpp"

using namespace psim;

namespace {

struct RepGeo {
    Geo g;       // g.n = Ntot (array stride), g.c0 = first cell of the replica
    Pot p;
    double radius;
    BrickGeo bg;   // the replica's bricks (brick kernel)
};

// replica owning original id o: last r with p0[r] <= o
__device__ __forceinline__ int replica_of(const int *__restrict__ p0, int nrep, int o)
{
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p0[mid] <= o) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Thread-per-particle pair forces for the replicas rmap[0..nrep): workgroup
// `it` of the launch covers slots p0[r] + (it - blk0[l]) BLOCK .. of replica
// r = rmap[l] (blk0 = prefix of those replicas' workgroup counts; the
// replicas of one occupancy class, see EnsClass), so a workgroup never spans
// two replicas and
// the replica's geometry and constants are the same for every lane (scalar
// loads, no per-lane replica search).  The forces go to nfrc, where the
// integrator (k_ens_integ, the same launch shape) reads them: as two
// kernels the 5.25 M-particle ensemble step takes 0.927 ms against 0.945
// with the integrator in this kernel's epilogue.
template <int MASK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ens_step(const RepGeo *__restrict__ reps,
                                                    const int *__restrict__ p0,
                                                    const int *__restrict__ blk0,
                                                    const int *__restrict__ rmap, int nrep,
                                                    int nblk, int per_xcd, int ntot, StepArgs a)
{
    __shared__ unsigned hits[kHitCap * BLOCK];
    const int it = xcd_cell(blockIdx.x, per_xcd);
    if (it >= nblk) return;
    const int rl = replica_of(blk0, nrep, it);   // index in the launch's replica list
    const int r = rmap[rl];
    const int s = p0[r] + (it - blk0[rl]) * BLOCK + (int)threadIdx.x;
    if (s >= p0[r + 1]) return;
    // references into the replica table: local copies passed on by
    // reference would live in scratch memory
    const Geo &g = reps[r].g;
    const Pot &p = reps[r].p;
    const double radius = reps[r].radius;
    const int N = ntot;
    const double xi = a.pos[s], yi = a.pos[N + s], zi = a.pos[2 * N + s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{a.pos, N, radius};
    StartTab tab{a.cstart};
    Orig orig{a.id, 1, 0};
    if (g.exact_mi)
        gather_force<MASK, true, BLOCK>(g, p, src, tab, s, xi, yi, zi, radius,
                                        hits + threadIdx.x, fx, fy, fz, a.err, orig);
    else
        gather_force<MASK, false, BLOCK>(g, p, src, tab, s, xi, yi, zi, radius,
                                         hits + threadIdx.x, fx, fy, fz, a.err, orig);
    a.nfrc[s] = fx;
    a.nfrc[N + s] = fy;
    a.nfrc[2 * N + s] = fz;
}

// The G&B step of every particle with the forces k_ens_step left in nfrc.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ens_integ(const RepGeo *__restrict__ reps,
                                                     const int *__restrict__ p0,
                                                     const int *__restrict__ blk0, int nrep,
                                                     int nblk, int per_xcd, int ntot, GB gb,
                                                     StepArgs a)
{
    const int it = xcd_cell(blockIdx.x, per_xcd);
    if (it >= nblk) return;
    const int r = replica_of(blk0, nrep, it);
    const int s = p0[r] + (it - blk0[r]) * BLOCK + (int)threadIdx.x;
    if (s >= p0[r + 1]) return;
    const int N = ntot;
    integrate_one<false>(reps[r].g, gb, a, s, a.nfrc[s], a.nfrc[N + s], a.nfrc[2 * N + s]);
}

// Brick kernel: workgroup `it` of the launch is brick it - item0[l] of
// replica r = rmap[l] (item0 = prefix of the listed replicas' brick counts);
// the forces go to nfrc for k_ens_integ.
template <int MASK>
__global__ __launch_bounds__(kBrickThreads) void k_ens_step_brick(
    const RepGeo *__restrict__ reps, const int *__restrict__ item0,
    const int *__restrict__ rmap, int nrep, int nitems, int per_xcd, int ntot, StepArgs a)
{
    __shared__ BrickShared sh;
    const int it = xcd_cell(blockIdx.x, per_xcd);
    if (it >= nitems) return;
    const int rl = replica_of(item0, nrep, it);
    const int r = rmap[rl];
    // references into the replica table: local copies passed on by
    // reference would live in scratch memory
    const Geo &g = reps[r].g;
    const Pot &p = reps[r].p;
    const double radius = reps[r].radius;
    const BrickGeo &bg = reps[r].bg;
    const int b = it - item0[rl];
    const int N = ntot;
    auto epi = [&](int s, double fx, double fy, double fz) {
        a.nfrc[s] = fx;
        a.nfrc[N + s] = fy;
        a.nfrc[2 * N + s] = fz;
    };
    if (g.exact_mi)
        brick_forces<MASK, true, 1>(g, p, radius, a.pos, a.cstart, a.id, a.err, bg, b, sh, epi);
    else
        brick_forces<MASK, false, 1>(g, p, radius, a.pos, a.cstart, a.id, a.err, bg, b, sh,
                                     epi);
}

// Neighbour-set statistics of the two ensemble pair kernels (their launch
// shapes and acceptance code; see NbrStats).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ens_stats(const RepGeo *__restrict__ reps,
                                                     const int *__restrict__ p0,
                                                     const int *__restrict__ blk0,
                                                     const int *__restrict__ rmap, int nrep,
                                                     int nblk, int per_xcd, int ntot,
                                                     const double *pos, const int *id,
                                                     const int *cstart, NbrStatsOut so)
{
    __shared__ unsigned hits[kHitCap * BLOCK];
    const int it = xcd_cell(blockIdx.x, per_xcd);
    if (it >= nblk) return;
    const int rl = replica_of(blk0, nrep, it);
    const int r = rmap[rl];
    const int s = p0[r] + (it - blk0[rl]) * BLOCK + (int)threadIdx.x;
    if (s >= p0[r + 1]) return;
    const Geo &g = reps[r].g;
    const Pot &p = reps[r].p;
    const double radius = reps[r].radius;
    const int N = ntot;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    SoAPos src{pos, N, radius};
    StartTab tab{cstart};
    Orig orig{id, 1, 0};
    NbrStats st;
    if (g.exact_mi)
        gather_force<F_CAL, true, BLOCK>(g, p, src, tab, s, xi, yi, zi, radius,
                                         hits + threadIdx.x, fx, fy, fz, nullptr, orig, 0, &st);
    else
        gather_force<F_CAL, false, BLOCK>(g, p, src, tab, s, xi, yi, zi, radius,
                                          hits + threadIdx.x, fx, fy, fz, nullptr, orig, 0, &st);
    so.count[id[s]] = st.count;
    so.hash[id[s]] = st.hash;
}

__global__ __launch_bounds__(kBrickThreads) void k_ens_stats_brick(
    const RepGeo *__restrict__ reps, const int *__restrict__ item0,
    const int *__restrict__ rmap, int nrep, int nitems, int per_xcd, const double *pos,
    const int *id, const int *cstart, NbrStatsOut so)
{
    __shared__ BrickShared sh;
    const int it = xcd_cell(blockIdx.x, per_xcd);
    if (it >= nitems) return;
    const int rl = replica_of(item0, nrep, it);
    const int r = rmap[rl];
    const Geo &g = reps[r].g;
    const Pot &p = reps[r].p;
    const double radius = reps[r].radius;
    const BrickGeo &bg = reps[r].bg;
    const int b = it - item0[rl];
    auto epi = [](int, double, double, double) {};
    if (g.exact_mi)
        brick_forces<F_CAL, true, 1, NbrStats>(g, p, radius, pos, cstart, id, nullptr, bg, b, sh,
                                               epi, &so);
    else
        brick_forces<F_CAL, false, 1, NbrStats>(g, p, radius, pos, cstart, id, nullptr, bg, b,
                                                sh, epi, &so);
}

__global__ void k_iota_ens(int n, int *p)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) p[s] = s;
}

// The lagged arrays gathered into the current slot order (k_settle), pos0
// (the positions the last step started from) re-imaged against pos in the
// box of the slot's replica (slots p0[r]..p0[r+1] belong to replica r).
__global__ void k_settle_ens(int n, const RepGeo *__restrict__ reps,
                             const int *__restrict__ p0, int nrep,
                             const int *__restrict__ prev, const double *__restrict__ pos0,
                             const double *__restrict__ frc, const double *__restrict__ mem,
                             const double *__restrict__ pos, double *__restrict__ pos0_out,
                             double *__restrict__ frc_out, double *__restrict__ mem_out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int p = prev[s];
    const double L = reps[replica_of(p0, nrep, s)].g.L;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        pos0_out[d * n + s] = safe_mod0(pos0[d * n + p], pos[d * n + s], L);
        frc_out[d * n + s] = frc[d * n + p];
        mem_out[d * n + s] = mem[d * n + p];
    }
}

// Initial hash of every particle into its replica's cells.
__global__ void k_ens_hash(const RepGeo *__restrict__ reps, const int *__restrict__ p0,
                           int nrep, int ntot, const double *pos, const int *id, int *key,
                           int *arank, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ntot) return;
    const Geo g = reps[replica_of(p0, nrep, id[s])].g;
    const int k = g.c0 + cell_1d(pos[s], g.cell, g.S) + g.S * cell_1d(pos[ntot + s], g.cell, g.S)
                + g.S2 * cell_1d(pos[2 * ntot + s], g.cell, g.S);
    key[s] = k;
    arank[s] = wave_aggregated_rank(count, k);
}

constexpr int kEnsBlocks[3] = {64, 128, 256};
// home particles per brick tried by the autotune (one lane per particle)
constexpr int kEnsTargets[3] = {96, 192, 384};

// Occupancy classes (particles per cell of a replica): [0, 0.25),
// [0.25, 0.5), [0.5, 1), [1, inf).  Each class's replicas get their own pair
// kernel and tile, which the autotune measures class by class.
constexpr int kEnsClasses = 4;
inline int occ_class(double occ) { return occ < 0.25 ? 0 : occ < 0.5 ? 1 : occ < 1.0 ? 2 : 3; }

struct EnsClass {
    std::vector<int> reps;     // replica ids, ascending
    int *rmap = nullptr;       // the same list on the device
    bool brick = false;        // brick kernel (else thread per particle)
    int block = 256;           // thread-per-particle workgroup size
    int target = 192;          // brick kernel: home particles per brick
    int nblk = 0, nitems = 0;  // workgroups / bricks over the class's replicas
    int *blk0 = nullptr, *item0 = nullptr;   // their prefixes over the list
};

}  // namespace

struct psim_ens {
    psim_params common;
    GB gb;
    int nrep = 0, ntot = 0, ncell = 0, nscan = 0;
    int mask = 0;
    bool brick = false;  // initial pair kernel of every class: bricks
    int block = 256;     // initial workgroup size / brick target
    int target = 192;
    // classes 0..kEnsClasses-1 by occupancy; cls[kEnsClasses] holds every
    // replica (one pair launch for all), used when `merged`
    EnsClass cls[kEnsClasses + 1];
    bool merged = false;
    int integ_nblk = 0;        // integrator: 256-lane workgroups over all replicas
    int *integ_blk0 = nullptr; // prefix of the replicas' workgroup counts
    std::vector<RepGeo> h_reps;
    std::vector<int> h_occ_n;
    hipStream_t stream = nullptr;
    RepGeo *reps = nullptr;
    int *p0 = nullptr;
    std::vector<int> h_p0, h_c0;
    std::vector<double> h_box;
    double *pos = nullptr, *pos0 = nullptr, *frc = nullptr, *mem = nullptr;
    double *npos = nullptr, *npos0 = nullptr, *nfrc = nullptr, *nmem = nullptr;
    double *vel = nullptr;
    int *id = nullptr, *nid = nullptr, *key = nullptr, *arank = nullptr, *perm = nullptr;
    int *count = nullptr, *cstart = nullptr, *bsum = nullptr;
    // pos0 / frc / mem stay in the slot order of the last step (pos0: the
    // positions it started from, re-imaged when read); prev[s] is the slot
    // in that order of the particle now in slot s
    int *prev = nullptr;
    bool lagged = false;
    DevErr *err = nullptr;
    StepCtl *ctl = nullptr;
    int *st_count = nullptr;                 // neighbour-set statistics
    unsigned long long *st_hash = nullptr;
    bool have_state = false;
    int graph_steps = 0;
    // captured step sequences, one per arrangement of the swapped buffers
    // they start from (an odd number of steps alternates between two)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    const void *gkey[2] = {nullptr, nullptr};   // pos0 buffer at the capture's start
};

#define ECK(x)                                                                   \
    do {                                                                         \
        hipError_t _e = (x);                                                     \
        if (_e != hipSuccess) {                                                  \
            std::fprintf(stderr, "psim_ens: %s failed: %s (%s:%d)\n", #x,        \
                         hipGetErrorString(_e), __FILE__, __LINE__);             \
            return PSIM_E_HIP;                                                   \
        }                                                                        \
    } while (0)

namespace {

template <class T>
int ealloc(T **p, size_t count)
{
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, count * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return PSIM_E_NOMEM;
    }
    return 0;
}

inline unsigned eblocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

StepArgs ens_args(psim_ens *e)
{
    StepArgs a{};
    a.pos = e->pos; a.pos0 = e->pos0; a.frc = e->frc; a.mem = e->mem;
    a.id = e->id; a.cstart = e->cstart;
    a.prev = e->lagged ? e->prev : nullptr;
    a.x0_raw = e->lagged ? 1 : 0;
    a.npos = e->npos; a.npos0 = e->npos0; a.nfrc = e->nfrc; a.nmem = e->nmem;
    a.nid = nullptr; a.vel = e->vel; a.key = e->key; a.arank = e->arank;
    a.count = e->count; a.err = e->err;
    a.noise = nullptr;
    a.seed = e->common.seed;
    a.ctl = e->ctl;
    return a;
}

template <int MASK, int BLOCK>
void launch_ens_pair(psim_ens *e, const EnsClass &cl, const StepArgs &a)
{
    const int per = (cl.nblk + 7) / 8;
    k_ens_step<MASK, BLOCK><<<8 * per, BLOCK, 0, e->stream>>>(
        e->reps, e->p0, cl.blk0, cl.rmap, (int)cl.reps.size(), cl.nblk, per, e->ntot, a);
}

// pair forces of one occupancy class's replicas into nfrc
template <int MASK>
void launch_ens_pair_class(psim_ens *e, const EnsClass &cl, const StepArgs &a)
{
    if (cl.reps.empty()) return;
    if (cl.brick) {
        const int per = (cl.nitems + 7) / 8;
        k_ens_step_brick<MASK><<<8 * per, kBrickThreads, 0, e->stream>>>(
            e->reps, cl.item0, cl.rmap, (int)cl.reps.size(), cl.nitems, per, e->ntot, a);
        return;
    }
    switch (cl.block) {
    case 64: launch_ens_pair<MASK, 64>(e, cl, a); break;
    case 128: launch_ens_pair<MASK, 128>(e, cl, a); break;
    default: launch_ens_pair<MASK, 256>(e, cl, a); break;
    }
}

// pair forces of every class, then the G&B step of every particle
template <int MASK>
void launch_ens_pair_block(psim_ens *e, const StepArgs &a)
{
    if (e->merged)
        launch_ens_pair_class<MASK>(e, e->cls[kEnsClasses], a);
    else
        for (int c = 0; c < kEnsClasses; ++c) launch_ens_pair_class<MASK>(e, e->cls[c], a);
    const int per = (e->integ_nblk + 7) / 8;
    k_ens_integ<256><<<8 * per, 256, 0, e->stream>>>(e->reps, e->p0, e->integ_blk0, e->nrep,
                                                     e->integ_nblk, per, e->ntot, e->gb, a);
}

// the buffer roles after one step's cell rebuild
void swap_lagged(psim_ens *e)
{
    std::swap(e->id, e->nid);
    std::swap(e->pos, e->pos0);
    std::swap(e->frc, e->nfrc);
    std::swap(e->mem, e->nmem);
}

void launch_build_ens(psim_ens *e, StepCtl *advance)
{
    const int n = e->ntot, nc = e->ncell;
    hipStream_t st = e->stream;
    const bool direct = e->nscan <= kScanDirectTiles;
    if (direct) {
        k_scan_direct<<<e->nscan, kScanThreads, 0, st>>>(e->count, nc, e->cstart, n);
    } else {
        launch_scan3(st, e->count, nc, e->nscan, e->bsum, e->cstart, n);
    }
    if (n >= kMoveIlpMin)
        k_scatter<kMoveIlp><<<eblocks(n, 256 * kMoveIlp), 256, 0, st>>>(
            n, e->key, e->arank, e->cstart, e->perm, direct ? e->count : nullptr, nc);
    else
        k_scatter<1><<<eblocks(n, 256), 256, 0, st>>>(n, e->key, e->arank, e->cstart, e->perm,
                                                      direct ? e->count : nullptr, nc);
    // positions and ids move into the new order, the positions into the
    // buffer of the last step's start positions (read by this step's
    // integrator, not after it); the positions this step started from become
    // pos0, and they, the forces and the memory stay where this step had or
    // wrote them: the next step reads them through prev
    if (n >= kMoveIlpMin)
        k_move_lagged<kMoveIlp><<<eblocks(n, 256 * kMoveIlp), 256, 0, st>>>(
            n, e->perm, e->key, e->cstart, e->npos, e->id, e->pos0, e->nid, e->prev, advance);
    else
        k_move_lagged<1><<<eblocks(n, 256), 256, 0, st>>>(
            n, e->perm, e->key, e->cstart, e->npos, e->id, e->pos0, e->nid, e->prev, advance);
    swap_lagged(e);
    e->lagged = true;
}

// The lagged arrays gathered into the current slot order, copied back so
// that the buffers stay where a captured step sequence expects them.
int settle_ens(psim_ens *e)
{
    if (!e->lagged) return 0;
    const int n = e->ntot;
    k_settle_ens<<<eblocks(n, 256), 256, 0, e->stream>>>(n, e->reps, e->p0, e->nrep, e->prev,
                                                         e->pos0, e->frc, e->mem, e->pos,
                                                         e->npos0, e->nfrc, e->nmem);
    ECK(hipGetLastError());
    const size_t b = sizeof(double) * 3 * (size_t)n;
    ECK(hipMemcpyAsync(e->pos0, e->npos0, b, hipMemcpyDeviceToDevice, e->stream));
    ECK(hipMemcpyAsync(e->frc, e->nfrc, b, hipMemcpyDeviceToDevice, e->stream));
    ECK(hipMemcpyAsync(e->mem, e->nmem, b, hipMemcpyDeviceToDevice, e->stream));
    e->lagged = false;
    return 0;
}

// Workgroups of `block` lanes per replica for the thread-per-particle
// kernel, and their prefix.
// Thread-per-particle workgroups of `block` lanes for the replicas of class c.
int set_blocks(psim_ens *e, int c, int block)
{
    EnsClass &cl = e->cls[c];
    const int m = (int)cl.reps.size();
    std::vector<int> b0(m + 1);
    int nb = 0;
    for (int l = 0; l < m; ++l) {
        b0[l] = nb;
        nb += (e->h_occ_n[cl.reps[l]] + block - 1) / block;
    }
    b0[m] = nb;
    if (m) ECK(hipMemcpy(cl.blk0, b0.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice));
    cl.brick = false;
    cl.block = block;
    cl.nblk = nb;
    return 0;
}

// The integrator's 256-lane workgroups over all replicas.
int set_integ_blocks(psim_ens *e)
{
    std::vector<int> b0(e->nrep + 1);
    int nb = 0;
    for (int r = 0; r < e->nrep; ++r) {
        b0[r] = nb;
        nb += (e->h_occ_n[r] + 255) / 256;
    }
    b0[e->nrep] = nb;
    ECK(hipMemcpy(e->integ_blk0, b0.data(), sizeof(int) * (e->nrep + 1),
                  hipMemcpyHostToDevice));
    e->integ_nblk = nb;
    return 0;
}

// The pair kernel of one class (forces into nfrc only).
int launch_class(psim_ens *e, const EnsClass &cl, const StepArgs &a)
{
    switch (e->mask) {
    case 1: launch_ens_pair_class<1>(e, cl, a); break;
    case 2: launch_ens_pair_class<2>(e, cl, a); break;
    case 3: launch_ens_pair_class<3>(e, cl, a); break;
    case 4: launch_ens_pair_class<4>(e, cl, a); break;
    case 5: launch_ens_pair_class<5>(e, cl, a); break;
    case 6: launch_ens_pair_class<6>(e, cl, a); break;
    case 7: launch_ens_pair_class<7>(e, cl, a); break;
    default: return PSIM_E_ARG;
    }
    return 0;
}

// One full step, enqueued on the ensemble's stream (also used while
// capturing the graph).
int enqueue_step(psim_ens *e)
{
    StepArgs a = ens_args(e);
    switch (e->mask) {
    case 1: launch_ens_pair_block<1>(e, a); break;
    case 2: launch_ens_pair_block<2>(e, a); break;
    case 3: launch_ens_pair_block<3>(e, a); break;
    case 4: launch_ens_pair_block<4>(e, a); break;
    case 5: launch_ens_pair_block<5>(e, a); break;
    case 6: launch_ens_pair_block<6>(e, a); break;
    case 7: launch_ens_pair_block<7>(e, a); break;
    default: return PSIM_E_ARG;
    }
    launch_build_ens(e, e->ctl);   // its last kernel also advances the step control
    ECK(hipGetLastError());
    return 0;
}

void drop_graph(psim_ens *e)
{
    for (int k = 0; k < 2; ++k) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
        e->exec[k] = nullptr;
        e->graph[k] = nullptr;
        e->gkey[k] = nullptr;
    }
}

// Cut every replica of class c into bricks of about `target` home particles
// and upload the replica table and the class's brick prefix.
int set_bricks(psim_ens *e, int c, int target)
{
    EnsClass &cl = e->cls[c];
    const int m = (int)cl.reps.size();
    std::vector<int> it0(m + 1);
    int nitems = 0;
    for (int l = 0; l < m; ++l) {
        const int r = cl.reps[l];
        RepGeo &rg = e->h_reps[r];
        const double occ = (double)e->h_occ_n[r] / rg.g.ncell;
        rg.bg = brick_shape(rg.g.S, occ, (double)target);
        it0[l] = nitems;
        nitems += rg.bg.nbricks;
    }
    it0[m] = nitems;
    ECK(hipMemcpy(e->reps, e->h_reps.data(), sizeof(RepGeo) * e->nrep,
                  hipMemcpyHostToDevice));
    if (m) ECK(hipMemcpy(cl.item0, it0.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice));
    cl.brick = true;
    cl.target = target;
    cl.nitems = nitems;
    return 0;
}

// The captured sequence that starts from the current arrangement of the
// swapped buffers (captured on first use).  Capturing enqueues nothing but
// advances the host's buffer bookkeeping as the replay will.
int graph_for_now(psim_ens *e, hipGraphExec_t *out)
{
    if (!e->lagged) {
        k_iota_ens<<<eblocks(e->ntot, 256), 256, 0, e->stream>>>(e->ntot, e->prev);
        ECK(hipGetLastError());
        e->lagged = true;
    }
    const void *key = e->pos0;
    for (int k = 0; k < 2; ++k)
        if (e->exec[k] && e->gkey[k] == key) {
            *out = e->exec[k];
            // the buffers the replay leaves behind
            for (int q = 0; q < e->graph_steps; ++q) swap_lagged(e);
            return 0;
        }
    const int slot = e->exec[0] ? 1 : 0;
    ECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < e->graph_steps && !rc; ++k) rc = enqueue_step(e);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) {
        std::fprintf(stderr, "psim_ens: graph capture failed: %s\n", hipGetErrorString(ce));
        return PSIM_E_HIP;
    }
    e->graph[slot] = g;
    ECK(hipGraphInstantiate(&e->exec[slot], e->graph[slot], nullptr, nullptr, 0));
    e->gkey[slot] = key;
    *out = e->exec[slot];
    return 0;
}

}  // namespace

extern "C" {

int psim_ens_create(const psim_params *common, const psim_replica *reps, int nrep,
                    psim_ens **out)
{
    if (!common || !reps || nrep <= 0 || !out) return PSIM_E_ARG;
    *out = nullptr;
    const psim_params &C = *common;
    if (C.force_mask <= 0 || C.force_mask > 7 || !(C.dt > 0) || !(C.mass > 0)) return PSIM_E_ARG;
    psim_ens *e = new (std::nothrow) psim_ens();
    if (!e) return PSIM_E_NOMEM;
    e->common = C;
    e->mask = C.force_mask;
    e->nrep = nrep;
    e->brick = C.kernel_path == 3;
    std::vector<RepGeo> h(nrep);
    long long ntot = 0, ncell = 0;
    e->h_p0.resize(nrep + 1);
    e->h_c0.resize(nrep + 1);
    e->h_box.resize(nrep);
    for (int r = 0; r < nrep; ++r) {
        const psim_replica &R = reps[r];
        if (R.n_particles <= 0 || R.cells_per_dim < 3 || R.cells_per_dim > 1290 ||
            !(R.box > 0) || !(R.cell_size > 0) || !(R.cutoff > 0)) {
            delete e;
            return PSIM_E_ARG;
        }
        psim_params P = C;
        P.n_particles = R.n_particles;
        P.cells_per_dim = R.cells_per_dim;
        P.box = R.box;
        P.cell_size = R.cell_size;
        P.half_uv = 0.0;
        P.cutoff = R.cutoff;
        P.lj_num = R.lj_num;
        P.lj_well_depth = R.lj_well_depth;
        P.yukawa_strength = R.yukawa_strength;
        P.debye_length = R.debye_length > 0 ? R.debye_length : 0.5;
        P.ao_well_depth = R.ao_well_depth;
        Geo g;
        Pot p;
        GB b;
        make_constants(P, g, p, b);
        if (r == 0) e->gb = b;
        g.c0 = (int)ncell;
        h[r].g = g;
        h[r].p = p;
        h[r].radius = C.radius;
        h[r].bg = BrickGeo{};
        e->h_p0[r] = (int)ntot;
        e->h_c0[r] = (int)ncell;
        e->h_box[r] = R.box;
        ntot += R.n_particles;
        ncell += (long long)g.ncell;
    }
    e->h_p0[nrep] = (int)ntot;
    e->h_c0[nrep] = (int)ncell;
    if (ntot >= (1 << kIdxBits) || ncell >= (1ll << 31) - 1) {
        delete e;
        return PSIM_E_UNSUPPORTED;
    }
    e->ntot = (int)ntot;
    e->ncell = (int)ncell;
    for (auto &rg : h) rg.g.n = e->ntot;   // SoA stride of the stacked arrays
    e->nscan = (e->ncell + kScanTile - 1) / kScanTile;
    const size_t n3 = 3 * (size_t)e->ntot;
    int rc = 0;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = PSIM_E_HIP;
    double **dbufs[] = {&e->pos, &e->pos0, &e->frc, &e->mem, &e->npos, &e->npos0, &e->nfrc,
                        &e->nmem, &e->vel};
    for (double **b : dbufs)
        if (!rc) rc = ealloc(b, n3);
    int **ibufs[] = {&e->id, &e->nid, &e->key, &e->arank, &e->perm, &e->prev};
    for (int **b : ibufs)
        if (!rc) rc = ealloc(b, (size_t)e->ntot);
    if (!rc) rc = ealloc(&e->count, (size_t)e->ncell);
    if (!rc) rc = ealloc(&e->cstart, (size_t)e->ncell + 1);
    if (!rc) rc = ealloc(&e->bsum, (size_t)e->nscan);
    if (!rc) rc = ealloc(&e->err, 1);
    if (!rc) rc = ealloc(&e->ctl, 1);
    if (!rc) rc = ealloc(&e->reps, (size_t)nrep);
    if (!rc) rc = ealloc(&e->p0, (size_t)nrep + 1);
    if (!rc) rc = ealloc(&e->integ_blk0, (size_t)nrep + 1);
    e->h_reps = h;
    e->h_occ_n.resize(nrep);
    for (int r = 0; r < nrep; ++r) {
        e->h_occ_n[r] = reps[r].n_particles;
        e->cls[occ_class((double)reps[r].n_particles / h[r].g.ncell)].reps.push_back(r);
        e->cls[kEnsClasses].reps.push_back(r);
    }
    for (int c = 0; c <= kEnsClasses && !rc; ++c) {
        EnsClass &cl = e->cls[c];
        if (cl.reps.empty()) continue;
        const size_t m = cl.reps.size();
        rc = ealloc(&cl.rmap, m);
        if (!rc) rc = ealloc(&cl.blk0, m + 1);
        if (!rc) rc = ealloc(&cl.item0, m + 1);
        if (!rc && hipMemcpy(cl.rmap, cl.reps.data(), sizeof(int) * m, hipMemcpyHostToDevice) !=
                       hipSuccess)
            rc = PSIM_E_HIP;
        if (!rc) rc = set_blocks(e, c, e->block);
        if (!rc) rc = set_bricks(e, c, e->target);   // the replica table with its bricks
        if (!rc && !e->brick) rc = set_blocks(e, c, e->block);
    }
    if (!rc) rc = set_integ_blocks(e);
    if (!rc && hipMemcpy(e->p0, e->h_p0.data(), sizeof(int) * (nrep + 1), hipMemcpyHostToDevice) != hipSuccess)
        rc = PSIM_E_HIP;
    if (!rc && hipMemset(e->count, 0, sizeof(int) * e->ncell) != hipSuccess) rc = PSIM_E_HIP;
    if (!rc && hipMemset(e->err, 0, sizeof(DevErr)) != hipSuccess) rc = PSIM_E_HIP;
    if (rc) {
        psim_ens_destroy(e);
        return rc;
    }
    *out = e;
    return 0;
}

void psim_ens_destroy(psim_ens *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    void *bufs[] = {e->pos, e->pos0, e->frc, e->mem, e->npos, e->npos0, e->nfrc, e->nmem,
                    e->vel, e->id, e->nid, e->key, e->arank, e->perm, e->count, e->cstart,
                    e->bsum, e->err, e->ctl, e->reps, e->p0, e->integ_blk0, e->st_count,
                    e->st_hash, e->prev};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (EnsClass &cl : e->cls) {   // the occupancy classes and the all-replica one
        void *cb[] = {cl.rmap, cl.blk0, cl.item0};
        for (void *b : cb)
            if (b) (void)hipFree(b);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int psim_ens_total_particles(psim_ens *e) { return e ? e->ntot : PSIM_E_ARG; }

int psim_ens_set_state(psim_ens *e, const double *pos, const double *vel, double current_time)
{
    if (!e || !pos) return PSIM_E_ARG;
    const int n = e->ntot;
    std::vector<double> h(3 * (size_t)n);
    auto up = [&](const double *src, double *dst, bool pos0) -> int {
        int r = 0;
        for (int i = 0; i < n; ++i) {
            while (i >= e->h_p0[r + 1]) ++r;
            const double L = e->h_box[r];
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
                if (src) {
                    v = src[3 * (size_t)i + d];
                } else if (pos0) {
                    // a fresh particle: pos0 = safeMod0(0, x) (particle.cpp:31-66)
                    const double x = pos[3 * (size_t)i + d];
                    v = (std::fabs(x) > (L / 2)) ? ((x < 0) ? -L : L) : 0.0;
                }
                h[(size_t)d * n + i] = v;
            }
        }
        ECK(hipMemcpy(dst, h.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
        return 0;
    };
    int rc;
    if ((rc = up(pos, e->npos, false))) return rc;
    if ((rc = up(nullptr, e->npos0, true))) return rc;
    if ((rc = up(vel, e->vel, false))) return rc;
    if ((rc = up(nullptr, e->nfrc, false))) return rc;
    if ((rc = up(nullptr, e->nmem, false))) return rc;
    std::vector<int> ids(n);
    for (int i = 0; i < n; ++i) ids[i] = i;
    ECK(hipMemcpy(e->id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    ECK(hipMemset(e->count, 0, sizeof(int) * e->ncell));
    StepCtl c{};
    c.step = 0;
    c.time = current_time;
    c.dt = e->common.dt;
    c.vel_counter = 0;
    c.vel_freq = e->common.vel_freq;
    c.first = (current_time == 0) ? 1 : 0;
    c.do_vel = 0;
    if (!c.first) {
        // as the host-driven engine does at the top of a step
        c.do_vel = (c.vel_freq == 0 || c.vel_counter == c.vel_freq) ? 1 : 0;
        c.vel_counter = (c.vel_counter == c.vel_freq) ? 0 : c.vel_counter + 1;
    }
    ECK(hipMemcpy(e->ctl, &c, sizeof(c), hipMemcpyHostToDevice));
    k_ens_hash<<<eblocks(n, 256), 256, 0, e->stream>>>(e->reps, e->p0, e->nrep, n, e->npos,
                                                       e->id, e->key, e->arank, e->count);
    launch_build_ens(e, nullptr);
    ECK(hipGetLastError());
    // the fresh previous positions (already re-imaged, upload order), forces
    // and memory gathered into the slot order; the state is then settled
    k_settle<<<eblocks(n, 256), 256, 0, e->stream>>>(n, e->prev, e->npos0, e->frc, e->mem,
                                                     e->pos0, e->nfrc, e->nmem);
    ECK(hipGetLastError());
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    ECK(hipMemcpyAsync(e->frc, e->nfrc, b3, hipMemcpyDeviceToDevice, e->stream));
    ECK(hipMemcpyAsync(e->mem, e->nmem, b3, hipMemcpyDeviceToDevice, e->stream));
    e->lagged = false;
    ECK(hipStreamSynchronize(e->stream));
    e->have_state = true;
    return 0;
}

int psim_ens_get_state(psim_ens *e, double *pos, double *pos0, double *vel, double *frc)
{
    if (!e || !e->have_state) return PSIM_E_ARG;
    const int n = e->ntot;
    int rs = settle_ens(e);
    if (rs) return rs;
    std::vector<int> ids(n);
    std::vector<double> h(3 * (size_t)n);
    ECK(hipStreamSynchronize(e->stream));
    ECK(hipMemcpy(ids.data(), e->id, sizeof(int) * n, hipMemcpyDeviceToHost));
    auto down = [&](const double *src, double *dst, bool by_slot) -> int {
        if (!dst) return 0;
        ECK(hipMemcpy(h.data(), src, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        for (int s = 0; s < n; ++s) {
            const int o = by_slot ? ids[s] : s;
            for (int d = 0; d < 3; ++d) dst[3 * (size_t)o + d] = h[(size_t)d * n + s];
        }
        return 0;
    };
    int rc;
    if ((rc = down(e->pos, pos, true))) return rc;
    if ((rc = down(e->pos0, pos0, true))) return rc;
    if ((rc = down(e->vel, vel, false))) return rc;
    if ((rc = down(e->frc, frc, true))) return rc;
    return 0;
}

int psim_ens_set_graph(psim_ens *e, int steps_per_graph)
{
    if (!e || steps_per_graph < 0) return PSIM_E_ARG;
    drop_graph(e);
    e->graph_steps = steps_per_graph;
    return 0;
}

int psim_ens_step(psim_ens *e, int nsteps)
{
    if (!e || nsteps < 0 || !e->have_state) return PSIM_E_ARG;
    int done = 0;
    if (e->graph_steps > 0) {
        for (; done + e->graph_steps <= nsteps; done += e->graph_steps) {
            hipGraphExec_t x = nullptr;
            int rc = graph_for_now(e, &x);
            if (rc) return rc;
            ECK(hipGraphLaunch(x, e->stream));
        }
    }
    for (; done < nsteps; ++done) {
        int rc = enqueue_step(e);
        if (rc) return rc;
    }
    return 0;
}

int psim_ens_sync(psim_ens *e)
{
    if (!e) return PSIM_E_ARG;
    ECK(hipStreamSynchronize(e->stream));
    return 0;
}

int psim_ens_last_error(psim_ens *e, int *code, int *i, int *j)
{
    if (!e) return PSIM_E_ARG;
    DevErr h{};
    ECK(hipStreamSynchronize(e->stream));
    ECK(hipMemcpy(&h, e->err, sizeof(h), hipMemcpyDeviceToHost));
    if (code) *code = h.code;
    if (i) *i = h.i;
    if (j) *j = h.j;
    return 0;
}

int psim_ens_pair_stats(psim_ens *e, int *count_by_orig, unsigned long long *hash_by_orig)
{
    if (!e || !e->have_state || !count_by_orig || !hash_by_orig) return PSIM_E_ARG;
    const int n = e->ntot;
    int rc = 0;
    if (!e->st_count) {
        if (!rc) rc = ealloc(&e->st_count, (size_t)n);
        if (!rc) rc = ealloc(&e->st_hash, (size_t)n);
        if (rc) return rc;
    }
    const NbrStatsOut so{e->st_count, e->st_hash};
    for (int c = 0; c <= kEnsClasses; ++c) {
        if ((c == kEnsClasses) != e->merged) continue;
        const EnsClass &cl = e->cls[c];
        if (cl.reps.empty()) continue;
        const int m = (int)cl.reps.size();
        if (cl.brick) {
            const int per = (cl.nitems + 7) / 8;
            k_ens_stats_brick<<<8 * per, kBrickThreads, 0, e->stream>>>(
                e->reps, cl.item0, cl.rmap, m, cl.nitems, per, e->pos, e->id, e->cstart, so);
            continue;
        }
        const int per = (cl.nblk + 7) / 8;
#define PSIM_ENS_STATS(BL)                                                                 \
    k_ens_stats<BL><<<8 * per, BL, 0, e->stream>>>(e->reps, e->p0, cl.blk0, cl.rmap, m,    \
                                                   cl.nblk, per, n, e->pos, e->id,         \
                                                   e->cstart, so)
        switch (cl.block) {
        case 64: PSIM_ENS_STATS(64); break;
        case 128: PSIM_ENS_STATS(128); break;
        default: PSIM_ENS_STATS(256); break;
        }
#undef PSIM_ENS_STATS
    }
    ECK(hipGetLastError());
    ECK(hipMemcpyAsync(count_by_orig, e->st_count, sizeof(int) * n, hipMemcpyDeviceToHost,
                       e->stream));
    ECK(hipMemcpyAsync(hash_by_orig, e->st_hash, sizeof(unsigned long long) * n,
                       hipMemcpyDeviceToHost, e->stream));
    ECK(hipStreamSynchronize(e->stream));
    return 0;
}

// For each occupancy class, and for all replicas as one launch: time every
// candidate pair kernel (three workgroup sizes of the thread-per-particle
// kernel, three brick sizes of the brick kernel) on those replicas alone,
// trial_steps launches after one warm-up (the pair kernels only write forces,
// so they repeat on one state), and keep the fastest.  Then 1 + trial_steps
// whole steps with one launch per class and as many with one launch for all
// decide between the two (the separate launches' tails can cost more than
// the per-class tiles save); *best_ms is the chosen one's.  *kernel and
// *tile report the choice for the class holding the most particles
// (psim_ens_class_info gives every class's).  The whole steps advance the
// ensemble like any other steps; no choice changes a result.
int psim_ens_autotune(psim_ens *e, int trial_steps, int *chosen_kernel, int *chosen_tile,
                      float *best_ms)
{
    if (!e || trial_steps <= 0 || !e->have_state) return PSIM_E_ARG;
    hipEvent_t t0, t1;
    ECK(hipEventCreate(&t0));
    ECK(hipEventCreate(&t1));
    int rc = 0;
    int big = -1;
    long long big_n = -1;
    const StepArgs a = ens_args(e);
    // each class's candidates on its own, and the same for all replicas in one
    // launch (class kEnsClasses)
    for (int c = 0; c <= kEnsClasses && !rc; ++c) {
        EnsClass &cl = e->cls[c];
        if (cl.reps.empty()) continue;
        float best = 0.f;
        bool best_brick = cl.brick;
        int best_tile = cl.brick ? cl.target : cl.block;
        for (int ci = 0; ci < 6 && !rc; ++ci) {
            const bool brick = ci >= 3;
            const int b = brick ? kEnsTargets[ci - 3] : kEnsBlocks[ci];
            rc = brick ? set_bricks(e, c, b) : set_blocks(e, c, b);
            if (!rc) rc = launch_class(e, cl, a);
            if (rc) break;
            ECK(hipEventRecord(t0, e->stream));
            for (int k = 0; k < trial_steps && !rc; ++k) rc = launch_class(e, cl, a);
            ECK(hipEventRecord(t1, e->stream));
            ECK(hipEventSynchronize(t1));
            ECK(hipGetLastError());
            float ms = 0.f;
            ECK(hipEventElapsedTime(&ms, t0, t1));
            if (best == 0.f || ms < best) {
                best = ms;
                best_tile = b;
                best_brick = brick;
            }
        }
        cl.brick = best_brick;
        (best_brick ? cl.target : cl.block) = best_tile;
        if (c == kEnsClasses) continue;
        long long nc = 0;
        for (int r : cl.reps) nc += e->h_occ_n[r];
        if (nc > big_n) {
            big_n = nc;
            big = c;
        }
    }
    // the replica table holds one brick cut per replica: (re)make the one of
    // the configuration about to run
    auto apply = [&](bool merged) {
        e->merged = merged;
        int r2 = 0;
        for (int c = 0; c <= kEnsClasses && !r2; ++c) {
            EnsClass &cl = e->cls[c];
            if (cl.reps.empty() || (c == kEnsClasses) != merged) continue;
            r2 = cl.brick ? set_bricks(e, c, cl.target) : set_blocks(e, c, cl.block);
        }
        return r2;
    };
    // whole steps per class and in one launch; keep the faster
    float ms_cls = 0.f, ms_all = 0.f;
    for (int m = 0; m < 2 && !rc; ++m) {
        rc = apply(m == 1);
        if (!rc) rc = enqueue_step(e);
        if (rc) break;
        ECK(hipEventRecord(t0, e->stream));
        for (int k = 0; k < trial_steps && !rc; ++k) rc = enqueue_step(e);
        ECK(hipEventRecord(t1, e->stream));
        ECK(hipEventSynchronize(t1));
        ECK(hipEventElapsedTime(m ? &ms_all : &ms_cls, t0, t1));
    }
    if (!rc) rc = apply(ms_all < ms_cls);
    drop_graph(e);   // recaptured with the chosen kernels on the next step
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc) return rc;
    const EnsClass &cb = e->cls[e->merged ? kEnsClasses : big];
    if (chosen_kernel) *chosen_kernel = cb.brick ? 3 : 1;
    if (chosen_tile) *chosen_tile = cb.brick ? cb.target : cb.block;
    if (best_ms) *best_ms = (e->merged ? ms_all : ms_cls) / trial_steps;
    return 0;
}

int psim_ens_class_info(psim_ens *e, int cls, int *nrep, int *kernel, int *tile)
{
    if (!e || cls < 0 || cls >= kEnsClasses) return PSIM_E_ARG;
    // the kernel that runs the class's replicas: its own, or the one launch
    // for all replicas when the autotune found that faster
    const EnsClass &cl = e->cls[e->merged ? kEnsClasses : cls];
    if (nrep) *nrep = (int)e->cls[cls].reps.size();
    if (kernel) *kernel = cl.brick ? 3 : 1;
    if (tile) *tile = cl.brick ? cl.target : cl.block;
    return 0;
}

int psim_ens_merged(psim_ens *e) { return e ? (e->merged ? 1 : 0) : PSIM_E_ARG; }

}  // extern "C"
