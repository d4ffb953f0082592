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
// One step (the order of system::run, system.cpp:187-211):
//   k_step      fused: 27-cell forces F(x_n) (IForce::getAcceleration),
//               push (frc0 <- frc), G&B integration (firstStep/normalStep/
//               velocityStep), wrap (particle::setPos), new cell hash and
//               an atomic per-cell count; writes the next state in the
//               current slot order to the "n*" buffers.
//   k_scan_*    counts -> cstart (exclusive scan), counts cleared: tile
//               totals, their prefix (one block), then each tile's scan;
//               one launch, k_scan_direct, for grids of up to 32 scan
//               tiles (k_scatter then clears the counts).
//   k_scatter   slot of each particle in its new cell (arbitrary order).
//   k_move_lagged  canonical within-cell order (ascending previous slot,
//               i.e. a stable counting sort); positions and ids move into
//               it; the positions the step started from (pos0 of the next
//               step, re-imaged when read), the forces and the memory stay
//               where the step had or wrote them (read through the slot map
//               prev).
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "psim_ctx.hpp"
#include "psim_build.hpp"
#include "psim_params.hpp"
#include "psim_step.hpp"

using namespace psim;


namespace {

using eng::blocks_for;
using eng::dalloc;

void fill_constants(psim_ctx *c) { make_constants(c->prm, c->geo, c->pot, c->gb); }

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
int launch_build(psim_ctx *c, StepCtl *advance)
{
    const int n = c->n, nc = c->ncell;
    hipStream_t st = c->stream;
    const bool direct = c->nscan <= kScanDirectTiles;
    if (direct) {
        k_scan_direct<<<c->nscan, kScanThreads, 0, st>>>(c->count, nc, c->cstart, n,
                                                         c->rec.cstart);
    } else {
        launch_scan3(st, c->count, nc, c->nscan, c->bsum, c->cstart, n, c->rec.cstart);
    }
    if (n >= kMoveIlpMin)
        k_scatter<kMoveIlp><<<blocks_for(n, 256 * kMoveIlp), 256, 0, st>>>(
            n, c->key, c->arank, c->cstart, c->perm, direct ? c->count : nullptr, nc);
    else
        k_scatter<1><<<blocks_for(n, 256), 256, 0, st>>>(n, c->key, c->arank, c->cstart,
                                                         c->perm, direct ? c->count : nullptr,
                                                         nc);
    // positions and ids move into the new order, the positions into the
    // buffer of the last step's start positions (read by this step's
    // integrator, not after it); the positions this step started from become
    // pos0, and they, the forces and the memory stay where this step had or
    // wrote them (old order): the next step reads them through prev
    if (n >= kMoveIlpMin)
        k_move_lagged<kMoveIlp><<<blocks_for(n, 256 * kMoveIlp), 256, 0, st>>>(
            n, c->perm, c->key, c->cstart, c->npos, c->id, c->pos0, c->nid, c->prev, advance,
            c->rec.pos, c->rec.id, c->rec.inv);
    else
        k_move_lagged<1><<<blocks_for(n, 256), 256, 0, st>>>(
            n, c->perm, c->key, c->cstart, c->npos, c->id, c->pos0, c->nid, c->prev, advance,
            c->rec.pos, c->rec.id, c->rec.inv);
    CK(hipGetLastError());
    std::swap(c->id, c->nid);
    std::swap(c->pos, c->pos0);
    std::swap(c->frc, c->nfrc);
    std::swap(c->mem, c->nmem);
    c->lagged = true;
    return 0;
}

int settle(psim_ctx *c)
{
    if (!c->lagged) return 0;
    k_settle<<<blocks_for(c->n, 256), 256, 0, c->stream>>>(c->n, c->prev, c->pos0, c->frc,
                                                           c->mem, c->npos0, c->nfrc, c->nmem,
                                                           c->pos, c->prm.box);
    CK(hipGetLastError());
    // copied back rather than swapped: the buffers stay where a captured
    // step sequence expects them
    const size_t b = sizeof(double) * 3 * (size_t)c->n;
    CK(hipMemcpyAsync(c->pos0, c->npos0, b, hipMemcpyDeviceToDevice, c->stream));
    CK(hipMemcpyAsync(c->frc, c->nfrc, b, hipMemcpyDeviceToDevice, c->stream));
    CK(hipMemcpyAsync(c->mem, c->nmem, b, hipMemcpyDeviceToDevice, c->stream));
    c->lagged = false;
    return 0;
}

// The common LJ exponents (ljNum 6, 12, 18) get kernels with the exponent
// fixed at compile time; any other value takes the general kernel.
template <int MASK, class Launch>
void with_lj_exponent(int lj_num, Launch launch)
{
    if constexpr ((MASK & F_LJ) != 0) {
        switch (lj_num) {
        case 18: launch(std::integral_constant<int, 18>{}); return;
        case 12: launch(std::integral_constant<int, 12>{}); return;
        case 6: launch(std::integral_constant<int, 6>{}); return;
        default: break;
        }
    }
    launch(std::integral_constant<int, 0>{});
}

template <int MASK>
void launch_step_mask(psim_ctx *c, const StepArgs &a)
{
    if (c->dense) {
        const int per = (c->ncell + 7) / 8;
        with_lj_exponent<MASK>(c->pot.lj_num, [&](auto ljn) {
            k_step_dense<MASK, decltype(ljn)::value><<<8 * per, kDenseThreads, 0, c->stream>>>(
                c->geo, c->pot, c->prm.radius, a.pos, a.id, a.cstart, a.nfrc, a.err, per);
        });
    } else if (c->brick) {
        const int per = (c->bg.nbricks + 7) / 8;
        if (c->brick_g == 1)
            k_step_brick<MASK, 1><<<8 * per, kBrickThreads, 0, c->stream>>>(
                c->geo, c->pot, c->gb, c->prm.radius, c->bg, per, a);
        else
            k_step_brick<MASK, kSmallSplit><<<8 * per, kBrickThreads, 0, c->stream>>>(
                c->geo, c->pot, c->gb, c->prm.radius, c->bg, per, a);
    } else if (c->n < kSmallSystem) {
        // few particles: small blocks and kSmallSplit lanes per particle so
        // that every CU gets work
        k_step<MASK, 64, kSmallSplit>
            <<<xcd_grid(blocks_for((long long)c->n * kSmallSplit, 64)), 64, 0, c->stream>>>(
                c->geo, c->pot, c->gb, c->prm.radius, a);
    } else {
        k_step<MASK, kBlock><<<xcd_grid(blocks_for(c->n, kBlock)), kBlock, 0, c->stream>>>(
            c->geo, c->pot, c->gb, c->prm.radius, a);
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
        with_lj_exponent<MASK>(c->pot.lj_num, [&](auto ljn) {
            k_forces_dense<MASK, decltype(ljn)::value>
                <<<8 * per, kDenseThreads, 0, c->stream>>>(c->geo, c->pot, c->prm.radius, c->pos,
                                                           c->id, c->cstart, c->fout, c->err,
                                                           per);
        });
    } else if (c->brick) {
        const int per = (c->bg.nbricks + 7) / 8;
        k_forces_brick<MASK><<<8 * per, kBrickThreads, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, c->bg, per, c->pos, c->id, c->cstart, c->fout,
            c->err);
    } else {
        k_forces<MASK><<<xcd_grid(blocks_for(c->n, kBlock)), kBlock, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, c->pos, c->id, c->cstart, c->fout, c->err);
    }
}

// Neighbour-set statistics with the launch shape launch_step_mask uses.
void launch_stats(psim_ctx *c, NbrStatsOut so)
{
    if (c->dense) {
        const int per = (c->ncell + 7) / 8;
        k_stats_dense<<<8 * per, kDenseThreads, 0, c->stream>>>(c->geo, c->pot, c->prm.radius,
                                                                c->pos, c->id, c->cstart, so,
                                                                per);
    } else if (c->brick) {
        const int per = (c->bg.nbricks + 7) / 8;
        if (c->brick_g == 1)
            k_stats_brick<1><<<8 * per, kBrickThreads, 0, c->stream>>>(
                c->geo, c->pot, c->prm.radius, c->bg, per, c->pos, c->id, c->cstart, so);
        else
            k_stats_brick<kSmallSplit><<<8 * per, kBrickThreads, 0, c->stream>>>(
                c->geo, c->pot, c->prm.radius, c->bg, per, c->pos, c->id, c->cstart, so);
    } else if (c->n < kSmallSystem) {
        k_stats<64, kSmallSplit>
            <<<xcd_grid(blocks_for((long long)c->n * kSmallSplit, 64)), 64, 0, c->stream>>>(
                c->geo, c->pot, c->prm.radius, c->pos, c->id, c->cstart, so);
    } else {
        k_stats<kBlock, 1><<<xcd_grid(blocks_for(c->n, kBlock)), kBlock, 0, c->stream>>>(
            c->geo, c->pot, c->prm.radius, c->pos, c->id, c->cstart, so);
    }
}

template <int MASK>
void launch_table_mask(psim_ctx *c)
{
    k_forces_table<MASK><<<blocks_for(c->n, kBlock), kBlock, 0, c->stream>>>(
        c->geo, c->pot, c->tsorted, c->thi, c->tcse, c->fout, c->err);
}

// The step arguments of the context's current buffers.  ctl: the kernels
// read the step index and the first-step / velocity-step flags from device
// memory (graph replay) instead of from the arguments.
StepArgs step_args(psim_ctx *c, const double *noise, const StepCtl *ctl)
{
    StepArgs a{};
    a.pos = c->pos; a.pos0 = c->pos0; a.frc = c->frc; a.mem = c->mem;
    a.id = c->id; a.cstart = c->cstart;
    a.prev = c->lagged ? c->prev : nullptr;
    a.x0_raw = c->lagged ? 1 : 0;
    a.npos = c->npos; a.npos0 = c->npos0; a.nfrc = c->nfrc; a.nmem = c->nmem;
    a.nid = nullptr; a.vel = c->vel; a.key = c->key; a.arank = c->arank;
    a.count = c->count; a.err = c->err;
    a.pbc = c->pbc;
    a.noise = noise;
    a.seed = c->prm.seed;
    a.step = c->step_index;
    a.ctl = ctl;
    return a;
}

// The first-step and velocity-step flags of the next step, advancing the
// velocity counter (brownianIntegrator.cpp:183, 288, 292).
void next_flags(psim_ctx *c, int *first, int *do_vel)
{
    *first = (c->time == 0) ? 1 : 0;
    *do_vel = 0;
    if (!*first) {
        *do_vel = (c->prm.vel_freq == 0 || c->vel_counter == c->prm.vel_freq) ? 1 : 0;
        if (c->vel_counter == c->prm.vel_freq) c->vel_counter = 0; else c->vel_counter++;
    }
}

// One step's kernels: forces (+ G&B), separate G&B on the dense path, cell
// rebuild.
int enqueue_step(psim_ctx *c, const StepArgs &a, psim_ctx::Mark *mk, StepCtl *advance = nullptr)
{
    if (mk) {
        hipEvent_t e = take_event(c, &mk->e0);
        if (e) CK(hipEventRecord(e, c->stream));
    }
    int rc = launch_step(c, a);
    if (rc) return rc;
    if (mk) {
        hipEvent_t e = take_event(c, &mk->e1);
        if (e) CK(hipEventRecord(e, c->stream));
    }
    if (c->dense) {
        k_integrate<<<blocks_for(c->n, 256), 256, 0, c->stream>>>(c->geo, c->gb, a);
        CK(hipGetLastError());
    }
    if (mk) {
        hipEvent_t e = take_event(c, &mk->e1b);
        if (e) CK(hipEventRecord(e, c->stream));
    }
    rc = launch_build(c, advance);
    if (rc) return rc;
    if (mk) {
        hipEvent_t e = take_event(c, &mk->e2);
        if (e) CK(hipEventRecord(e, c->stream));
    }
    return 0;
}

// What a captured step sequence bakes into its kernel arguments.
GraphKey graph_key(const psim_ctx *c)
{
    GraphKey k{};
    k.pos = c->pos; k.pos0 = c->pos0; k.frc = c->frc; k.mem = c->mem;
    k.npos = c->npos; k.npos0 = c->npos0; k.nfrc = c->nfrc; k.nmem = c->nmem;
    k.id = c->id; k.nid = c->nid; k.pbc = c->pbc;
    k.geo = c->geo; k.pot = c->pot; k.gb = c->gb; k.seed = c->prm.seed;
    return k;
}

__global__ void k_iota(int n, int *p)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) p[s] = s;
}

void drop_graph(psim_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

// graph_steps steps from one graph launch (captured on first use, and again
// whenever the buffers or constants it baked in have changed).  The step
// sequence swaps the lagged buffers once per step, so an even number of
// steps leaves them where the capture found them.
int replay_graph(psim_ctx *c)
{
    if (!c->lagged) {
        // the captured steps read pos0 / frc / mem through prev
        k_iota<<<blocks_for(c->n, 256), 256, 0, c->stream>>>(c->n, c->prev);
        CK(hipGetLastError());
        c->lagged = true;
    }
    const GraphKey key = graph_key(c);
    if (c->gexec && std::memcmp(&key, &c->gkey, sizeof key) != 0) drop_graph(c);
    // step control of the first captured step; each step's last kernel
    // prepares the next
    StepCtl h{};
    h.step = c->step_index;
    h.time = c->time;
    h.dt = c->prm.dt;
    h.vel_freq = c->prm.vel_freq;
    next_flags(c, &h.first, &h.do_vel);
    h.vel_counter = c->vel_counter;
    CK(hipMemcpyAsync(c->ctl, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    if (!c->gexec) {
        CK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        for (int k = 0; k < c->graph_steps && !rc; ++k) {
            const StepArgs a = step_args(c, nullptr, c->ctl);
            rc = enqueue_step(c, a, nullptr, c->ctl);
        }
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(c->stream, &g);
        if (rc || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            std::fprintf(stderr, "psim_gpu: step capture failed: %s\n", hipGetErrorString(ce));
            return rc ? rc : PSIM_E_HIP;
        }
        c->graph = g;
        CK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        c->gkey = key;
    }
    CK(hipGraphLaunch(c->gexec, c->stream));
    // the host's copy of the step control, as the eager loop leaves it
    for (int k = 0; k < c->graph_steps; ++k) {
        if (k > 0) {
            int f, v;
            next_flags(c, &f, &v);
        }
        c->step_index += 1;
        c->time += c->prm.dt;   // system.cpp:210
    }
    return 0;
}

int run_steps(psim_ctx *c, int nsteps, const double *d_noise)
{
    if (!c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    int s = 0;
    if (c->graph_steps > 0 && !d_noise && !c->profiling) {
        for (; s + c->graph_steps <= nsteps; s += c->graph_steps) {
            const int rc = replay_graph(c);
            if (rc) return rc;
        }
    }
    for (; s < nsteps; ++s) {
        StepArgs a = step_args(c, d_noise ? d_noise + (size_t)s * 6 * n : nullptr, nullptr);
        next_flags(c, &a.first, &a.do_vel);
        psim_ctx::Mark mk{};
        int rc = enqueue_step(c, a, c->profiling ? &mk : nullptr);
        if (rc) return rc;
        if (c->profiling) {
            c->marks.push_back(mk);
            if (c->ev_used > 3000) {
                rc = collect_times(c);
                if (rc) return rc;
            }
        }
        c->step_index += 1;
        c->time += c->prm.dt;   // system.cpp:210
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

namespace psim {
namespace eng {

int dalloc_bytes(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        std::fprintf(stderr, "psim_gpu: hipMalloc(%zu B) failed: %s\n", bytes,
                     hipGetErrorString(e));
        *p = nullptr;
        return PSIM_E_NOMEM;
    }
    return 0;
}

// Upload a state (AoS3 by original index).  orig_by_slot (a permutation, or
// null for ascending original index) is the slot order the particles are
// staged in: the cell build keeps it within each cell, so a state saved with
// its slot order continues with the same force summation order.
int set_state(psim_ctx *c, const double *pos, const double *pos0, const double *vel,
              const double *frc, const double *memx, double current_time, int vel_counter,
              unsigned long long step_index, const int *orig_by_slot)
{
    if (!c || !pos) return PSIM_E_ARG;
    const int n = c->n;
    const double L = c->prm.box;
    std::vector<int> ids(n);
    if (orig_by_slot) {
        std::vector<char> seen(n, 0);
        for (int s = 0; s < n; ++s) {
            const int o = orig_by_slot[s];
            if (o < 0 || o >= n || seen[o]) return PSIM_E_ARG;
            seen[o] = 1;
            ids[s] = o;
        }
    } else {
        for (int s = 0; s < n; ++s) ids[s] = s;
    }
    // host AoS3 (original order) -> SoA into the n* staging arrays
    std::vector<double> h(3 * (size_t)n);
    auto soa = [&](const double *src, double *dst, bool is_pos0, bool by_slot) -> int {
        for (int s = 0; s < n; ++s)
            for (int d = 0; d < 3; ++d) {
                const int i = by_slot ? ids[s] : s;
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
                h[(size_t)d * n + s] = v;
            }
        CK(hipMemcpyAsync(dst, h.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                          c->stream));
        CK(hipStreamSynchronize(c->stream));
        return 0;
    };
    int rc;
    if ((rc = soa(pos, c->npos, false, true))) return rc;
    if ((rc = soa(pos0, c->npos0, true, true))) return rc;
    if ((rc = soa(vel, c->vel, false, false))) return rc;   // velocities by original index
    if ((rc = soa(frc, c->nfrc, false, true))) return rc;
    if ((rc = soa(memx, c->nmem, false, true))) return rc;
    CK(hipMemcpyAsync(c->id, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    CK(hipMemsetAsync(c->count, 0, sizeof(int) * c->ncell, c->stream));
    k_hash_count<<<blocks_for(n, 256), 256, 0, c->stream>>>(c->geo, c->npos, c->key,
                                                            c->arank, c->count);
    CK(hipGetLastError());
    rc = launch_build(c);
    if (rc) return rc;
    // the uploaded previous positions (already re-imaged, staging order),
    // forces and memory gathered into the slot order; the state is then
    // settled
    k_settle<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, c->prev, c->npos0, c->frc, c->mem,
                                                        c->pos0, c->nfrc, c->nmem);
    CK(hipGetLastError());
    const size_t b3 = sizeof(double) * 3 * (size_t)n;
    CK(hipMemcpyAsync(c->frc, c->nfrc, b3, hipMemcpyDeviceToDevice, c->stream));
    CK(hipMemcpyAsync(c->mem, c->nmem, b3, hipMemcpyDeviceToDevice, c->stream));
    c->lagged = false;
    CK(hipStreamSynchronize(c->stream));
    c->time = current_time;
    c->vel_counter = vel_counter;
    c->step_index = step_index;
    c->have_state = true;
    return 0;
}

int launch_build(psim_ctx *c) { return ::launch_build(c, nullptr); }
int settle(psim_ctx *c) { return ::settle(c); }
int run_steps(psim_ctx *c, int nsteps, const double *d_noise)
{
    return ::run_steps(c, nsteps, d_noise);
}
int read_error(psim_ctx *c, int *code, int *i, int *j) { return ::read_error(c, code, i, j); }

}  // namespace eng
}  // namespace psim

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
    case PSIM_E_IO: return "file could not be read or written";
    default: return "unknown status";
    }
}

int psim_gpu_create(const psim_params *params, psim_ctx **out)
{
    if (!params || !out) return PSIM_E_ARG;
    *out = nullptr;
    const psim_params &P = *params;
    const int chk = check_params(P);
    if (chk) return chk;
    psim_ctx *c = new (std::nothrow) psim_ctx();
    if (!c) return PSIM_E_NOMEM;
    c->prm = P;
    fill_constants(c);
    c->n = P.n_particles;
    c->ncell = c->geo.ncell;
    c->nscan = (c->ncell + kScanTile - 1) / kScanTile;
    const PathChoice path = choose_path(P, c->geo);
    c->dense = path.dense;
    c->brick = path.brick;
    c->brick_g = path.brick_g;
    c->bg = path.bg;
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
    if (!rc) rc = dalloc(&c->prev, c->n);
    if (!rc) rc = dalloc(&c->ctl, 1);
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

int psim_gpu_set_graph(psim_ctx *c, int steps_per_graph)
{
    if (!c || steps_per_graph < 0 || (steps_per_graph & 1)) return PSIM_E_ARG;
    CK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->graph_steps = steps_per_graph;
    return 0;
}

int psim_gpu_get_params(psim_ctx *c, psim_params *out)
{
    if (!c || !out) return PSIM_E_ARG;
    *out = c->prm;
    return 0;
}

int psim_gpu_kernel_info(psim_ctx *c, int *path, int brick[4])
{
    if (!c || !path) return PSIM_E_ARG;
    *path = c->dense ? 2 : (c->brick ? 3 : 1);
    if (brick) {
        brick[0] = c->brick ? c->bg.bx : 0;
        brick[1] = c->brick ? c->bg.by : 0;
        brick[2] = c->brick ? c->bg.bz : 0;
        brick[3] = c->brick ? c->brick_g : (c->dense ? 0 : (c->n < kSmallSystem ? kSmallSplit : 1));
    }
    return 0;
}

void psim_gpu_destroy(psim_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    void *bufs[] = {c->pos, c->pos0, c->frc, c->mem, c->npos, c->npos0, c->nfrc,
                    c->nmem, c->vel, c->fout, c->id, c->nid, c->key, c->arank,
                    c->perm, c->prev, c->ctl, c->count, c->cstart, c->bsum, c->err, c->noise,
                    c->tsorted, c->thi, c->tcse, c->st_count, c->st_hash, c->pbc,
                    c->track0, c->hg.key,
                    c->hg.rank, c->hg.perm, c->hg.count, c->hg.cstart, c->hg.bsum,
                    c->hg.hpos,
                    c->work, c->iwork};
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
    // frc holds the force of the previous step (the next push makes it frc0);
    // frc0 of the caller is not needed by the integrator after the push
    (void)frc0;
    return psim::eng::set_state(c, pos, pos0, vel, frc, memx, current_time, vel_counter,
                                step_index, nullptr);
}

int psim_gpu_get_state(psim_ctx *c, double *pos, double *pos0, double *vel, double *frc,
                       double *frc0, double *memx, double *current_time,
                       int *vel_counter, unsigned long long *step_index)
{
    if (!c || !c->have_state) return PSIM_E_ARG;
    int rs = psim::eng::settle(c);
    if (rs) return rs;
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

int psim_gpu_pair_stats(psim_ctx *c, int which, int *count_by_orig,
                        unsigned long long *hash_by_orig)
{
    if (!c || !c->have_state || !count_by_orig || !hash_by_orig) return PSIM_E_ARG;
    if (which != PSIM_PAIRS_STEP && which != PSIM_PAIRS_ADJOINT) return PSIM_E_ARG;
    const int n = c->n;
    int rc = 0;
    if (!c->st_count) {
        if (!rc) rc = dalloc(&c->st_count, n);
        if (!rc) rc = dalloc(&c->st_hash, n);
        if (rc) return rc;
    }
    const NbrStatsOut so{c->st_count, c->st_hash};
    if (which == PSIM_PAIRS_STEP)
        launch_stats(c, so);
    else
        psim::eng::launch_adj_stats(c, so);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(count_by_orig, c->st_count, sizeof(int) * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipMemcpyAsync(hash_by_orig, c->st_hash, sizeof(unsigned long long) * n,
                      hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return 0;
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

int psim_gpu_set_force_params(psim_ctx *c, const psim_params *params)
{
    if (!c || !params) return PSIM_E_ARG;
    const psim_params &P = *params;
    if (!(P.cutoff > 0) || P.force_mask <= 0 || P.force_mask > 7) return PSIM_E_ARG;
    if ((P.force_mask & PSIM_FORCE_LJ) && !(P.debye_length > 0)) return PSIM_E_ARG;
    CK(hipStreamSynchronize(c->stream));
    c->prm.force_mask = P.force_mask;
    c->prm.lj_num = P.lj_num;
    c->prm.cutoff = P.cutoff;
    c->prm.lj_well_depth = P.lj_well_depth;
    c->prm.yukawa_strength = P.yukawa_strength;
    c->prm.debye_length = P.debye_length;
    c->prm.ao_well_depth = P.ao_well_depth;
    fill_constants(c);
    return 0;
}

}  // extern "C"
