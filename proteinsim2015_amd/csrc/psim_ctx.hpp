// psim_ctx.hpp -- the single-system context shared by the engine, analysis
// and adjoint translation units, and the engine internals they call.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "psim_brick_shape.hpp"
#include "psim_device.hpp"
#include "psim_gpu.h"

// The buffers and constants a captured step sequence bakes into its kernels.
struct GraphKey {
    const void *pos, *pos0, *frc, *mem, *npos, *npos0, *nfrc, *nmem, *id, *nid, *pbc;
    psim::Geo geo;
    psim::Pot pot;
    psim::GB gb;
    unsigned long long seed;
};

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct psim_ctx {
    psim_params prm;
    psim::Geo geo;
    psim::Pot pot;
    psim::GB gb;
    int n = 0, ncell = 0, nscan = 0;
    bool dense = false;   // cell-per-workgroup pair kernel (else thread per particle)
    bool brick = false;   // brick-per-workgroup pair kernel (sparse systems)
    psim::BrickGeo bg{};  // its bricks
    int brick_g = 1;      // its lanes per particle
    hipStream_t stream = nullptr;
    double *pos = nullptr, *pos0 = nullptr, *frc = nullptr, *mem = nullptr;
    double *npos = nullptr, *npos0 = nullptr, *nfrc = nullptr, *nmem = nullptr;
    double *vel = nullptr, *fout = nullptr;
    int *id = nullptr, *nid = nullptr, *key = nullptr, *arank = nullptr, *perm = nullptr;
    // pos0 / frc / mem stay in the slot order of the last step: pos0 the
    // positions that step started from (not yet re-imaged), frc and mem what
    // it wrote; prev[s] is the slot in that order of the particle now in
    // slot s (lagged), until settle() gathers them into the current order,
    // pos0 re-imaged against pos (npos0 is settle()'s staging buffer)
    int *prev = nullptr;
    bool lagged = false;
    // graph replay of graph_steps steps (0: eager launches); ctl is the
    // device-side step control the captured kernels read
    int graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey{};
    psim::StepCtl *ctl = nullptr;
    int *count = nullptr, *cstart = nullptr, *bsum = nullptr;
    psim::DevErr *err = nullptr;
    double *noise = nullptr;
    size_t noise_cap = 0;
    double *tsorted = nullptr;
    int *thi = nullptr, *tcse = nullptr;
    int *st_count = nullptr;                 // neighbour-set statistics
    unsigned long long *st_hash = nullptr;
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
    // displacement tracker (analysisTracker.cpp:26-85), by original index:
    // periodic crossing counters and start positions, [3][N]
    int *pbc = nullptr;
    double *track0 = nullptr;
    // pair-histogram grid (its own cell edge >= the histogram's reach)
    struct HistGrid {
        int S = 0, ncell = 0, nscan = 0;
        double cell = 0.0;
        int *key = nullptr, *rank = nullptr, *perm = nullptr;
        int *count = nullptr, *cstart = nullptr, *bsum = nullptr;
        double *hpos = nullptr;   // positions in grid order, [3][n]
        size_t cap_cells = 0;
    } hg;
    // when set, the cell rebuild also writes the new positions, ids and cell
    // table there (the adjoint's record of a recomputed step)
    struct Record {
        double *pos = nullptr;
        int *id = nullptr, *cstart = nullptr;
        int *inv = nullptr;   // original index -> slot
    } rec;
    double *work = nullptr;     // analysis scratch (partials, histograms)
    size_t work_cap = 0;        // doubles
    int *iwork = nullptr;       // analysis scratch (labels, flags)
    size_t iwork_cap = 0;       // ints
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


namespace psim {
namespace eng {

int dalloc_bytes(void **p, size_t bytes);
template <class T>
inline int dalloc(T **p, size_t count)
{
    return dalloc_bytes((void **)p, (count ? count : 1) * sizeof(T));
}
inline unsigned blocks_for(long long n, int threads)
{
    return (unsigned)((n + threads - 1) / threads);
}
// counts (already accumulated) -> cstart, scatter, canonical move of the
// positions and ids (the state arrays become lagged)
int launch_build(psim_ctx *c);
// gather lagged pos0 / frc / mem into the current slot order (no-op when
// they are in it already)
int settle(psim_ctx *c);
// nsteps of system::run with the context's step bookkeeping
int run_steps(psim_ctx *c, int nsteps, const double *d_noise);
int read_error(psim_ctx *c, int *code, int *i, int *j);
// neighbour-set statistics of the adjoint's pair visitor (launch shape of
// its backward step)
void launch_adj_stats(psim_ctx *c, NbrStatsOut so);
// psim_gpu_set_state with an explicit staging slot order (null = identity)
int set_state(psim_ctx *c, const double *pos, const double *pos0, const double *vel,
              const double *frc, const double *memx, double current_time, int vel_counter,
              unsigned long long step_index, const int *orig_by_slot);

}  // namespace eng
}  // namespace psim
