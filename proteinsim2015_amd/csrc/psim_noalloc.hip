// psim_noalloc.hip -- forces at caller-held positions with no device
// allocation: psim_gpu_forces_noalloc (include/psim_gpu.h).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  Inputs and
// outputs are the caller's device arrays; the cell build's working arrays
// (keys, ranks, counts, the cell start table, the slot permutation, the
// sorted positions, the error record) live in `scratch`, a file-scope
// device array of PSIM_NOALLOC_SCRATCH_BYTES that each GPU holds until the
// process ends.  Calls on one GPU share it: calls that overlap on one GPU
// race, and every call clears what must start at zero.
//
// The result equals psim_gpu_forces on a context created with the same
// parameters and given the same positions, bit for bit: the same cell
// build (within a cell, ascending original index), the same kernel choice
// and the same force kernels, with the output written as N x 3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <type_traits>

#include "psim_build.hpp"
#include "psim_ctx.hpp"
#include "psim_params.hpp"
#include "psim_step.hpp"

using namespace psim;

namespace {

constexpr size_t kScratchBytes = PSIM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU (each holds its own copy).
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

int scratch_for_current_device(void **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return PSIM_E_HIP;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            std::fprintf(stderr, "psim_gpu_forces_noalloc: hipGetSymbolAddress failed: %s\n",
                         hipGetErrorString(e));
            return PSIM_E_HIP;
        }
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return 0;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Carving of the scratch array for n particles and ncell cells.
struct Carve {
    size_t key, rank, perm, id, count, cstart, bsum, pos, err, total;
};
Carve carve(size_t n, size_t ncell, size_t nscan)
{
    Carve c{};
    size_t o = 0;
    c.key = o;    o += align256(4 * n);
    c.rank = o;   o += align256(4 * n);
    c.perm = o;   o += align256(4 * n);
    c.id = o;     o += align256(4 * n);
    c.count = o;  o += align256(4 * ncell);
    c.cstart = o; o += align256(4 * (ncell + 1));
    c.bsum = o;   o += align256(4 * (nscan ? nscan : 1));
    c.pos = o;    o += align256(8 * 3 * n);
    c.err = o;    o += align256(sizeof(DevErr));
    c.total = o;
    return c;
}

// Zero fill (the build's counts and the error record).
__global__ void k_fill_zero(int *p, int count)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        p[i] = 0;
}

// Cell hash (systemHandling.cpp:33-48) of particle s = original index s from
// N x 3 positions, and its rank within the cell.
__global__ void k_hash_aos(Geo g, const double *pos3, int *key, int *rank, int *count)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.n) return;
    const size_t b = 3 * (size_t)s;
    const int k = cell_1d(pos3[b], g.cell, g.S) + g.S * cell_1d(pos3[b + 1], g.cell, g.S)
                + g.S2 * cell_1d(pos3[b + 2], g.cell, g.S);
    key[s] = k;
    rank[s] = wave_aggregated_rank(count, k);
}

// Within a cell, ascending original index (the order a context's cell build
// gives an identity start); gathers the positions into slot order [3][N].
__global__ void k_move_aos(int n, const int *__restrict__ perm, const int *__restrict__ key,
                           const int *__restrict__ cstart, const double *__restrict__ pos3,
                           double *__restrict__ pos, int *__restrict__ id)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm[t];
    const int c = key[s];
    const int b = cstart[c], e = cstart[c + 1];
    int rank = 0;
    for (int u = b; u < e; ++u) rank += (perm[u] < s);
    const int dst = b + rank;
    const size_t q = 3 * (size_t)s;
    pos[dst] = pos3[q];
    pos[n + dst] = pos3[q + 1];
    pos[2 * n + dst] = pos3[q + 2];
    id[dst] = s;
}

template <int MASK>
void launch_forces_aos(const psim_params &P, const Geo &geo, const Pot &pot,
                       const PathChoice &path, const double *pos, const int *id,
                       const int *cstart, double *fout, DevErr *err, hipStream_t st)
{
    const int n = geo.n;
    if (path.dense) {
        const int per = (geo.ncell + 7) / 8;
        auto launch = [&](auto ljn) {
            k_forces_dense<MASK, decltype(ljn)::value, true>
                <<<8 * per, kDenseThreads, 0, st>>>(geo, pot, P.radius, pos, id, cstart, fout,
                                                    err, per);
        };
        if constexpr ((MASK & F_LJ) != 0) {
            switch (pot.lj_num) {
            case 18: launch(std::integral_constant<int, 18>{}); return;
            case 12: launch(std::integral_constant<int, 12>{}); return;
            case 6: launch(std::integral_constant<int, 6>{}); return;
            default: break;
            }
        }
        launch(std::integral_constant<int, 0>{});
    } else if (path.brick) {
        const int per = (path.bg.nbricks + 7) / 8;
        k_forces_brick<MASK, true><<<8 * per, kBrickThreads, 0, st>>>(
            geo, pot, P.radius, path.bg, per, pos, id, cstart, fout, err);
    } else {
        k_forces<MASK, true><<<xcd_grid(eng::blocks_for(n, kBlock)), kBlock, 0, st>>>(
            geo, pot, P.radius, pos, id, cstart, fout, err);
    }
}

}  // namespace

extern "C" int psim_gpu_forces_noalloc(const psim_params *params, const double *d_pos,
                                       double *d_force, void *stream)
{
    if (!params || !d_pos || !d_force) return PSIM_E_ARG;
    const psim_params &P = *params;
    const int chk = check_params(P);
    if (chk) return chk;
    Geo geo;
    Pot pot;
    GB gb;
    make_constants(P, geo, pot, gb);
    const int n = geo.n, ncell = geo.ncell;
    const int nscan = (ncell + kScanTile - 1) / kScanTile;
    const Carve cv = carve((size_t)n, (size_t)ncell, (size_t)nscan);
    if (cv.total > kScratchBytes) {
        std::fprintf(stderr,
                     "psim_gpu_forces_noalloc: %d particles in %d cells need %zu bytes of "
                     "scratch; the maximum is %zu (use psim_gpu_create + psim_gpu_forces)\n",
                     n, ncell, cv.total, kScratchBytes);
        return PSIM_E_NOMEM;
    }
    void *base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    unsigned char *b = static_cast<unsigned char *>(base);
    int *key = reinterpret_cast<int *>(b + cv.key);
    int *rank = reinterpret_cast<int *>(b + cv.rank);
    int *perm = reinterpret_cast<int *>(b + cv.perm);
    int *id = reinterpret_cast<int *>(b + cv.id);
    int *count = reinterpret_cast<int *>(b + cv.count);
    int *cstart = reinterpret_cast<int *>(b + cv.cstart);
    int *bsum = reinterpret_cast<int *>(b + cv.bsum);
    double *pos = reinterpret_cast<double *>(b + cv.pos);
    DevErr *err = reinterpret_cast<DevErr *>(b + cv.err);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const PathChoice path = choose_path(P, geo);

    // the previous call's contents: counts and the error record start at zero
    const int fill_blocks = (int)std::min<long long>(1024, eng::blocks_for(ncell, 256));
    k_fill_zero<<<fill_blocks, 256, 0, st>>>(count, ncell);
    k_fill_zero<<<1, 64, 0, st>>>(reinterpret_cast<int *>(err), (int)(sizeof(DevErr) / 4));
    k_hash_aos<<<eng::blocks_for(n, 256), 256, 0, st>>>(geo, d_pos, key, rank, count);
    const bool direct = nscan <= kScanDirectTiles;
    if (direct) {
        k_scan_direct<<<nscan, kScanThreads, 0, st>>>(count, ncell, cstart, n);
    } else {
        launch_scan3(st, count, ncell, nscan, bsum, cstart, n);
    }
    k_scatter<1><<<eng::blocks_for(n, 256), 256, 0, st>>>(n, key, rank, cstart, perm,
                                                       direct ? count : nullptr, ncell);
    k_move_aos<<<eng::blocks_for(n, 256), 256, 0, st>>>(n, perm, key, cstart, d_pos, pos, id);
    switch (pot.mask) {
    case 1: launch_forces_aos<1>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    case 2: launch_forces_aos<2>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    case 3: launch_forces_aos<3>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    case 4: launch_forces_aos<4>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    case 5: launch_forces_aos<5>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    case 6: launch_forces_aos<6>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    default: launch_forces_aos<7>(P, geo, pot, path, pos, id, cstart, d_force, err, st); break;
    }
    CK(hipGetLastError());
    DevErr h{};
    CK(hipMemcpyAsync(&h, err, sizeof(DevErr), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    return h.code;
}
