// psim_analysis.hip -- on-device analysis of a context's current state:
// kinetic temperature, mean displacement, the tracked mean-square
// displacement, contact counts (coordination), contact clusters, and a
// Gaussian-smoothed pair-distance histogram with its position gradient.
//
// Definitions follow the reference's analysis manager:
//   temperature        analysisManager::getTemperature (analysisGeneral.cpp:26-44)
//   mean_disp2         analysisManager::meanDisplacement (analysisGeneral.cpp:46-56)
//   msd                analysisManager::trackedDisplacement
//                      (analysisTracker.cpp:26-85; FP64 here, float there)
//   contacts           system::updateInteractions / iterateParticleInteractions
//                      (systemHandling.cpp:94-161): r < contact radius over
//                      the 27 cells of the force grid
//   clusters           analysisManager::findClusters / writeClusters
//                      (analysisCluster.cpp:26-100)
// The pair histogram has no reference counterpart (the reference writes no
// g(r)); it is the observable of the structure-fitting loss.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "psim_build.hpp"
#include "psim_ctx.hpp"
#include "psim_pairs.hpp"

using namespace psim;
using eng::blocks_for;
using eng::dalloc;

namespace {

constexpr int kAnaBlock = 256;
constexpr int kMaxBins = 2048;        // bins held in LDS by the histogram kernel
constexpr double kWindow = 8.0;       // Gaussian support in units of sigma
constexpr int kMaxHistCells = 256;    // cells per dimension of the histogram grid

// ---- block reductions (256 lanes = 4 waves) ---------------------------------
template <int NQ>
__device__ void block_sums(double (&v)[NQ], double *out)
{
    __shared__ double red[NQ][kAnaBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = v[q];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) x += __shfl_down(x, d, 64);
        if (lane == 0) red[q][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double t = 0.0;
        for (int k = 0; k < kAnaBlock / 64; ++k) t += red[threadIdx.x][k];
        out[(size_t)blockIdx.x * NQ + threadIdx.x] = t;
    }
}

// ---- displacement tracker ---------------------------------------------------
// analysisManager::initDisplacementTracker (analysisTracker.cpp:26-36)
__global__ void k_track_init(int n, const double *pos, const int *id, double *track0, int *pbc)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int o = id[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        track0[(size_t)d * n + o] = pos[(size_t)d * n + s];
        pbc[(size_t)d * n + o] = 0;
    }
}

// Contacts of slot s: the 27 cells of the force grid around the particle's
// own cell, r = sqrt(pbcDist) < rc (systemHandling.cpp:94-122, 124-161).
template <class Fn>
__device__ void visit_contacts(const Geo &g, const double *pos, const int *cstart, int s,
                               double rc, Fn fn)
{
    const int N = g.n, S = g.S;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    const int cx = cell_1d(xi, g.cell, S), cy = cell_1d(yi, g.cell, S),
              cz = cell_1d(zi, g.cell, S);
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                int rx = (cx + ox) % S, ry = (cy + oy) % S, rz = (cz + oz) % S;
                rx = rx < 0 ? rx + S : rx;
                ry = ry < 0 ? ry + S : ry;
                rz = rz < 0 ? rz + S : rz;
                const int h = rx + S * ry + g.S2 * rz;
                for (int j = cstart[h]; j < cstart[h + 1]; ++j) {
                    if (j == s) continue;
                    const double r2 = pbc_dist_ref(xi, yi, zi, pos[j], pos[N + j],
                                                   pos[2 * N + j], g.L, g.half_pbc);
                    if (sqrt(r2) < rc) fn(j);
                }
            }
}

// One pass: per-particle coordination (by original index) and per-block
// partial sums of v^2, |x - x0|^2, the tracked squared displacement and the
// contact count.
__global__ __launch_bounds__(kAnaBlock) void k_observables(
    Geo g, const double *pos, const double *pos0, const double *vel, const int *id,
    const int *cstart, const double *track0, const int *pbc, double contact_r, int *coord,
    double *partials)
{
    const int s = blockIdx.x * kAnaBlock + threadIdx.x;
    const int N = g.n;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (s < N) {
        const int o = id[s];
        // getTemperature: vtot += vx*vx; += vy*vy; += vz*vz
        double vt = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double vd = vel[(size_t)d * N + o];
            vt += (vd * vd);
        }
        v[0] = vt;
        // meanDisplacement: dx*dx + dy*dy + dz*dz with dx = x - x0
        const double dx = pos[s] - pos0[s], dy = pos[N + s] - pos0[N + s],
                     dz = pos[2 * N + s] - pos0[2 * N + s];
        v[1] = dx * dx + dy * dy + dz * dz;
        if (track0) {
            // trackedDisplacement: dx = x - xStart; dx += boxSize * xPBC
            double m = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double e = pos[(size_t)d * N + s] - track0[(size_t)d * N + o];
                e += g.L * (double)pbc[(size_t)d * N + o];
                m += (e * e);
            }
            v[2] = m;
        }
        int nc = 0;
        if (contact_r > 0.0) visit_contacts(g, pos, cstart, s, contact_r, [&](int) { ++nc; });
        if (coord) coord[o] = nc;
        v[3] = (double)nc;
    }
    block_sums<4>(v, partials);
}

// ---- contact clusters (connected components) -------------------------------
__global__ void k_cc_init(int n, int *lab)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o < n) lab[o] = o;
}

// Minimum-label propagation over the contact graph.  Labels are always the
// original index of a member of the same component and never increase.
__global__ __launch_bounds__(kAnaBlock) void k_cc_hook(Geo g, const double *pos,
                                                       const int *id, const int *cstart,
                                                       double rc, int *lab, int *changed)
{
    const int s = blockIdx.x * kAnaBlock + threadIdx.x;
    if (s >= g.n) return;
    const int o = id[s];
    int m = lab[o];
    visit_contacts(g, pos, cstart, s, rc, [&](int j) { m = min(m, lab[id[j]]); });
    if (m < lab[o]) {
        atomicMin(&lab[o], m);
        atomicMin(&lab[lab[o]], m);
        changed[0] = 1;
    }
}

// Pointer jumping: lab[o] <- root of its label chain (lab[x] <= x, so every
// chain strictly decreases until it reaches a fixed point).
__global__ void k_cc_jump(int n, int *lab)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    int l = lab[o];
    for (int k = 0; k < n; ++k) {
        const int nl = lab[l];
        if (nl == l) break;
        l = nl;
    }
    lab[o] = l;
}

// ---- pair-distance histogram -----------------------------------------------
// Gauss–Legendre points per sub-interval of a bin (sigma > 0): with
// sub-intervals of at most 1.25 sigma the rule matches the CDF difference to
// the last bits (2.2e-16 at 1 sigma, 7.8e-16 at 1.25 against scipy's ndtr).
constexpr int kGLNodes = 8;
constexpr double kGLMaxStep = 1.25;

struct HistGeo {
    int n;
    int S, S2;          // histogram grid (S >= 3), or S == 0 for all pairs
    double L, half, cell;
    double r_lo, width, sigma, inv_sigma, reach2;
    double delta;       // width / sigma
    int nbins;
    // bin masses: gl_m sub-intervals of gl_h = delta / gl_m per bin, each by
    // the kGLNodes-point Gauss–Legendre rule (nodes gl_x in [0, 1], weights
    // gl_w summing to 1); gl_c = e^{-h^2}, gl_a[k] = e^{-h^2 (x_k + 1/2)}
    int gl_m;
    double gl_h, gl_c;
    double gl_x[kGLNodes], gl_w[kGLNodes], gl_a[kGLNodes];
};

__device__ inline double norm_cdf(double t) { return 0.5 * erfc(-t * 0.70710678118654752440); }
__device__ inline double norm_pdf(double t) { return 0.39894228040143267794 * exp(-0.5 * t * t); }

// visit every j != s within reach of slot s (histogram grid, or all pairs)
template <class Fn>
__device__ void visit_reach(const HistGeo &h, const double *pos, const int *perm,
                            const int *cstart, int s, Fn fn)
{
    const int N = h.n;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    auto test = [&](int j) {
        if (j == s) return;
        const double xj = pos[j], yj = pos[N + j], zj = pos[2 * N + j];
        const double r2 = pbc_dist_ref(xi, yi, zi, xj, yj, zj, h.L, h.half);
        if (r2 < h.reach2) fn(j, xj, yj, zj, r2);
    };
    if (h.S == 0) {
        for (int j = 0; j < N; ++j) test(j);
        return;
    }
    const int S = h.S;
    const int cx = cell_1d(xi, h.cell, S), cy = cell_1d(yi, h.cell, S),
              cz = cell_1d(zi, h.cell, S);
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                int rx = cx + ox, ry = cy + oy, rz = cz + oz;
                rx = rx < 0 ? rx + S : (rx >= S ? rx - S : rx);
                ry = ry < 0 ? ry + S : (ry >= S ? ry - S : ry);
                rz = rz < 0 ? rz + S : (rz >= S ? rz - S : rz);
                const int c = rx + S * ry + h.S2 * rz;
                for (int u = cstart[c]; u < cstart[c + 1]; ++u) test(perm[u]);
            }
}

// edge window of a pair at distance r: edges k in [k0, k1] carry the
// Gaussian's mass; below k0 the CDF is 0, above k1 it is 1 (to 6e-16)
__device__ inline void edge_window(const HistGeo &h, double r, int &k0, int &k1)
{
    const double lo = (r - kWindow * h.sigma - h.r_lo) / h.width;
    const double hi = (r + kWindow * h.sigma - h.r_lo) / h.width;
    k0 = (int)fmax(0.0, ceil(lo));
    k1 = (int)fmin((double)h.nbins, floor(hi));
}

// The Gaussian's mass (mean r, width sigma > 0) in each bin it reaches,
// handed to add(bin, mass): bin k0 - 1 takes the mass below edge k0 (the CDF
// there), bins k0 .. k1 - 1 the mass between their edges, bin k1 the mass
// above edge k1 (k0, k1 from edge_window), as the CDF differences at the
// edges would give them.  The mass between edges is the Gauss–Legendre rule
// over gl_m sub-intervals, with the density at each node carried from one
// sub-interval to the next by phi(u + h) = phi(u) e^{-h u - h^2/2}: one CDF
// and kGLNodes + 1 exponentials per pair instead of one CDF per edge.
template <class Add>
__device__ __forceinline__ void pair_masses(const HistGeo &h, double r, Add add)
{
    int k0, k1;
    edge_window(h, r, k0, k1);
    const int nb = h.nbins;
    if (k0 > k1) {
        if (k1 >= 0 && k1 < nb) add(k1, 1.0);
        return;
    }
    const double t0 = (h.r_lo + h.width * k0 - r) * h.inv_sigma;
    double E = norm_cdf(t0);   // the mass below edge k0
    if (k0 - 1 >= 0 && k0 - 1 < nb) add(k0 - 1, E);
    if (k0 < k1) {
        double phi[kGLNodes], g[kGLNodes];
        const double A = exp(-h.gl_h * t0);
#pragma unroll
        for (int k = 0; k < kGLNodes; ++k) {
            phi[k] = norm_pdf(fma(h.gl_h, h.gl_x[k], t0));
            g[k] = A * h.gl_a[k];
        }
        for (int q = k0; q < k1; ++q) {   // bin q: edges q and q + 1
            double m = 0.0;
            for (int sub = 0; sub < h.gl_m; ++sub) {
#pragma unroll
                for (int k = 0; k < kGLNodes; ++k) {
                    m = fma(h.gl_w[k], phi[k], m);
                    phi[k] *= g[k];
                    g[k] *= h.gl_c;
                }
            }
            m *= h.gl_h;
            if (q < nb) add(q, m);
            E += m;
        }
    }
    if (k1 >= 0 && k1 < nb) add(k1, 1.0 - E);
}

// h_b = sum over ordered pairs i != j of the Gaussian mass of r_ij in bin b
// (sigma > 0), or the count of r_ij in [e_b, e_b+1) (sigma == 0).
__global__ __launch_bounds__(kAnaBlock) void k_pair_hist(HistGeo h, const double *pos,
                                                         const int *perm, const int *cstart,
                                                         double *hist)
{
    __shared__ double sh[kMaxBins];
    for (int b = threadIdx.x; b < h.nbins; b += kAnaBlock) sh[b] = 0.0;
    __syncthreads();
    const int t = blockIdx.x * kAnaBlock + threadIdx.x;
    if (t < h.n) {
        const int s = h.S ? perm[t] : t;
        visit_reach(h, pos, perm, cstart, s, [&](int, double, double, double, double r2) {
            const double r = sqrt(r2);
            if (h.sigma == 0.0) {
                const double q = floor((r - h.r_lo) / h.width);
                if (q >= 0.0 && q < (double)h.nbins) atomicAdd(&sh[(int)q], 1.0);
                return;
            }
            pair_masses(h, r, [&](int b, double m) { atomicAdd(&sh[b], m); });
        });
    }
    __syncthreads();
    for (int b = threadIdx.x; b < h.nbins; b += kAnaBlock)
        if (sh[b] != 0.0) atomicAdd(&hist[b], sh[b]);
}

// grad_i = d(sum_b w_b h_b)/dx_i = -2 sum_j S(r_ij) d_ij / r_ij with
// S(r) = sum_b w_b dm_b/dr and d_ij the minimum-image x_j - x_i.
__global__ __launch_bounds__(kAnaBlock) void k_pair_hist_grad(HistGeo h, const double *pos,
                                                              const int *perm,
                                                              const int *cstart,
                                                              const int *id, const double *w,
                                                              double *grad)
{
    const int t = blockIdx.x * kAnaBlock + threadIdx.x;
    if (t >= h.n) return;
    const int N = h.n;
    const int s = h.S ? perm[t] : t;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    visit_reach(h, pos, perm, cstart, s, [&](int, double xj, double yj, double zj, double r2) {
        const double r = sqrt(r2);
        int k0, k1;
        edge_window(h, r, k0, k1);
        double sum = 0.0;
        for (int k = k0; k <= k1; ++k) {
            const double e = h.r_lo + h.width * k;
            const double wk = (k < h.nbins) ? w[k] : 0.0;
            const double wm = (k >= 1) ? w[k - 1] : 0.0;
            sum += norm_pdf((e - r) * h.inv_sigma) * (wk - wm);
        }
        const double S = sum * h.inv_sigma;
        double dx = xj - xi, dy = yj - yi, dz = zj - zi;
        dx = dx > h.half ? dx - h.L : (dx < -h.half ? dx + h.L : dx);
        dy = dy > h.half ? dy - h.L : (dy < -h.half ? dy + h.L : dy);
        dz = dz > h.half ? dz - h.L : (dz < -h.half ? dz + h.L : dz);
        const double c = -2.0 * S / r;
        gx += c * dx;
        gy += c * dy;
        gz += c * dz;
    });
    const int o = id[s];
    grad[o] = gx;
    grad[N + o] = gy;
    grad[2 * N + o] = gz;
}

// canonical within-cell order of the histogram grid (ascending slot), so the
// neighbour visit order does not depend on atomic timing
// (and the positions copied into that order, hpos [3][n])
__global__ void k_hg_canon(int n, const int *perm_in, const int *key, const int *cstart,
                           int *perm_out, const double *pos, double *hpos)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int s = perm_in[t];
    const int c = key[s];
    const int b = cstart[c], e = cstart[c + 1];
    int rank = 0;
    for (int u = b; u < e; ++u) rank += (perm_in[u] < s);
    perm_out[b + rank] = s;
#pragma unroll
    for (int d = 0; d < 3; ++d) hpos[(size_t)d * n + b + rank] = pos[(size_t)d * n + s];
}

// ---- pair histogram on the histogram grid (S >= 3) ---------------------------
// SPLIT lanes per particle (grid order t): lane `sub` walks the x-runs sub,
// sub + SPLIT, ... of the 27 grid cells (psim_pairs.hpp x_run) over the
// grid-ordered positions; every j != t within reach is visited with the
// reference's minimum-image distance.
constexpr int kHistSplit = 8;

template <int SPLIT, class Fn>
__device__ __forceinline__ void visit_grid(const HistGeo &h, const double *__restrict__ hpos,
                                           const int *__restrict__ cstart, int t, int sub,
                                           double xi, double yi, double zi, Fn fn)
{
    const int N = h.n, S = h.S;
    const int cx = cell_1d(xi, h.cell, S), cy = cell_1d(yi, h.cell, S),
              cz = cell_1d(zi, h.cell, S);
    const StartTab tab{cstart};
    for (int q = sub; q < 9; q += SPLIT) {
        int ry = cy + q / 3 - 1, rz = cz + q % 3 - 1;
        ry = ry < 0 ? ry + S : (ry >= S ? ry - S : ry);
        rz = rz < 0 ? rz + S : (rz >= S ? rz - S : rz);
        const XRun xr = x_run(tab, S * ry + h.S2 * rz, cx, S);
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
            const int e = sg ? xr.e1 : xr.e0;
            for (int u = sg ? xr.b1 : xr.b0; u < e; ++u) {
                if (u == t) continue;
                const double xj = hpos[u], yj = hpos[N + u], zj = hpos[2 * N + u];
                const double r2 = pbc_dist_ref(xi, yi, zi, xj, yj, zj, h.L, h.half);
                if (r2 < h.reach2) fn(xj, yj, zj, r2);
            }
        }
    }
}

// Per-block partial histograms (smoothed, sigma > 0).  Each lane keeps the
// r^2 of its pairs within reach in a short list in LDS and evaluates them
// (the edge CDFs) when some lane of the wave has filled its list, and once
// at the end, so that the costly part runs with most lanes busy.  Each wave
// adds into its own copy of the histogram in LDS; the block's sum goes to
// part[blockIdx.x][*], summed in a fixed order by k_hist_sum.
constexpr int kHistCap = 8;
template <int SPLIT>
__global__ __launch_bounds__(kAnaBlock) void k_pair_hist_grid(HistGeo h,
                                                              const double *__restrict__ hpos,
                                                              const int *__restrict__ cstart,
                                                              double *part)
{
    extern __shared__ double lds[];   // lists [kHistCap][kAnaBlock], then [4][nbins]
    double *lst = lds;
    double *wh = lds + kHistCap * kAnaBlock;
    const int nb = h.nbins;
    for (int b = threadIdx.x; b < (kAnaBlock / 64) * nb; b += kAnaBlock) wh[b] = 0.0;
    __syncthreads();
    double *mine = wh + (threadIdx.x >> 6) * nb;
    const int gt = blockIdx.x * kAnaBlock + threadIdx.x;
    const int t = gt / SPLIT, sub = gt % SPLIT;
    auto drain = [&](int nh) {
        for (int k = 0; k < nh; ++k) {
            const double r = sqrt(lst[k * kAnaBlock + threadIdx.x]);
            pair_masses(h, r, [&](int b, double m) { atomicAdd(&mine[b], m); });
        }
    };
    int nh = 0;
    if (t < h.n) {
        const int N = h.n, S = h.S;
        const double xi = hpos[t], yi = hpos[N + t], zi = hpos[2 * N + t];
        const int cx = cell_1d(xi, h.cell, S), cy = cell_1d(yi, h.cell, S),
                  cz = cell_1d(zi, h.cell, S);
        const StartTab tab{cstart};
        for (int q = sub; q < 9; q += SPLIT) {
            int ry = cy + q / 3 - 1, rz = cz + q % 3 - 1;
            ry = ry < 0 ? ry + S : (ry >= S ? ry - S : ry);
            rz = rz < 0 ? rz + S : (rz >= S ? rz - S : rz);
            const XRun xr = x_run(tab, S * ry + h.S2 * rz, cx, S);
#pragma unroll
            for (int sg = 0; sg < 2; ++sg) {
                const int e = sg ? xr.e1 : xr.e0;
                for (int u = sg ? xr.b1 : xr.b0; u < e; ++u) {
                    if (u != t) {
                        const double r2 = pbc_dist_ref(xi, yi, zi, hpos[u], hpos[N + u],
                                                       hpos[2 * N + u], h.L, h.half);
                        if (r2 < h.reach2) {
                            lst[nh * kAnaBlock + threadIdx.x] = r2;
                            ++nh;
                        }
                    }
                    if (__any(nh == kHistCap)) {
                        drain(nh);
                        nh = 0;
                    }
                }
            }
        }
    }
    drain(nh);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kAnaBlock) {
        double v = 0.0;
        for (int w = 0; w < kAnaBlock / 64; ++w) v += wh[w * nb + b];
        part[(size_t)blockIdx.x * nb + b] = v;
    }
}

// hist[b] = sum over the nparts block partials of bin b: block b, lane t
// sums partials t, t + 256, ... in order, then a fixed tree (reproducible).
__global__ __launch_bounds__(256) void k_hist_sum(const double *part, int nparts, int nb,
                                                  double *hist)
{
    __shared__ double red[256];
    const int b = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (int k = t; k < nparts; k += 256) v += part[(size_t)k * nb + b];
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) hist[b] = red[0];
}

// Gradient of sum_b w_b h_b on the histogram grid (as k_pair_hist_grad);
// dw[k] = w_k - w_{k-1} in LDS, and along a pair's edge window the Gaussian
// density by the recurrence phi(t + delta) = phi(t) e^{-delta t} e^{-delta^2/2}
// (two exponentials per pair instead of one per edge) when delta <= 1.
template <int SPLIT>
__global__ __launch_bounds__(kAnaBlock) void k_pair_hist_grad_grid(
    HistGeo h, const double *__restrict__ hpos, const int *__restrict__ perm,
    const int *__restrict__ cstart, const int *__restrict__ id, const double *__restrict__ w,
    double *grad)
{
    extern __shared__ double dw[];   // nbins + 1
    const int nb = h.nbins;
    for (int k = threadIdx.x; k <= nb; k += kAnaBlock)
        dw[k] = (k < nb ? w[k] : 0.0) - (k >= 1 ? w[k - 1] : 0.0);
    __syncthreads();
    const int gt = blockIdx.x * kAnaBlock + threadIdx.x;
    const int t = gt / SPLIT, sub = gt % SPLIT;
    if (t >= h.n) return;   // whole lane groups
    const int N = h.n;
    const double xi = hpos[t], yi = hpos[N + t], zi = hpos[2 * N + t];
    const bool recur = h.delta <= 1.0;
    const double dmul = exp(-h.delta * h.delta);          // c_{j+1} / c_j
    const double c0 = exp(-0.5 * h.delta * h.delta);      // c_0
    double gx = 0.0, gy = 0.0, gz = 0.0;
    visit_grid<SPLIT>(h, hpos, cstart, t, sub, xi, yi, zi,
                      [&](double xj, double yj, double zj, double r2) {
        const double r = sqrt(r2);
        int k0, k1;
        edge_window(h, r, k0, k1);
        double sum = 0.0;
        if (recur) {
            const double t0 = (h.r_lo + h.width * k0 - r) * h.inv_sigma;
            double phi = norm_pdf(t0);
            const double A = exp(-h.delta * t0);
            double cj = c0;
            for (int k = k0; k <= k1; ++k) {
                sum = fma(phi, dw[k], sum);
                phi *= A * cj;
                cj *= dmul;
            }
        } else {
            for (int k = k0; k <= k1; ++k)
                sum = fma(norm_pdf((h.r_lo + h.width * k - r) * h.inv_sigma), dw[k], sum);
        }
        const double S = sum * h.inv_sigma;
        double dx = xj - xi, dy = yj - yi, dz = zj - zi;
        dx = dx > h.half ? dx - h.L : (dx < -h.half ? dx + h.L : dx);
        dy = dy > h.half ? dy - h.L : (dy < -h.half ? dy + h.L : dy);
        dz = dz > h.half ? dz - h.L : (dz < -h.half ? dz + h.L : dz);
        const double c = -2.0 * S / r;
        gx = fma(c, dx, gx);
        gy = fma(c, dy, gy);
        gz = fma(c, dz, gz);
    });
#pragma unroll
    for (int d = 1; d < SPLIT; d <<= 1) {
        gx += __shfl_xor(gx, d, 64);
        gy += __shfl_xor(gy, d, 64);
        gz += __shfl_xor(gz, d, 64);
    }
    if (sub == 0) {
        const int o = id[perm[t]];
        grad[o] = gx;
        grad[N + o] = gy;
        grad[2 * N + o] = gz;
    }
}

// ---- host helpers -------------------------------------------------------------
int ensure_work(psim_ctx *c, size_t doubles, size_t ints)
{
    if (doubles > c->work_cap) {
        if (c->work) CK(hipFree(c->work));
        c->work = nullptr;
        c->work_cap = 0;
        int rc = dalloc(&c->work, doubles);
        if (rc) return rc;
        c->work_cap = doubles;
    }
    if (ints > c->iwork_cap) {
        if (c->iwork) CK(hipFree(c->iwork));
        c->iwork = nullptr;
        c->iwork_cap = 0;
        int rc = dalloc(&c->iwork, ints);
        if (rc) return rc;
        c->iwork_cap = ints;
    }
    return 0;
}

// Nodes x (ascending) and weights w (summing to 1) of the n-point
// Gauss–Legendre rule on [0, 1]: Newton's iteration on the roots of P_n,
// started from cos(pi (i + 3/4) / (n + 1/2)), w = 2 / ((1 - z^2) P_n'(z)^2)
// on [-1, 1] (Abramowitz & Stegun 25.4.29).
void gauss_legendre01(int n, double *x, double *w)
{
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            dp = n * (z * p1 - p2) / (z * z - 1.0);
            const double z1 = z;
            z = z1 - p1 / dp;
            if (std::fabs(z - z1) <= 1e-16) break;
        }
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

int check_hist_args(psim_ctx *c, double r_lo, double width, int nbins, double sigma)
{
    if (!(r_lo >= 0.0) || !(width > 0.0) || nbins <= 0 || nbins > kMaxBins || !(sigma >= 0.0))
        return PSIM_E_ARG;
    const double reach = r_lo + width * nbins + kWindow * sigma;
    // pairs beyond half the box have no unique minimum image
    if (reach > 0.5 * c->geo.L) return PSIM_E_UNSUPPORTED;
    return 0;
}

// Build the histogram grid for the current positions; fills h.
int build_hist_grid(psim_ctx *c, double r_lo, double width, int nbins, double sigma,
                    HistGeo &h)
{
    const int n = c->n;
    const double L = c->geo.L;
    const double reach = r_lo + width * nbins + kWindow * sigma;
    h.n = n;
    h.L = L;
    h.half = c->geo.half_pbc;
    h.r_lo = r_lo;
    h.width = width;
    h.sigma = sigma;
    h.inv_sigma = sigma > 0.0 ? 1.0 / sigma : 0.0;
    h.delta = width * h.inv_sigma;
    h.reach2 = reach * reach;
    h.nbins = nbins;
    gauss_legendre01(kGLNodes, h.gl_x, h.gl_w);
    h.gl_m = sigma > 0.0 ? std::max(1, (int)std::ceil(h.delta / kGLMaxStep)) : 1;
    h.gl_h = h.delta / h.gl_m;
    h.gl_c = std::exp(-h.gl_h * h.gl_h);
    for (int k = 0; k < kGLNodes; ++k) h.gl_a[k] = std::exp(-h.gl_h * h.gl_h * (h.gl_x[k] + 0.5));
    int S = (int)std::floor(L / reach);
    if (S > kMaxHistCells) S = kMaxHistCells;
    if (S < 3) {
        h.S = h.S2 = 0;
        h.cell = L;
        return 0;
    }
    h.S = S;
    h.S2 = S * S;
    h.cell = L / S;
    auto &g = c->hg;
    const int ncell = S * S * S;
    if ((size_t)ncell > g.cap_cells || !g.key) {
        void *bufs[] = {g.key, g.rank, g.perm, g.count, g.cstart, g.bsum, g.hpos};
        for (void *b : bufs)
            if (b) CK(hipFree(b));
        g = psim_ctx::HistGrid{};
        const int nscan = (ncell + kScanTile - 1) / kScanTile;
        int rc = 0;
        if (!rc) rc = dalloc(&g.key, n);
        if (!rc) rc = dalloc(&g.rank, n);
        if (!rc) rc = dalloc(&g.perm, 2 * (size_t)n);
        if (!rc) rc = dalloc(&g.count, ncell);
        if (!rc) rc = dalloc(&g.cstart, (size_t)ncell + 1);
        if (!rc) rc = dalloc(&g.bsum, nscan);
        if (!rc) rc = dalloc(&g.hpos, 3 * (size_t)n);
        if (rc) return rc;
        CK(hipMemsetAsync(g.count, 0, sizeof(int) * ncell, c->stream));
        g.cap_cells = ncell;
    }
    g.S = S;
    g.ncell = ncell;
    g.nscan = (ncell + kScanTile - 1) / kScanTile;
    g.cell = h.cell;
    Geo hg = c->geo;
    hg.S = S;
    hg.S2 = S * S;
    hg.ncell = ncell;
    hg.cell = h.cell;
    hg.c0 = 0;
    hipStream_t st = c->stream;
    k_hash_count<<<blocks_for(n, 256), 256, 0, st>>>(hg, c->pos, g.key, g.rank, g.count);
    launch_scan3(st, g.count, ncell, g.nscan, g.bsum, g.cstart, n);
    k_scatter<1><<<blocks_for(n, 256), 256, 0, st>>>(n, g.key, g.rank, g.cstart, g.perm + n,
                                                  nullptr, 0);
    k_hg_canon<<<blocks_for(n, 256), 256, 0, st>>>(n, g.perm + n, g.key, g.cstart, g.perm,
                                                   c->pos, g.hpos);
    CK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int psim_gpu_track_start(psim_ctx *c)
{
    if (!c || !c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    if (!c->pbc) {
        int rc = dalloc(&c->pbc, 3 * (size_t)n);
        if (!rc) rc = dalloc(&c->track0, 3 * (size_t)n);
        if (rc) return rc;
    }
    k_track_init<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, c->pos, c->id, c->track0, c->pbc);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

int psim_gpu_get_tracker(psim_ctx *c, double *start, int *crossings)
{
    if (!c || !c->have_state) return PSIM_E_ARG;
    if (!c->pbc) return PSIM_E_UNSUPPORTED;   // tracker not started
    const size_t n = (size_t)c->n;
    std::vector<double> hs(3 * n);
    std::vector<int> hp(3 * n);
    CK(hipMemcpyAsync(hs.data(), c->track0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipMemcpyAsync(hp.data(), c->pbc, sizeof(int) * 3 * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    for (size_t o = 0; o < n; ++o)
        for (int d = 0; d < 3; ++d) {
            if (start) start[3 * o + d] = hs[(size_t)d * n + o];
            if (crossings) crossings[3 * o + d] = hp[(size_t)d * n + o];
        }
    return 0;
}

int psim_gpu_set_tracker(psim_ctx *c, const double *start, const int *crossings)
{
    if (!c || !c->have_state || !start || !crossings) return PSIM_E_ARG;
    const size_t n = (size_t)c->n;
    if (!c->pbc) {
        int rc = dalloc(&c->pbc, 3 * n);
        if (!rc) rc = dalloc(&c->track0, 3 * n);
        if (rc) return rc;
    }
    std::vector<double> hs(3 * n);
    std::vector<int> hp(3 * n);
    for (size_t o = 0; o < n; ++o)
        for (int d = 0; d < 3; ++d) {
            hs[(size_t)d * n + o] = start[3 * o + d];
            hp[(size_t)d * n + o] = crossings[3 * o + d];
        }
    CK(hipMemcpyAsync(c->track0, hs.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipMemcpyAsync(c->pbc, hp.data(), sizeof(int) * 3 * n, hipMemcpyHostToDevice,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

int psim_gpu_observables(psim_ctx *c, double contact_r, psim_observables *out,
                         int *coord_by_orig)
{
    if (!c || !out || !c->have_state || !(contact_r >= 0.0)) return PSIM_E_ARG;
    const int n = c->n;
    const unsigned nb = blocks_for(n, kAnaBlock);
    int rc = psim::eng::settle(c);   // the mean displacement reads pos0
    if (rc) return rc;
    rc = ensure_work(c, (size_t)nb * 4, (size_t)n);
    if (rc) return rc;
    k_observables<<<nb, kAnaBlock, 0, c->stream>>>(c->geo, c->pos, c->pos0, c->vel, c->id,
                                                   c->cstart, c->track0, c->pbc, contact_r,
                                                   c->iwork, c->work);
    CK(hipGetLastError());
    std::vector<double> part((size_t)nb * 4);
    CK(hipMemcpyAsync(part.data(), c->work, sizeof(double) * part.size(),
                      hipMemcpyDeviceToHost, c->stream));
    if (coord_by_orig)
        CK(hipMemcpyAsync(coord_by_orig, c->iwork, sizeof(int) * n, hipMemcpyDeviceToHost,
                          c->stream));
    CK(hipStreamSynchronize(c->stream));
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (unsigned b = 0; b < nb; ++b)
        for (int q = 0; q < 4; ++q) sum[q] += part[(size_t)b * 4 + q];
    out->temperature = sum[0] / 3.0 / n;   // vtot / 3.0 / nParticles
    out->mean_disp2 = sum[1] / (double)n;
    out->msd = c->track0 ? sum[2] / n : std::nan("");
    out->contacts = (long long)std::llround(sum[3]);
    out->mean_coordination = (double)out->contacts / (double)n;
    out->time = c->time;
    out->step = c->step_index;
    return 0;
}

int psim_gpu_clusters(psim_ctx *c, double contact_r, int min_size, int *labels_by_orig,
                      int *n_clusters, int *mean_size)
{
    if (!c || !c->have_state || !(contact_r > 0.0)) return PSIM_E_ARG;
    const int n = c->n;
    const unsigned nb = blocks_for(n, kAnaBlock);
    int rc = ensure_work(c, 1, (size_t)n + 1);
    if (rc) return rc;
    int *lab = c->iwork, *changed = c->iwork + n;
    k_cc_init<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, lab);
    CK(hipGetLastError());
    // labels only decrease and each round lowers at least one: at most n rounds
    for (int it = 0; it <= n; ++it) {
        CK(hipMemsetAsync(changed, 0, sizeof(int), c->stream));
        k_cc_hook<<<nb, kAnaBlock, 0, c->stream>>>(c->geo, c->pos, c->id, c->cstart, contact_r,
                                                   lab, changed);
        k_cc_jump<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, lab);
        CK(hipGetLastError());
        int hc = 0;
        CK(hipMemcpyAsync(&hc, changed, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        CK(hipStreamSynchronize(c->stream));
        if (!hc) break;
    }
    std::vector<int> lh(n);
    CK(hipMemcpyAsync(lh.data(), lab, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    if (labels_by_orig) std::memcpy(labels_by_orig, lh.data(), sizeof(int) * n);
    // writeClusters (analysisCluster.cpp:86-100): clusters of more than
    // min_size - 1 members, average size by integer division
    std::vector<int> size(n, 0);
    for (int o = 0; o < n; ++o) size[lh[o]]++;
    long long total = 0;
    int count = 0;
    for (int o = 0; o < n; ++o)
        if (size[o] >= min_size && size[o] > 0) {
            total += size[o];
            ++count;
        }
    if (n_clusters) *n_clusters = count;
    if (mean_size) *mean_size = count ? (int)(total / count) : 0;
    return 0;
}

int psim_gpu_pair_histogram(psim_ctx *c, double r_lo, double width, int nbins, double sigma,
                            double *hist)
{
    if (!c || !hist || !c->have_state) return PSIM_E_ARG;
    int rc = check_hist_args(c, r_lo, width, nbins, sigma);
    if (rc) return rc;
    HistGeo h{};
    rc = build_hist_grid(c, r_lo, width, nbins, sigma, h);
    if (rc) return rc;
    const int n = c->n;
    if (h.S && sigma > 0.0) {
        // per-block partials, summed in block order (reproducible)
        const unsigned nblk = blocks_for((long long)n * kHistSplit, kAnaBlock);
        rc = ensure_work(c, (size_t)nblk * nbins + nbins, 1);
        if (rc) return rc;
        double *part = c->work, *out = c->work + (size_t)nblk * nbins;
        const size_t lds = sizeof(double) * (kHistCap * kAnaBlock + (kAnaBlock / 64) * nbins);
        k_pair_hist_grid<kHistSplit><<<nblk, kAnaBlock, lds, c->stream>>>(h, c->hg.hpos,
                                                                        c->hg.cstart, part);
        k_hist_sum<<<nbins, 256, 0, c->stream>>>(part, (int)nblk, nbins, out);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(hist, out, sizeof(double) * nbins, hipMemcpyDeviceToHost,
                          c->stream));
        CK(hipStreamSynchronize(c->stream));
        return 0;
    }
    rc = ensure_work(c, (size_t)nbins, 1);
    if (rc) return rc;
    CK(hipMemsetAsync(c->work, 0, sizeof(double) * nbins, c->stream));
    k_pair_hist<<<blocks_for(n, kAnaBlock), kAnaBlock, 0, c->stream>>>(
        h, c->pos, c->hg.perm, c->hg.cstart, c->work);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(hist, c->work, sizeof(double) * nbins, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    return 0;
}

int psim_gpu_pair_histogram_grad(psim_ctx *c, double r_lo, double width, int nbins,
                                 double sigma, const double *w, double *grad)
{
    if (!c || !w || !grad || !c->have_state || !(sigma > 0.0)) return PSIM_E_ARG;
    int rc = check_hist_args(c, r_lo, width, nbins, sigma);
    if (rc) return rc;
    HistGeo h{};
    rc = build_hist_grid(c, r_lo, width, nbins, sigma, h);
    if (rc) return rc;
    const int n = c->n;
    rc = ensure_work(c, (size_t)nbins + 3 * (size_t)n, 1);
    if (rc) return rc;
    double *dw = c->work, *dg = c->work + nbins;
    CK(hipMemcpyAsync(dw, w, sizeof(double) * nbins, hipMemcpyHostToDevice, c->stream));
    if (h.S) {
        const size_t lds = sizeof(double) * (nbins + 1);
        k_pair_hist_grad_grid<kHistSplit>
            <<<blocks_for((long long)n * kHistSplit, kAnaBlock), kAnaBlock, lds, c->stream>>>(
                h, c->hg.hpos, c->hg.perm, c->hg.cstart, c->id, dw, dg);
    } else {
        k_pair_hist_grad<<<blocks_for(n, kAnaBlock), kAnaBlock, 0, c->stream>>>(
            h, c->pos, c->hg.perm, c->hg.cstart, c->id, dw, dg);
    }
    CK(hipGetLastError());
    std::vector<double> hg(3 * (size_t)n);
    CK(hipMemcpyAsync(hg.data(), dg, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                      c->stream));
    CK(hipStreamSynchronize(c->stream));
    for (int o = 0; o < n; ++o)
        for (int d = 0; d < 3; ++d) grad[3 * (size_t)o + d] = hg[(size_t)d * n + o];
    return 0;
}

}  // extern "C"
