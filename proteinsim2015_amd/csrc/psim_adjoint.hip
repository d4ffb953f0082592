// psim_adjoint.hip -- reverse-mode derivative of a G&B trajectory with
// respect to the pair-force parameters, for structure fitting.
//
// With fixed noise, the positions follow (brownianIntegrator.cpp:191-296)
//   x_1     = x_0 + v_0 coEff1 dt + alpha_0 F(x_0) + noise            (first step)
//   x_{n+1} = (1+c0) x_n - c0 x_{n-1} + (c1+c2) F(x_n) - c2 F(x_{n-1}) + noise
// (c0 = coEff0, c1 = dt^2 coEff1 / m, c2 = dt^2 coEff2 / m,
// alpha_0 = dt^2 coEff3 / m; the periodic wrap is a translation and the
// velocity step leaves the positions as setPos does).  For a loss L(x_K),
// with a_n = dL/dx_n and a_{K+1} = 0:
//   lambda_n = alpha_n a_{n+1} - c2 a_{n+2}       (alpha_n = c1 + c2 after step 0)
//   a_n      = beta_n a_{n+1} - c0 a_{n+2} + J_n lambda_n   (beta_0 = 1 on a
//                                                           first step, else 1 + c0)
//   dL/dtheta = sum_n lambda_n . dF(x_n)/dtheta
// J_n = dF/dx at x_n (symmetric): for a pair force F_i = sum_j g(r) d_ij
// with d_ij = x_j - x_i, (J lambda)_i = sum_j g (lambda_j - lambda_i)
//   + (g'/r) (d . (lambda_j - lambda_i)) d.
// By default the derivative is pathwise: the set of pairs inside the cutoff
// is the one the forward trajectory used (a change of the cutoff moves no
// pair across it), which is the exact derivative of the discrete trajectory
// wherever no pair sits exactly at the cutoff.  The force jumps at the cutoff
// (fNet(c) != 0), so the loss of a large system is a staircase in c whose
// secant slope holds the pairs that cross.  With a cutoff window w > 0 the
// cutoff derivative also carries that boundary term,
//   dF_i/dc += sum_j delta(r_ij - c) fNet(r_ij) d_ij / r_ij,
// with delta replaced by the box 1{|r - c| < w} / (2 w): the linearised
// secant over [c - w, c + w].
//
// Memory: a checkpoint of the full state every `every` steps; the backward
// pass recomputes each segment from its checkpoint (the in-kernel Philox
// noise is keyed by step and particle, so the recomputation is bit-identical)
// and records the positions and cell tables of that segment only.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "psim_ctx.hpp"
#include "psim_pairs.hpp"

using namespace psim;
using eng::blocks_for;
using eng::dalloc;

namespace {

constexpr int kAdjBlock = 256;
constexpr int kAdjSmall = 256 * 1024;   // below this N, 8 lanes per particle
constexpr int kNTheta = 5;   // AO wellDepth, AO cutOff, LJ wellDepth, yukawaStrength, debyeLength

// Host-side derivatives of the pair constants (AOPotential.cpp:50-55).
struct AdjPot {
    double neg_kt;                      // -kT
    double ao_dc1_dphi, ao_dc2_dphi;    // dC1/dwellDepth, dC2/dwellDepth
    double ao_dc1_dc, ao_dc2_dc;        // dC1/dcutOff, dC2/dcutOff
    // cutoff window: pairs with |r - c| < win add fNet / (2 win) to the
    // cutoff derivative; win_hi2 = (c + win)^2 (0 without a window)
    double cut, win, win_hi2, win_scale;
};

AdjPot make_adj(const psim_params &P, double win)
{
    AdjPot q{};
    q.cut = P.cutoff;
    q.win = win;
    q.win_hi2 = win > 0.0 ? (P.cutoff + win) * (P.cutoff + win) : 0.0;
    q.win_scale = win > 0.0 ? 0.5 / win : 0.0;
    q.neg_kt = -P.kT;
    const double rc = P.cutoff;
    const double K = -P.kT * std::pow(rc / (rc - 1.0), 3);
    const double a1 = K * P.ao_well_depth;
    const double a2 = -3.0 / (2.0 * rc);
    const double a3 = 1.0 / (2.0 * rc * rc * rc);
    q.ao_dc1_dphi = -K * a2;
    q.ao_dc2_dphi = -3.0 * K * a3;
    const double da1 = -3.0 * a1 / (rc * (rc - 1.0));
    const double da2 = 3.0 / (2.0 * rc * rc);
    const double da3 = -3.0 / (2.0 * rc * rc * rc * rc);
    q.ao_dc1_dc = -(da1 * a2 + a1 * da2);
    q.ao_dc2_dc = -3.0 * (da1 * a3 + a1 * da3);
    return q;
}

// fNet(r), dfNet/dr and dfNet/dtheta of the enabled modules (the forward
// expressions of pair_fr, LJPotential.cpp:97-109, AOPotential.cpp:95-102,
// Calibration.cpp:75-81).
template <int MASK>
__device__ __forceinline__ void pair_fnet_d(const Pot &p, const AdjPot &q, double r,
                                            double rinv, double size, double &f, double &fp,
                                            double (&dth)[kNTheta])
{
    f = 0.0;
    fp = 0.0;
#pragma unroll
    for (int k = 0; k < kNTheta; ++k) dth[k] = 0.0;
    const double r2inv = rinv * rinv;
    if (MASK & F_CAL) {
        const double r37 = pow_const<37>(rinv);
        f += -(36.0 * r37);
        fp += 1332.0 * r37 * rinv;
    }
    if (MASK & F_LJ) {
        const double n = (double)p.lj_num;
        const double sr = size * rinv;
        const double lj = lj_power_x(sr * sr, p.lj_num, size, rinv);
        const double attract = fma(2.0, lj, -1.0) * (p.lj_four_n * rinv * lj);
        const double dattract = -p.lj_four_n * lj * r2inv * (n * fma(4.0, lj, -1.0) + fma(2.0, lj, -1.0));
        const double lam = p.debye_len;
        const double e = exp_neg(r * p.debye_inv);
        const double yb = e * (r2inv * (lam + r));   // repel / yukawaStrength
        const double repel = yb * p.yuk_a;
        const double drepel = p.yuk_a * e * r2inv * (-r * p.debye_inv - 2.0 * lam * rinv - 2.0);
        f += p.lj_neg_kt_eps * (attract + repel);
        fp += p.lj_neg_kt_eps * (dattract + drepel);
        dth[2] += q.neg_kt * (attract + repel);
        dth[3] += p.lj_neg_kt_eps * yb;
        dth[4] += p.lj_neg_kt_eps * p.yuk_a * e * r2inv * (r * (lam + r) * p.debye_inv * p.debye_inv + 1.0);
    }
    if (MASK & F_AO) {
        const double r38 = pow_const<19>(r2inv);
        f += -(36.0 * r38 + p.ao_c1 * rinv + p.ao_c2 * r);
        fp += 1368.0 * r38 * rinv + p.ao_c1 * r2inv - p.ao_c2;
        dth[0] += -(q.ao_dc1_dphi * rinv + q.ao_dc2_dphi * r);
        dth[1] += -(q.ao_dc1_dc * rinv + q.ao_dc2_dc * r);
    }
}

// Pairs of slot s inside the cutoff, with the forward kernels' neighbour
// test and displacement (psim_pairs.hpp gather_force / drain_hits), as
// fn(j, dx, dy, dz, r, 1/r, true); with win_hi2 > 0 also the pairs outside
// it with r^2 < win_hi2, as fn(..., false).  Lane `sub` of the SPLIT lanes
// that share the particle visits the x-runs sub, sub + SPLIT, ... of the 9
// (gather_force's walk).
template <bool EXACT, int SPLIT, class Fn>
__device__ void visit_pairs(const Geo &g, const Pot &p, const double *pos, const int *cstart,
                            int s, int sub, double win_hi2, Fn fn)
{
    const int N = g.n, S = g.S;
    const double xi = pos[s], yi = pos[N + s], zi = pos[2 * N + s];
    const int cx = cell_1d(xi, g.cell, S), cy = cell_1d(yi, g.cell, S),
              cz = cell_1d(zi, g.cell, S);
    const StartTab tab{cstart};
    for (int q = sub; q < 9; q += SPLIT) {
        int ry = cy + q / 3 - 1, ky = 0;
        if (ry < 0) { ry += S; ky = 1; } else if (ry >= S) { ry -= S; ky = 2; }
        int rz = cz + q % 3 - 1, kz = 0;
        if (rz < 0) { rz += S; kz = 1; } else if (rz >= S) { rz -= S; kz = 2; }
        const double ys = shifted(yi, ky, g.L), zs = shifted(zi, kz, g.L);
        const XRun xr = x_run(tab, g.c0 + S * ry + g.S2 * rz, cx, S);
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
            const double xs = shifted(xi, sg ? xr.k1 : xr.k0, g.L);
            const int e = sg ? xr.e1 : xr.e0;
            for (int j = sg ? xr.b1 : xr.b0; j < e; ++j) {
                if (j == s) continue;
                const double xj = pos[j], yj = pos[N + j], zj = pos[2 * N + j];
                double dx, dy, dz, r2, r, rinv;
                bool inside = true;
                if (!EXACT) {
                    dx = xj - xs;
                    dy = yj - ys;
                    dz = zj - zs;
                    r2 = fma(dx, dx, fma(dy, dy, dz * dz));
                    if (!(r2 < p.rc2_hi) && !(r2 < win_hi2)) continue;
                    if (r2 >= p.rc2_lo)
                        inside = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc) < p.rc2;
                    if (!inside && !(r2 < win_hi2)) continue;
                    rinv = rsqrt_nr(r2);
                    r = r2 * rinv;
                } else {
                    r2 = pbc_dist_ref(xi, yi, zi, xj, yj, zj, g.L, g.half_pbc);
                    inside = r2 < p.rc2;
                    if (!inside && !(r2 < win_hi2)) continue;
                    r = sqrt(r2);
                    rinv = 1.0 / r;
                    dx = min_image_uv(xj - xi, g.L, g.half_uv);
                    dy = min_image_uv(yj - yi, g.L, g.half_uv);
                    dz = min_image_uv(zj - zi, g.L, g.half_uv);
                }
                fn(j, dx, dy, dz, r, rinv, inside);
            }
        }
    }
}

// One backward step at the recorded state x_n (SPLIT lanes per particle).
// lam (lambda_n) is in the slot order of that state, so a neighbour's value
// is read at its slot; lambda_{n-1} goes to lam_prev in the slot order of
// the state before (inv: original index -> slot there) or, with inv null, by
// original index.  The a_n are kept by original index.
template <int MASK, bool EXACT, int BLOCK, int SPLIT>
__global__ __launch_bounds__(BLOCK) void k_adj_step(
    Geo g, Pot p, AdjPot q, double radius, const double *pos, const int *id,
    const int *cstart, const double *lam, const double *a1, const double *a2, double *a0,
    double *lam_prev, const int *inv, double beta, double coeff0, double alpha_prev,
    double c2, double *tpart)
{
    const int gt = xcd_block() * BLOCK + threadIdx.x;
    const int s = gt / SPLIT, sub = gt % SPLIT;
    const int N = g.n;
    double th[kNTheta] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double jx = 0.0, jy = 0.0, jz = 0.0;
    if (s < N) {   // whole lane groups
        const double lx = lam[s], ly = lam[N + s], lz = lam[2 * N + s];
        const double size = 2.0 * radius;
        visit_pairs<EXACT, SPLIT>(g, p, pos, cstart, s, sub, q.win_hi2,
                                  [&](int j, double dx, double dy, double dz, double r,
                                      double rinv, bool inside) {
            double f, fp, dth[kNTheta];
            pair_fnet_d<MASK>(p, q, r, rinv, size, f, fp, dth);
            const double ld = (lx * dx + ly * dy + lz * dz) * rinv;
            if (fabs(r - q.cut) < q.win) th[1] += f * q.win_scale * ld;
            if (!inside) return;
            const double ex = lam[j] - lx, ey = lam[N + j] - ly, ez = lam[2 * N + j] - lz;
            const double gr = f * rinv;
            const double gp = (fp - gr) * rinv;
            const double dd = dx * ex + dy * ey + dz * ez;
            const double k2 = gp * rinv * dd;
            jx += gr * ex + k2 * dx;
            jy += gr * ey + k2 * dy;
            jz += gr * ez + k2 * dz;
#pragma unroll
            for (int k = 0; k < kNTheta; ++k) th[k] += dth[k] * ld;
        });
#pragma unroll
        for (int d = 1; d < SPLIT; d <<= 1) {
            jx += __shfl_xor(jx, d, 64);
            jy += __shfl_xor(jy, d, 64);
            jz += __shfl_xor(jz, d, 64);
        }
    }
    if (s < N && sub == 0) {
        const int o = id[s];
        const double n0x = beta * a1[o] - coeff0 * a2[o] + jx;
        const double n0y = beta * a1[N + o] - coeff0 * a2[N + o] + jy;
        const double n0z = beta * a1[2 * N + o] - coeff0 * a2[2 * N + o] + jz;
        a0[o] = n0x;
        a0[N + o] = n0y;
        a0[2 * N + o] = n0z;
        if (lam_prev) {
            const int w = inv ? inv[o] : o;
            lam_prev[w] = alpha_prev * n0x - c2 * a1[o];
            lam_prev[N + w] = alpha_prev * n0y - c2 * a1[N + o];
            lam_prev[2 * N + w] = alpha_prev * n0z - c2 * a1[2 * N + o];
        }
    }
    __shared__ double red[kNTheta][BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kNTheta; ++k) {
        double x = th[k];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) x += __shfl_down(x, d, 64);
        if (lane == 0) red[k][w] = x;
    }
    __syncthreads();
    if (threadIdx.x < kNTheta) {
        double t = 0.0;
        for (int k = 0; k < BLOCK / 64; ++k) t += red[threadIdx.x][k];
        tpart[(size_t)blockIdx.x * kNTheta + threadIdx.x] = t;
    }
}

// acc[k] = sum of tpart[b][k] over all nb block partials of all steps:
// block k, lane t sums b = t, t + 256, ... in order, then a fixed tree
// (deterministic)
__global__ __launch_bounds__(256) void k_theta_sum(const double *tpart, long long nb,
                                                   double *acc)
{
    __shared__ double red[256];
    const int k = blockIdx.x, t = threadIdx.x;
    double v = 0.0;
    for (long long b = t; b < nb; b += 256) v += tpart[(size_t)b * kNTheta + k];
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) acc[k] = red[0];
}

// First stage of the dL/dtheta sum: block (c, k) sums the block partials
// b of chunk c (b = c, c + kThetaChunks, ...) of parameter k into
// part[c][k], which k_theta_sum then sums in chunk order (deterministic).
constexpr int kThetaChunks = 256;
__global__ __launch_bounds__(256) void k_theta_chunks(const double *tpart, long long nb,
                                                      double *part)
{
    __shared__ double red[256];
    const int c = blockIdx.x, k = blockIdx.y, t = threadIdx.x;
    double v = 0.0;
    for (long long b = c + (long long)kThetaChunks * t; b < nb; b += (long long)kThetaChunks * 256)
        v += tpart[(size_t)b * kNTheta + k];
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) part[(size_t)c * kNTheta + k] = red[0];
}

// Several device-to-device copies in one launch (the state arrays of a
// checkpoint or of a recorded step): segment k copies words[k] 4-byte words.
constexpr int kMaxSeg = 8;
struct CopySegs {
    const unsigned *src[kMaxSeg];
    unsigned *dst[kMaxSeg];
    long long end[kMaxSeg];   // cumulative word counts
    int nseg;
};

__global__ __launch_bounds__(256) void k_copy_segs(CopySegs cs)
{
    const long long total = cs.end[cs.nseg - 1];
    for (long long w = blockIdx.x * 256LL + threadIdx.x; w < total;
         w += (long long)gridDim.x * 256) {
        int k = 0;
        while (w >= cs.end[k]) ++k;
        const long long off = w - (k ? cs.end[k - 1] : 0);
        cs.dst[k][off] = cs.src[k][off];
    }
}

// inv[id[s]] = s: the slot of each original index
__global__ void k_inv(int n, const int *id, int *inv)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[id[s]] = s;
}

// out[d][s] = in[d][id[s]]: a vector by original index into slot order
__global__ void k_gather_slots(int n, const int *id, const double *in, double *out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int o = id[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) out[(size_t)d * n + s] = in[(size_t)d * n + o];
}

// lam = alpha * a1 (the last step: a_{K+1} = 0)
__global__ void k_scale(int n3, double alpha, const double *a, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) out[i] = alpha * a[i];
}

}  // namespace

struct psim_adjoint {
    psim_ctx *c = nullptr;
    int K = 0, every = 0, ncp = 0;
    struct Meta {
        double time;
        int vel_counter;
        unsigned long long step;
    };
    std::vector<Meta> meta;        // ncp + 1 (the last = final state)
    double *cp_d = nullptr;        // (ncp + 1) x 15N: pos, pos0, frc, mem, vel
    int *cp_i = nullptr;           // (ncp + 1) x (N + ncell + 1): id, cstart
    // records of the states of one segment (every of them), or of the whole
    // trajectory (K, record_all: written by the forward pass itself)
    double *rec_pos = nullptr;     // nrec x 3N
    int *rec_i = nullptr;          // nrec x (N + ncell + 1)
    int *rec_inv = nullptr;        // nrec x N: original index -> slot of each record
    int nrec = 0;
    bool record_all = false;
    double *a[3] = {nullptr, nullptr, nullptr};
    double *lam[2] = {nullptr, nullptr};
    double *tpart = nullptr, *acc = nullptr;
    std::vector<int> first;        // per step: first-step flag
    bool have_forward = false;
    psim_params prm_used{};
    double win = 0.0;   // cutoff window (0: pathwise)
};

namespace {

size_t cp_stride_d(const psim_ctx *c) { return 15 * (size_t)c->n; }
size_t cp_stride_i(const psim_ctx *c) { return (size_t)c->n + c->ncell + 1; }

// Launch the copies collected in cs (one kernel; the grid covers the
// largest system in a few passes).
int copy_segs(psim_ctx *c, CopySegs &cs)
{
    const long long total = cs.end[cs.nseg - 1];
    const long long blocks = (total + 255) / 256;
    k_copy_segs<<<(unsigned)(blocks < 4096 ? blocks : 4096), 256, 0, c->stream>>>(cs);
    CK(hipGetLastError());
    return 0;
}

void add_seg(CopySegs &cs, const void *src, void *dst, size_t bytes)
{
    cs.src[cs.nseg] = (const unsigned *)src;
    cs.dst[cs.nseg] = (unsigned *)dst;
    cs.end[cs.nseg] = (cs.nseg ? cs.end[cs.nseg - 1] : 0) + (long long)(bytes / 4);
    ++cs.nseg;
}

// Record buffers for nrec states (the previous ones are freed).
int alloc_records(psim_adjoint *A, int nrec)
{
    psim_ctx *c = A->c;
    void *old[] = {A->rec_pos, A->rec_i, A->rec_inv};
    for (void *b : old)
        if (b) (void)hipFree(b);
    A->rec_pos = nullptr;
    A->rec_i = nullptr;
    A->rec_inv = nullptr;
    A->nrec = 0;
    int rc = dalloc(&A->rec_pos, (size_t)nrec * 3 * (size_t)c->n);
    if (!rc) rc = dalloc(&A->rec_i, (size_t)nrec * cp_stride_i(c));
    if (!rc) rc = dalloc(&A->rec_inv, (size_t)nrec * c->n);
    if (!rc) A->nrec = nrec;
    return rc;
}

// Record r of the records: the state's positions, ids, cell table and
// inverse slot map.
psim_ctx::Record record_at(psim_adjoint *A, int r)
{
    psim_ctx *c = A->c;
    const size_t si = cp_stride_i(c);
    psim_ctx::Record R{};
    R.pos = A->rec_pos + (size_t)r * 3 * (size_t)c->n;
    R.id = A->rec_i + (size_t)r * si;
    R.cstart = A->rec_i + (size_t)r * si + c->n;
    R.inv = A->rec_inv + (size_t)r * c->n;
    return R;
}

// The context's current state (settled: current slot order) as record r.
int record_now(psim_adjoint *A, int r)
{
    psim_ctx *c = A->c;
    const psim_ctx::Record R = record_at(A, r);
    CopySegs cs{};
    add_seg(cs, c->pos, R.pos, sizeof(double) * 3 * (size_t)c->n);
    add_seg(cs, c->id, R.id, sizeof(int) * c->n);
    add_seg(cs, c->cstart, R.cstart, sizeof(int) * (c->ncell + 1));
    int rc = copy_segs(c, cs);
    if (rc) return rc;
    k_inv<<<blocks_for(c->n, 256), 256, 0, c->stream>>>(c->n, R.id, R.inv);
    CK(hipGetLastError());
    return 0;
}

int save_cp(psim_adjoint *A, int k)
{
    psim_ctx *c = A->c;
    const size_t n3 = 3 * (size_t)c->n;
    double *d = A->cp_d + (size_t)k * cp_stride_d(c);
    int *ip = A->cp_i + (size_t)k * cp_stride_i(c);
    int rc = eng::settle(c);   // checkpoints hold the state in the current slot order
    if (rc) return rc;
    const double *src[5] = {c->pos, c->pos0, c->frc, c->mem, c->vel};
    CopySegs cs{};
    for (int q = 0; q < 5; ++q) add_seg(cs, src[q], d + q * n3, sizeof(double) * n3);
    add_seg(cs, c->id, ip, sizeof(int) * c->n);
    add_seg(cs, c->cstart, ip + c->n, sizeof(int) * (c->ncell + 1));
    A->meta[k] = {c->time, c->vel_counter, c->step_index};
    return copy_segs(c, cs);
}

int restore_cp(psim_adjoint *A, int k)
{
    psim_ctx *c = A->c;
    const size_t n3 = 3 * (size_t)c->n;
    const double *d = A->cp_d + (size_t)k * cp_stride_d(c);
    const int *ip = A->cp_i + (size_t)k * cp_stride_i(c);
    double *dst[5] = {c->pos, c->pos0, c->frc, c->mem, c->vel};
    CopySegs cs{};
    for (int q = 0; q < 5; ++q) add_seg(cs, d + q * n3, dst[q], sizeof(double) * n3);
    add_seg(cs, ip, c->id, sizeof(int) * c->n);
    add_seg(cs, ip + c->n, c->cstart, sizeof(int) * (c->ncell + 1));
    int rc = copy_segs(c, cs);
    if (rc) return rc;
    c->lagged = false;
    c->time = A->meta[k].time;
    c->vel_counter = A->meta[k].vel_counter;
    c->step_index = A->meta[k].step;
    return 0;
}

// The displacement tracker counts crossings of the trajectory that is
// actually advanced; the adjoint's replays are not, so it is paused.
struct TrackerPause {
    psim_ctx *c;
    int *saved;
    explicit TrackerPause(psim_ctx *cc) : c(cc), saved(cc->pbc) { c->pbc = nullptr; }
    ~TrackerPause() { c->pbc = saved; }
};

bool same_forces(const psim_params &a, const psim_params &b)
{
    return a.force_mask == b.force_mask && a.lj_num == b.lj_num && a.cutoff == b.cutoff &&
           a.lj_well_depth == b.lj_well_depth && a.yukawa_strength == b.yukawa_strength &&
           a.debye_length == b.debye_length && a.ao_well_depth == b.ao_well_depth;
}

int adj_block(const psim_ctx *) { return kAdjBlock; }
// lanes per particle: small systems spread each particle's 27 cells over 8
// lanes, so that every CU gets work and each lane waits on fewer lookups
#ifndef PSIM_ADJ_SPLIT
#define PSIM_ADJ_SPLIT 8
#endif
int adj_split(const psim_ctx *c) { return c->n < kAdjSmall ? PSIM_ADJ_SPLIT : 1; }
unsigned adj_blocks(const psim_ctx *c)
{
    return xcd_grid(blocks_for((long long)c->n * adj_split(c), adj_block(c)));
}

// Neighbour-set statistics of visit_pairs with the backward step's launch
// shape (each accepted pair counted; see NbrStats).
template <bool EXACT, int BLOCK, int SPLIT>
__global__ __launch_bounds__(BLOCK) void k_adj_stats(Geo g, Pot p, const double *pos,
                                                     const int *id, const int *cstart,
                                                     NbrStatsOut so)
{
    const int gt = xcd_block() * BLOCK + threadIdx.x;
    const int s = gt / SPLIT, sub = gt % SPLIT;
    if (s >= g.n) return;   // whole lane groups
    NbrStats st;
    visit_pairs<EXACT, SPLIT>(g, p, pos, cstart, s, sub, 0.0,
                              [&](int j, double, double, double, double, double, bool) {
                                  st.add(id[j]);
                              });
    split_stats<SPLIT>(st);
    if (sub == 0) {
        so.count[id[s]] = st.count;
        so.hash[id[s]] = st.hash;
    }
}

// one backward step; its block partials of dL/dtheta go to tp
template <int MASK>
void launch_adj_mask(psim_adjoint *A, const AdjPot &q, const double *pos, const int *id,
                     const int *cs, const double *lam, const double *a1, const double *a2,
                     double *a0, double *lam_prev, const int *inv, double beta,
                     double alpha_prev, double *tp)
{
    psim_ctx *c = A->c;
    const unsigned nb = adj_blocks(c);
#define PSIM_ADJ_LAUNCH(EX, BL, SP)                                                        \
    k_adj_step<MASK, EX, BL, SP><<<nb, BL, 0, c->stream>>>(                                \
        c->geo, c->pot, q, c->prm.radius, pos, id, cs, lam, a1, a2, a0, lam_prev, inv,     \
        beta, c->gb.coeff0, alpha_prev, c->gb.c2, tp)
    const bool split = adj_split(c) > 1;
    if (c->geo.exact_mi) {
        if (split) PSIM_ADJ_LAUNCH(true, kAdjBlock, PSIM_ADJ_SPLIT);
        else PSIM_ADJ_LAUNCH(true, kAdjBlock, 1);
    } else {
        if (split) PSIM_ADJ_LAUNCH(false, kAdjBlock, PSIM_ADJ_SPLIT);
        else PSIM_ADJ_LAUNCH(false, kAdjBlock, 1);
    }
#undef PSIM_ADJ_LAUNCH
}

}  // namespace

namespace psim {
namespace eng {

void launch_adj_stats(psim_ctx *c, NbrStatsOut so)
{
    const unsigned nb = adj_blocks(c);
    const bool split = adj_split(c) > 1;
#define PSIM_ADJ_STATS(EX, SP)                                                             \
    k_adj_stats<EX, kAdjBlock, SP><<<nb, kAdjBlock, 0, c->stream>>>(c->geo, c->pot, c->pos, \
                                                                    c->id, c->cstart, so)
    if (c->geo.exact_mi) {
        if (split) PSIM_ADJ_STATS(true, PSIM_ADJ_SPLIT);
        else PSIM_ADJ_STATS(true, 1);
    } else {
        if (split) PSIM_ADJ_STATS(false, PSIM_ADJ_SPLIT);
        else PSIM_ADJ_STATS(false, 1);
    }
#undef PSIM_ADJ_STATS
}

}  // namespace eng
}  // namespace psim

extern "C" {

int psim_adj_create(psim_ctx *c, int nsteps, int checkpoint_every, psim_adjoint **out)
{
    if (!c || !out || !c->have_state || nsteps <= 0 || checkpoint_every <= 0) return PSIM_E_ARG;
    *out = nullptr;
    psim_adjoint *A = new (std::nothrow) psim_adjoint();
    if (!A) return PSIM_E_NOMEM;
    A->c = c;
    A->K = nsteps;
    A->every = checkpoint_every < nsteps ? checkpoint_every : nsteps;
    A->ncp = (nsteps + A->every - 1) / A->every;
    A->meta.resize(A->ncp + 1);
    A->first.assign(nsteps, 0);
    const size_t n3 = 3 * (size_t)c->n;
    // block partials of dL/dtheta for every step, summed once at the end
    const size_t nb = adj_blocks(c);
    int rc = 0;
    if (!rc) rc = dalloc(&A->cp_d, (size_t)(A->ncp + 1) * cp_stride_d(c));
    if (!rc) rc = dalloc(&A->cp_i, (size_t)(A->ncp + 1) * cp_stride_i(c));
    if (!rc) rc = alloc_records(A, A->every);
    for (int k = 0; k < 3 && !rc; ++k) rc = dalloc(&A->a[k], n3);
    for (int k = 0; k < 2 && !rc; ++k) rc = dalloc(&A->lam[k], n3);
    if (!rc) rc = dalloc(&A->tpart, (size_t)nsteps * nb * kNTheta);
    if (!rc) rc = dalloc(&A->acc, kNTheta * (1 + (size_t)kThetaChunks));
    if (!rc) rc = save_cp(A, 0);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PSIM_E_HIP;
    if (rc) {
        psim_adj_destroy(A);
        return rc;
    }
    *out = A;
    return 0;
}

void psim_adj_destroy(psim_adjoint *A)
{
    if (!A) return;
    if (A->c && A->c->stream) (void)hipStreamSynchronize(A->c->stream);
    void *bufs[] = {A->cp_d, A->cp_i, A->rec_pos, A->rec_i, A->rec_inv, A->a[0], A->a[1], A->a[2],
                    A->lam[0], A->lam[1], A->tpart, A->acc};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete A;
}

int psim_adj_set_record_all(psim_adjoint *A, int on)
{
    if (!A) return PSIM_E_ARG;
    const int want = on ? A->K : A->every;
    if (A->nrec != want) {
        CK(hipStreamSynchronize(A->c->stream));
        const int rc = alloc_records(A, want);
        if (rc) {
            // back to the segment records (smaller) so the object stays usable
            A->record_all = false;
            if (alloc_records(A, A->every)) A->nrec = 0;
            A->have_forward = false;
            return rc;
        }
    }
    A->record_all = on != 0;
    A->have_forward = false;   // the records are made by the next forward
    return 0;
}

int psim_adj_set_cutoff_window(psim_adjoint *A, double w)
{
    if (!A || !(w >= 0.0) || !std::isfinite(w)) return PSIM_E_ARG;
    A->win = w;
    return 0;
}

int psim_adj_set_initial(psim_adjoint *A)
{
    if (!A) return PSIM_E_ARG;
    int rc = save_cp(A, 0);
    if (rc) return rc;
    CK(hipStreamSynchronize(A->c->stream));
    A->have_forward = false;
    return 0;
}

int psim_adj_forward(psim_adjoint *A)
{
    if (!A) return PSIM_E_ARG;
    if (A->nrec == 0) return PSIM_E_NOMEM;
    psim_ctx *c = A->c;
    TrackerPause pause(c);
    int rc = restore_cp(A, 0);
    if (rc) return rc;
    double t = c->time;
    for (int n = 0; n < A->K; ++n) {
        A->first[n] = (t == 0) ? 1 : 0;   // brownianIntegrator.cpp:183
        t += c->prm.dt;
    }
    if (A->record_all) {
        // every state of the trajectory recorded on the way (the cell
        // rebuild of each step writes the next one): no checkpoints between
        // the start and the end, and no recomputation in the backward pass
        if ((rc = record_now(A, 0))) return rc;
        for (int n = 0; n < A->K; ++n) {
            if (n + 1 < A->K) c->rec = record_at(A, n + 1);
            rc = eng::run_steps(c, 1, nullptr);
            c->rec = psim_ctx::Record{};
            if (rc) return rc;
        }
    } else {
        for (int k = 0; k < A->ncp; ++k) {
            if (k > 0 && (rc = save_cp(A, k))) return rc;
            const int n0 = k * A->every;
            const int n1 = n0 + A->every < A->K ? n0 + A->every : A->K;
            if ((rc = eng::run_steps(c, n1 - n0, nullptr))) return rc;
        }
    }
    if ((rc = save_cp(A, A->ncp))) return rc;
    int code = 0;
    if ((rc = eng::read_error(c, &code, nullptr, nullptr))) return rc;
    A->prm_used = c->prm;
    A->have_forward = code == 0;
    return code;
}

int psim_adj_backward(psim_adjoint *A, const double *gx, double *dtheta, double *gx0)
{
    if (!A || !gx || !dtheta) return PSIM_E_ARG;
    psim_ctx *c = A->c;
    if (!A->have_forward || !same_forces(A->prm_used, c->prm)) return PSIM_E_ARG;
    if (A->nrec == 0) return PSIM_E_NOMEM;
    // the window's pairs must lie in the 27 cells the walk visits
    if (A->win > 0.0 && c->prm.cutoff + A->win > c->geo.cell) return PSIM_E_ARG;
    TrackerPause pause(c);
    const int n = c->n;
    const size_t n3 = 3 * (size_t)n;
    const GB &gb = c->gb;
    const double alpha_first = gb.coeff3 * gb.dt * gb.dt * gb.m;
    const double alpha_normal = gb.c1 + gb.c2;
    auto alpha = [&](int step) { return A->first[step] ? alpha_first : alpha_normal; };
    auto beta = [&](int step) { return A->first[step] ? 1.0 : gb.c0; };
    const AdjPot q = make_adj(c->prm, A->win);
    // a_K = gx, a_{K+1} = 0, lambda_{K-1} = alpha_{K-1} a_K
    {
        std::vector<double> h(n3);
        for (int o = 0; o < n; ++o)
            for (int d = 0; d < 3; ++d) h[(size_t)d * n + o] = gx[3 * (size_t)o + d];
        CK(hipMemcpyAsync(A->a[1], h.data(), sizeof(double) * n3, hipMemcpyHostToDevice,
                          c->stream));
        CK(hipStreamSynchronize(c->stream));
    }
    CK(hipMemsetAsync(A->a[2], 0, sizeof(double) * n3, c->stream));
    const size_t nbk = adj_blocks(c);
    k_scale<<<blocks_for((long long)n3, 256), 256, 0, c->stream>>>((int)n3, alpha(A->K - 1),
                                                                   A->a[1], A->lam[0]);
    CK(hipGetLastError());
    double *a0 = A->a[0], *a1 = A->a[1], *a2 = A->a[2];
    double *lam = A->lam[0], *lam_prev = A->lam[1];
    const size_t si = cp_stride_i(c);
    int rc = 0;
    // with every state recorded by the forward pass the trajectory is one
    // segment with nothing to recompute
    const int nseg = A->record_all ? 1 : A->ncp;
    for (int k = nseg - 1; k >= 0; --k) {
        const int n0 = A->record_all ? 0 : k * A->every;
        const int n1 = A->record_all ? A->K : (n0 + A->every < A->K ? n0 + A->every : A->K);
        if (!A->record_all) {
            if ((rc = restore_cp(A, k))) return rc;
            // the checkpoint's state is record 0; each recomputed step's
            // cell rebuild writes the next record itself
            CopySegs cs{};
            add_seg(cs, c->pos, A->rec_pos, sizeof(double) * n3);
            add_seg(cs, c->id, A->rec_i, sizeof(int) * n);
            add_seg(cs, c->cstart, A->rec_i + n, sizeof(int) * (c->ncell + 1));
            if ((rc = copy_segs(c, cs))) return rc;
            k_inv<<<blocks_for(n, 256), 256, 0, c->stream>>>(n, A->rec_i, A->rec_inv);
            for (int s = n0; s + 1 < n1; ++s) {
                c->rec = record_at(A, s - n0 + 1);
                rc = eng::run_steps(c, 1, nullptr);
                c->rec = psim_ctx::Record{};
                if (rc) return rc;
            }
        }
        // lambda_{n1-1} arrives by original index (from the segment after,
        // or the start): into the slot order of the segment's last record
        k_gather_slots<<<blocks_for(n, 256), 256, 0, c->stream>>>(
            n, A->rec_i + (size_t)(n1 - 1 - n0) * si, lam, lam_prev);
        std::swap(lam, lam_prev);
        for (int s = n1 - 1; s >= n0; --s) {
            const int r = s - n0;
            const double *pos = A->rec_pos + (size_t)r * n3;
            const int *id = A->rec_i + (size_t)r * si;
            const int *cs = id + n;
            double *lp = s >= 1 ? lam_prev : nullptr;
            // lambda_{s-1} in the slot order of record r - 1 within the
            // segment, by original index across its start
            const int *inv = r >= 1 ? A->rec_inv + (size_t)(r - 1) * n : nullptr;
            const double ap = s >= 1 ? alpha(s - 1) : 0.0;
            double *tp = A->tpart + (size_t)s * nbk * kNTheta;
            switch (c->pot.mask) {
            case 1: launch_adj_mask<1>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 2: launch_adj_mask<2>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 3: launch_adj_mask<3>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 4: launch_adj_mask<4>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 5: launch_adj_mask<5>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 6: launch_adj_mask<6>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            case 7: launch_adj_mask<7>(A, q, pos, id, cs, lam, a1, a2, a0, lp, inv, beta(s), ap, tp); break;
            default: return PSIM_E_ARG;
            }
            CK(hipGetLastError());
            // rotate: a_{n+2} <- a_{n+1}, a_{n+1} <- a_n; lambda_n <- lambda_{n-1}
            double *t = a2;
            a2 = a1;
            a1 = a0;
            a0 = t;
            t = lam;
            lam = lam_prev;
            lam_prev = t;
        }
    }
    k_theta_chunks<<<dim3(kThetaChunks, kNTheta), 256, 0, c->stream>>>(
        A->tpart, (long long)A->K * nbk, A->acc + kNTheta);
    k_theta_sum<<<kNTheta, 256, 0, c->stream>>>(A->acc + kNTheta, kThetaChunks, A->acc);
    CK(hipGetLastError());
    // a1 now holds a_0 = dL/dx_0
    std::vector<double> h(gx0 ? n3 : 0);
    if (gx0)
        CK(hipMemcpyAsync(h.data(), a1, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    CK(hipMemcpyAsync(dtheta, A->acc, sizeof(double) * kNTheta, hipMemcpyDeviceToHost,
                      c->stream));
    // leave the context at the end of the trajectory
    if ((rc = restore_cp(A, A->ncp))) return rc;
    CK(hipStreamSynchronize(c->stream));
    if (gx0)
        for (int o = 0; o < n; ++o)
            for (int d = 0; d < 3; ++d) gx0[3 * (size_t)o + d] = h[(size_t)d * n + o];
    return 0;
}

}  // extern "C"
