// psim_integrator.hip -- the G&B step over a host's own particle arrays:
// the GPU side of the IIntegrator adaptor (psim_int_*, include/psim_gpu.h).
//
// brownianIntegrator::nextSystem (brownianIntegrator.cpp:181-356) advances
// the particles the host owns, in the host's particle order: firstStep when
// currentTime == 0, else normalStep with velocityStep every velFreq + 1
// steps, each new position stored through particle::setPos (single wrap,
// previous position re-imaged next to the new one, bounds check;
// particle.cpp:61-99).  Here one lane per particle does the same arithmetic
// in the same operation order (the file is built with -ffp-contract=off),
// with the particle's own mass, and keeps memX / memCorr on the device
// between calls.
//
// Noise: PSIM_NOISE_REFERENCE runs the reference's own generator per
// particle on the device -- fastRandom::psdrand's subtractive lagged
// generator (Knuth's ran3 scheme: a 55-entry table, lags 21 and 31,
// modulus 10^9) seeded with threadSeed = seed * (i + 1) in 32-bit
// arithmetic, and g250's polar Box-Muller with the spare deviate kept as a
// float (fastRandom.cpp:40-124, fastRandom.h:16) -- so a run draws the
// reference's streams.  PSIM_NOISE_PHILOX draws Philox4x32-10 normals keyed
// by (seed, call count, particle index) instead.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "psim_ctx.hpp"
#include "psim_params.hpp"

using namespace psim;

namespace {

constexpr int kFrDim = 55;          // table length
constexpr int kFrShort = 21;        // initialisation stride
constexpr int kFrLong = 30;         // refill offset (lag 31)
constexpr int kFrMod = 1000000000;  // modulus
// A polar draw rejects a pair with probability 1 - pi/4; 256 rejections in a
// row do not happen for a working stream.  A stream that re-seeds on every
// call (negative threadSeed, the 32-bit overflow of seed * (i + 1)) repeats
// one uniform and can reject forever: the lane stops and reports instead of
// spinning.
constexpr int kPolarTries = 256;

// Per-particle generator state in device memory: table entry k of particle
// i at tab[k * n + i] (k = 0..55), the read index, flags (1 = seeded,
// 2 = a spare deviate is held) and the spare deviate.
struct FrState {
    int *tab;
    int *jrand;
    int *flags;
    float *spare;
    int n;
};

struct FrLane {
    int jrand;
    int flags;
    float spare;
    bool stuck;
};

__device__ void fr_refill(int *t, int n, int i)
{
    for (int q = 1; q <= kFrDim; ++q) {
        const int inx = (q + kFrLong) % kFrDim;
        int v = t[q * n + i] - t[(inx + 1) * n + i];
        if (v < 0) v += kFrMod;
        t[q * n + i] = v;
    }
}

// fastRandom::psdrand (fastRandom.cpp:40-89)
__device__ double fr_uniform(const FrState &r, FrLane &f, int i, int iseed)
{
    int *t = r.tab;
    const int n = r.n;
    if (!(f.flags & 1) || iseed < 0) {
        // abs() in 32-bit two's complement
        iseed = (int)(iseed < 0 ? 0u - (unsigned)iseed : (unsigned)iseed);
        t[kFrDim * n + i] = iseed;
        int j = iseed, k = 1;
        for (int q = 1; q <= kFrDim - 1; ++q) {
            const int inx = (q * kFrShort) % kFrDim;
            t[inx * n + i] = k;
            k = j - k;
            if (k < 0) k += kFrMod;
            j = t[inx * n + i];
        }
        for (int rep = 0; rep < 3; ++rep) fr_refill(t, n, i);
        f.jrand = 0;
        f.flags |= 1;
    }
    f.jrand += 1;
    if (f.jrand > kFrDim) {
        fr_refill(t, n, i);
        f.jrand = 1;
    }
    return ((double)t[f.jrand * n + i]) / 1.0E9;
}

// fastRandom::g250 with mu = 0, sigma = 1 (fastRandom.cpp:92-124)
__device__ double fr_normal(const FrState &r, FrLane &f, int i, int seed)
{
    if (!(f.flags & 2)) {
        double v1, v2, rsq;
        int tries = 0;
        do {
            if (++tries > kPolarTries) {
                f.stuck = true;
                return 0.0;
            }
            v1 = 2.0 * fr_uniform(r, f, i, seed) - 1.0;
            v2 = 2.0 * fr_uniform(r, f, i, seed) - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while (rsq >= 1.0 || rsq == 0.0);
        const double polar = sqrt(-2.0 * log(rsq) / rsq);
        f.spare = (float)(v1 * polar);
        f.flags |= 2;
        return v2 * polar * 1.0 + 0.0;
    }
    f.flags &= ~2;
    return (double)f.spare * 1.0 + 0.0;
}

struct IntCoef {
    double coeff0, coeff1, coeff2, coeff3;
    double c0;                 // 1 + coEff0
    double sig1, sig2, corr, dev;
    double goy2, goy3, hn;
    double dt, dt_inv;
};

struct IntArgs {
    int n;
    int first;        // currentTime == 0
    int do_vel;       // velocityStep this call
    int noise;        // PSIM_NOISE_*
    int seed;
    unsigned long long philox_seed, call;
    double L;         // boxSize (the reference's int, as a double)
    double *pos, *pos0, *vel;           // n x 3, items order
    const double *frc, *frc0, *mass;    // n x 3, n x 3, n
    double *memx, *memcorr;             // [3][n]
    FrState fr;
    DevErr *err;
};

// particle::setX/Y/Z (particle.cpp:61-99): single wrap, the previous
// position re-imaged next to the new one, bounds check
__device__ __forceinline__ bool set_coord(double *pos, double *pos0, size_t k, double val,
                                          double L)
{
    const double old = pos[k];
    const double x = safe_mod(val, L);
    pos[k] = x;
    pos0[k] = safe_mod0(old, x, L);
    return (x < 0.0) || (x >= L);
}

__global__ __launch_bounds__(256) void k_int_step(IntCoef c, IntArgs a)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int n = a.n;
    const size_t b = 3 * (size_t)i;
    // brownianIntegrator.cpp:204, 236 (int arithmetic)
    const int thread_seed = (int)((unsigned)a.seed * (unsigned)(i + 1));
    FrLane f{a.fr.jrand[i], a.fr.flags[i], a.fr.spare[i], false};
    double z[6];
    if (a.noise == PSIM_NOISE_PHILOX) philox_normals6(a.philox_seed, a.call, (uint32_t)i, z);
    auto draw = [&](int k) {
        return a.noise == PSIM_NOISE_PHILOX ? z[k] : fr_normal(a.fr, f, i, thread_seed);
    };
    const double m = 1.0 / a.mass[i];
    double x[3], x0[3], F[3], F0[3], M[3], Cr[3], xn[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d] = a.pos[b + d];
        x0[d] = a.pos0[b + d];
        F[d] = a.frc[b + d];
    }
    bool bad = false;
    if (a.first) {
        // firstStep (:191-223)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            Cr[d] = 0.0;
            M[d] = draw(3 + d);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
            xn[d] = x[d] + (a.vel[b + d] * c.coeff1 * c.dt)
                  + (F[d] * c.coeff3 * c.dt * c.dt * m) + (c.sig1 * M[d]);
#pragma unroll
        for (int d = 0; d < 3; ++d) bad |= set_coord(a.pos, a.pos0, b + d, xn[d], a.L);
    } else {
        // normalStep (:225-296): three A draws, their correlation with the
        // previous memX, then three B draws
#pragma unroll
        for (int d = 0; d < 3; ++d) Cr[d] = draw(d);
#pragma unroll
        for (int d = 0; d < 3; ++d)
            Cr[d] = c.sig2 * ((c.corr * a.memx[d * n + i]) + (c.dev * Cr[d]));
#pragma unroll
        for (int d = 0; d < 3; ++d) M[d] = draw(3 + d);
        const double dt2 = c.dt * c.dt;
        const double c1 = m * dt2 * c.coeff1;
        const double c2 = m * dt2 * c.coeff2;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            F0[d] = a.frc0[b + d];
            double q = (c.c0 * x[d]);
            q -= (c.coeff0 * x0[d]);
            q += (c1 * F[d]);
            q += (c2 * (F[d] - F0[d]));
            q += (c.sig1 * M[d]) + (c.coeff0 * Cr[d]);
            xn[d] = q;
        }
        if (a.do_vel) {
            // velocityStep (:298-356), then setPos
            const double dt3 = c.dt * dt2;
            const double g2 = m * dt2 * c.goy2;
            const double g3 = m * dt3 * c.goy3;
            const double g4 = c.hn * c.dt_inv;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double dx0 = x[d] - x0[d];
                const double xN = safe_mod(xn[d], a.L);
                const double xo = safe_mod0(x[d], xN, a.L);
                const double dx = xN - xo;
                double v = dx + dx0;
                v += (g2 * F[d]);
                v -= (g3 * (F[d] - F0[d]));
                v += (Cr[d] - c.sig1 * M[d]);
                v *= g4;
                a.vel[b + d] = v;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) bad |= set_coord(a.pos, a.pos0, b + d, xn[d], a.L);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        a.memx[d * n + i] = M[d];
        a.memcorr[d * n + i] = Cr[d];
    }
    a.fr.jrand[i] = f.jrand;
    a.fr.flags[i] = f.flags;
    a.fr.spare[i] = f.spare;
    if (f.stuck) report_error(a.err, PSIM_E_UNSUPPORTED, i, -1);
    if (bad) report_error(a.err, PSIM_ERR_BOUNDS, i, -1);
}

__global__ void k_int_zero(int *p, size_t count)
{
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0;
}

}  // namespace

struct psim_integrator {
    int n = 0;
    int noise = PSIM_NOISE_REFERENCE;
    int seed = 0;
    unsigned long long philox_seed = 0, calls = 0;
    int vel_freq = 0, vel_counter = 0;
    IntCoef coef{};
    hipStream_t stream = nullptr;
    double *d_pos = nullptr, *d_pos0 = nullptr, *d_vel = nullptr;
    double *d_frc = nullptr, *d_frc0 = nullptr, *d_mass = nullptr;
    double *memx = nullptr, *memcorr = nullptr;
    int *tab = nullptr, *jrand = nullptr, *flags = nullptr;
    float *spare = nullptr;
    DevErr *err = nullptr;
};

extern "C" {

int psim_int_create(const psim_params *P, int noise, psim_integrator **out)
{
    if (!P || !out) return PSIM_E_ARG;
    *out = nullptr;
    if (P->n_particles <= 0 || !(P->dt > 0) || P->gamma < 0) return PSIM_E_ARG;
    if (noise != PSIM_NOISE_REFERENCE && noise != PSIM_NOISE_PHILOX) return PSIM_E_ARG;
    psim_integrator *g = new (std::nothrow) psim_integrator();
    if (!g) return PSIM_E_NOMEM;
    g->n = P->n_particles;
    g->noise = noise;
    g->seed = (int)P->seed;   // the reference's int seed (brownianIntegrator.cpp:83)
    g->philox_seed = P->seed;
    g->vel_freq = P->vel_freq;
    const GBCoeffs k = gb_coefficients(P->kT, P->gamma, P->dt);
    g->coef = IntCoef{k.coeff0, k.coeff1, k.coeff2, k.coeff3, 1.0 + k.coeff0, k.sig1, k.sig2,
                      k.corr, k.dev, k.goy2, k.goy3, k.hn, P->dt, 1.0 / P->dt};
    const size_t n = (size_t)g->n, n3 = 3 * n;
    int rc = 0;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) rc = PSIM_E_HIP;
    if (!rc) rc = eng::dalloc(&g->d_pos, n3);
    if (!rc) rc = eng::dalloc(&g->d_pos0, n3);
    if (!rc) rc = eng::dalloc(&g->d_vel, n3);
    if (!rc) rc = eng::dalloc(&g->d_frc, n3);
    if (!rc) rc = eng::dalloc(&g->d_frc0, n3);
    if (!rc) rc = eng::dalloc(&g->d_mass, n);
    if (!rc) rc = eng::dalloc(&g->memx, n3);
    if (!rc) rc = eng::dalloc(&g->memcorr, n3);
    if (!rc) rc = eng::dalloc(&g->tab, (kFrDim + 1) * n);
    if (!rc) rc = eng::dalloc(&g->jrand, n);
    if (!rc) rc = eng::dalloc(&g->flags, n);
    if (!rc) rc = eng::dalloc(&g->spare, n);
    if (!rc) rc = eng::dalloc(&g->err, 1);
    if (!rc) {
        // unseeded generators, no spare deviates, zero memory (the
        // constructor's fresh fastRandom objects and arrays)
        k_int_zero<<<1024, 256, 0, g->stream>>>(g->jrand, n);
        k_int_zero<<<1024, 256, 0, g->stream>>>(g->flags, n);
        k_int_zero<<<1024, 256, 0, g->stream>>>(reinterpret_cast<int *>(g->memx), 2 * n3);
        k_int_zero<<<1024, 256, 0, g->stream>>>(reinterpret_cast<int *>(g->memcorr), 2 * n3);
        k_int_zero<<<1, 64, 0, g->stream>>>(reinterpret_cast<int *>(g->err), 4);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess)
            rc = PSIM_E_HIP;
    }
    if (rc) {
        psim_int_destroy(g);
        return rc;
    }
    *out = g;
    return 0;
}

void psim_int_destroy(psim_integrator *g)
{
    if (!g) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    void *bufs[] = {g->d_pos, g->d_pos0, g->d_vel, g->d_frc, g->d_frc0, g->d_mass, g->memx,
                    g->memcorr, g->tab, g->jrand, g->flags, g->spare, g->err};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

int psim_int_next(psim_integrator *g, double current_time, int box, double *pos, double *pos0,
                  double *vel, const double *frc, const double *frc0, const double *mass,
                  int *err_index)
{
    if (!g || !pos || !pos0 || !vel || !frc || !frc0 || !mass || box <= 0) return PSIM_E_ARG;
    if (err_index) *err_index = -1;
    const size_t n = (size_t)g->n, n3 = 3 * n;
    IntArgs a{};
    a.n = g->n;
    a.first = (current_time == 0) ? 1 : 0;   // brownianIntegrator.cpp:183
    a.do_vel = 0;
    if (!a.first) {
        // :288 and the counter update of :292
        a.do_vel = (g->vel_freq == 0 || g->vel_counter == g->vel_freq) ? 1 : 0;
        if (g->vel_counter == g->vel_freq) g->vel_counter = 0; else g->vel_counter++;
    }
    a.noise = g->noise;
    a.seed = g->seed;
    a.philox_seed = g->philox_seed;
    a.call = g->calls++;
    a.L = (double)box;
    a.pos = g->d_pos;
    a.pos0 = g->d_pos0;
    a.vel = g->d_vel;
    a.frc = g->d_frc;
    a.frc0 = g->d_frc0;
    a.mass = g->d_mass;
    a.memx = g->memx;
    a.memcorr = g->memcorr;
    a.fr = FrState{g->tab, g->jrand, g->flags, g->spare, g->n};
    a.err = g->err;
    hipStream_t st = g->stream;
    CK(hipMemcpyAsync(g->d_pos, pos, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(g->d_pos0, pos0, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(g->d_vel, vel, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(g->d_frc, frc, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(g->d_frc0, frc0, sizeof(double) * n3, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(g->d_mass, mass, sizeof(double) * n, hipMemcpyHostToDevice, st));
    k_int_step<<<eng::blocks_for((long long)n, 256), 256, 0, st>>>(g->coef, a);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(pos, g->d_pos, sizeof(double) * n3, hipMemcpyDeviceToHost, st));
    CK(hipMemcpyAsync(pos0, g->d_pos0, sizeof(double) * n3, hipMemcpyDeviceToHost, st));
    if (a.do_vel) CK(hipMemcpyAsync(vel, g->d_vel, sizeof(double) * n3, hipMemcpyDeviceToHost, st));
    DevErr h{};
    CK(hipMemcpyAsync(&h, g->err, sizeof(DevErr), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (h.code && err_index) *err_index = h.i;
    return h.code;
}

}  // extern "C"
