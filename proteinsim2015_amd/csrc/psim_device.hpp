// psim_device.hpp -- device-side building blocks of the gfx950 hot path.
//
// Arithmetic contract (the file is compiled with -ffp-contract=off):
//   * every expression that must round like the reference (integrator,
//     wrap, cell hash, exact pair distance) is written in the reference's
//     operation order with no FMA;
//   * the pair loop uses explicit fma() where rounding does not matter.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psim {

constexpr int kBlock = 256;        // threads per block of the pair kernels
constexpr int kHitCap = 32;        // per-lane capacity of the LDS hit list
constexpr unsigned kIdxBits = 26;  // slot index bits in a hit-list entry
constexpr unsigned kIdxMask = (1u << kIdxBits) - 1u;

enum { F_CAL = 1, F_LJ = 2, F_AO = 4 };

// Box and cell geometry (system.cpp:219-227; systemState.h:7-19).
struct Geo {
    int n;          // particles in this system
    int S;          // cells per dimension (cellScale)
    int S2;
    int ncell;
    double L;       // box edge
    double half_pbc;  // L / 2.0 as used by pbcDist (utilities.cpp:85)
    double half_uv;   // L / 2 as used by unitVectorAdv (utilities.cpp:124)
    double cell;      // cell edge used by the hash
    int exact_mi;     // 1: per-pair reference minimum image (S < 4)
    int pad;
};

// Pair-force constants (LJPotential.cpp:34-55, AOPotential.cpp:40-55).
struct Pot {
    int mask;
    int lj_num;
    double rc2;       // cutOff*cutOff, strict test r^2 < rc2
    double rc2_hi;    // cull acceptance (guard band above rc2)
    double rc2_lo;    // below this the fast r^2 decides; above, exact
    double ov_factor; // 0.8 (overlap threshold factor, LJPotential.cpp:87)
    double ov_band;   // relative band for the exact overlap re-test
    double lj_neg_kt_eps;   // (-kT)*wellDepth
    double lj_four_n;       // 4.0*ljNum
    double debye_len;
    double debye_inv;
    double yuk_a;
    double ao_c1, ao_c2;
};

// G&B integrator coefficients (brownianIntegrator.cpp:53-81,111-175),
// precomputed on the host with the reference's formulas.
struct GB {
    double coeff0, coeff1, coeff3;
    double c0;        // 1 + coEff0
    double c1, c2;    // m*dt^2*coEff1, m*dt^2*coEff2
    double sig1, sig2, corr, dev;
    double m;         // 1/mass
    double dt;
    double g2, g3, g4;  // velocityStep (:328-330)
};

struct DevErr {
    int code;
    int i;
    int j;
    int pad;
};

__device__ inline void report_error(DevErr *e, int code, int i, int j)
{
    if (atomicCAS(&e->code, 0, code) == 0) {
        e->i = i;
        e->j = j;
    }
}

// floor(x / cellSize) clamped to [0, S-1] (systemHandling.cpp:39-41; the
// reference leaves the x/cellSize == S case undefined).
__device__ inline int cell_1d(double x, double cell, int S)
{
    double q = floor(x / cell);
    // NaN and negative -> 0, >= S -> S-1: the cell index always stays in range
    if (!(q >= 0.0)) return 0;
    return q < (double)S ? (int)q : S - 1;
}

// util::pbcDist with the reference's operation order (utilities.cpp:79-104).
__device__ inline double pbc_dist_ref(double X, double Y, double Z, double X1,
                                      double Y1, double Z1, double L,
                                      double half)
{
    double dx = fabs(X - X1);
    double dy = fabs(Y - Y1);
    double dz = fabs(Z - Z1);
    if (dx > half) dx -= L;
    if (dy > half) dy -= L;
    if (dz > half) dz -= L;
    return (dx * dx) + (dy * dy) + (dz * dz);
}

// util::unitVectorAdv's signed minimum-image vector (utilities.cpp:118-136).
__device__ inline double min_image_uv(double d, double L, double half_uv)
{
    if (fabs(d) > half_uv) d = (d < 0) ? d + L : d - L;
    return d;
}

template <int E>
__device__ inline double pow_const(double b)
{
    double a = 1.0;
    int e = E;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        if (e & 1) a *= b;
        e >>= 1;
        b *= b;
    }
    return a;
}

__device__ inline double pow_int(double b, int e)
{
    double a = 1.0;
    while (e) {
        if (e & 1) a *= b;
        e >>= 1;
        b *= b;
    }
    return a;
}

// Pair force magnitude (positive = attractive) summed over the enabled
// modules.  r2, rinv = 1/sqrt(r2), r = r2*rinv; size = r_i + r_j.
// Sets *nan_ao when the AO term is NaN (AOPotential.cpp:105-109).
template <int MASK>
__device__ inline double pair_fnet(const Pot &p, double r2, double rinv,
                                   double r, double size, bool *nan_ao)
{
    double f = 0.0;
    if (MASK & F_CAL) {
        // Calibration.cpp:75-81: -36 * rInv^37
        f += -(36.0 * pow_const<37>(rinv));
    }
    if (MASK & F_LJ) {
        // LJPotential.cpp:97-109
        double lj = pow_int(size * rinv, p.lj_num);
        double attract = fma(2.0, lj, -1.0) * (p.lj_four_n * rinv * lj);
        double repel = 0.0;
        if (p.yuk_a != 0.0) {
            double ye = exp(-(r * p.debye_inv));
            repel = ye * (rinv * rinv * (p.debye_len + r) * p.yuk_a);
        }
        f += p.lj_neg_kt_eps * (attract + repel);
    }
    if (MASK & F_AO) {
        // AOPotential.cpp:95-102: -(36 r^-38 + C1/r + C2 r)
        double r36 = pow_const<36>(rinv);
        double fa = -(fma(36.0 * r36, rinv * rinv, fma(p.ao_c1, rinv, p.ao_c2 * r)));
        if (isnan(fa)) *nan_ao = true;
        f += fa;
    }
    return f;
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller normals.
// counter = (id, step_lo, step_hi, block), key = (seed_lo, seed_hi).
// ---------------------------------------------------------------------------
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double u01(uint32_t v)
{
    return ((double)v + 0.5) * 2.3283064365386963e-10;
}

__device__ inline void philox_normals6(unsigned long long seed,
                                       unsigned long long step, uint32_t id,
                                       double z[6])
{
    uint32_t a[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    uint32_t b[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), 1u};
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    philox4x32_10(a, k0, k1);
    philox4x32_10(b, k0, k1);
    uint32_t u[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    const double two_pi = 6.283185307179586;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double rr = sqrt(-2.0 * log(u01(u[2 * q])));
        double th = two_pi * u01(u[2 * q + 1]);
        z[2 * q] = rr * cos(th);
        z[2 * q + 1] = rr * sin(th);
    }
}

// util::safeMod (utilities.cpp:26-39)
__device__ inline double safe_mod(double v, double L)
{
    if (v < 0) return v + L;
    if (v >= L) return v - L;
    return v;
}

// util::safeMod0 (utilities.cpp:41-50)
__device__ inline double safe_mod0(double v0, double v, double L)
{
    double dx = v - v0;
    return (fabs(dx) > (L / 2)) ? ((dx < 0) ? v0 - L : v0 + L) : v0;
}

}  // namespace psim
