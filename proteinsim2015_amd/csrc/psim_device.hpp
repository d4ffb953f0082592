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

// Box and cell geometry (system.cpp:50-64; systemState.h:7-19).
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
    int c0;           // offset of this system's cells in the cell-index space
                      // (0 for one system; replicas of an ensemble are
                      // stacked, each owning S^3 consecutive cells)
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
    double lj_att;          // (-kT)*wellDepth*4*ljNum (the kernels' folded factors)
    double lj_rep;          // (-kT)*wellDepth*yukawaStrength
    double lj_rep_lam;      // lj_rep * debyeLength
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

// Step control kept in device memory so that a captured sequence of steps
// can be replayed: the step kernels read it, the last kernel of each step
// (the cell rebuild's move) advances it.
// (brownianIntegrator.cpp:183 first-step test, :288-292 velocity counter,
// system.cpp:210 time advance)
struct StepCtl {
    unsigned long long step;   // Philox counter of the current step
    double time;               // currentTime at the start of the step
    double dt;
    int vel_counter;
    int vel_freq;
    int first;                 // currentTime == 0
    int do_vel;                // velocityStep this step
};

// Flags of the step about to run from its clock and velocity counter
// (brownianIntegrator.cpp:183 first step, :288-292 velocity counter).
__device__ __forceinline__ void ctl_prepare(StepCtl *c)
{
    c->first = (c->time == 0) ? 1 : 0;
    c->do_vel = 0;
    if (!c->first) {
        c->do_vel = (c->vel_freq == 0 || c->vel_counter == c->vel_freq) ? 1 : 0;
        if (c->vel_counter == c->vel_freq) c->vel_counter = 0; else c->vel_counter++;
    }
}

// End of step: advance the clock and set the flags of the next step
// (system.cpp:210).
__device__ __forceinline__ void ctl_advance(StepCtl *c)
{
    c->step += 1;
    c->time += c->dt;
    ctl_prepare(c);
}

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

// ---------------------------------------------------------------------------
// Neighbour-set statistics.  A pair kernel instantiated with NbrStats instead
// of NoStats records, for every particle, the pairs its acceptance test
// admits (the strict r^2 < cutOff^2 of LJPotential.cpp:82 after the
// guard-banded cull and util::pbcDist's recheck) and skips the force
// arithmetic: the count, and the sum mod 2^64 of a mixing hash of each
// partner's original index (the finaliser of the SplitMix64 generator of
// Steele, Lea and Flood, OOPSLA 2014), which differs between two partner
// sets of equal size unless the hashes collide.
// ---------------------------------------------------------------------------
__host__ __device__ inline unsigned long long nbr_hash(unsigned o)
{
    unsigned long long z = (unsigned long long)o + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct NoStats {
    static constexpr bool kOn = false;
    __device__ void add(int) {}
};

struct NbrStats {
    static constexpr bool kOn = true;
    int count = 0;
    unsigned long long hash = 0;
    __device__ void add(int o)
    {
        ++count;
        hash += nbr_hash((unsigned)o);
    }
};

// Where a statistics kernel writes, by original particle index.
struct NbrStatsOut {
    int *count;
    unsigned long long *hash;
};

// Sum of the statistics of the SPLIT consecutive lanes that share a particle.
template <int SPLIT>
__device__ __forceinline__ void split_stats(NbrStats &st)
{
#pragma unroll
    for (int d = 1; d < SPLIT; d <<= 1) {
        st.count += __shfl_xor(st.count, d, 64);
        st.hash += (unsigned long long)__shfl_xor((long long)st.hash, d, 64);
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
__device__ __forceinline__ double pow_const(double b)
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

__device__ __forceinline__ double pow_int(double b, int e)
{
    double a = 1.0;
    while (e) {
        if (e & 1) a *= b;
        e >>= 1;
        b *= b;
    }
    return a;
}

// 1/sqrt(x) for normal x: the hardware estimate (v_rsq_f64, measured on
// gfx950 within 2.5e8 ulp, i.e. ~6e-8 relative) refined by one third-order
// step y <- y + y e (1/2 + 3/8 e), e = 1 - x y^2 (the series of
// (1 - e)^(-1/2) to e^2), which leaves a relative error ~(5/16) e^3 below
// the rounding of the last operations: within about 1 ulp, at five
// instructions against the seven of two Newton steps.
__device__ __forceinline__ double rsqrt_nr(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

// e^{-a} for 0 <= a < 700: a = k ln2 + f with |f| <= ln2/2 (two-part ln2),
// e^{-f} by its Taylor series to degree 13 (truncation < 3e-17 relative),
// then scaled by 2^{-k}.
__device__ __forceinline__ double exp_neg(double a)
{
    const double log2e = 1.4426950408889634;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double k = rint(a * log2e);
    double f = fma(-k, ln2_hi, a);
    f = fma(-k, ln2_lo, f);
    const double x = -f;
    double p = 1.0 / 6227020800.0;            // 1/13!
    p = fma(p, x, 1.0 / 479001600.0);         // 1/12!
    p = fma(p, x, 1.0 / 39916800.0);
    p = fma(p, x, 1.0 / 3628800.0);
    p = fma(p, x, 1.0 / 362880.0);
    p = fma(p, x, 1.0 / 40320.0);
    p = fma(p, x, 1.0 / 5040.0);
    p = fma(p, x, 1.0 / 720.0);
    p = fma(p, x, 1.0 / 120.0);
    p = fma(p, x, 1.0 / 24.0);
    p = fma(p, x, 1.0 / 6.0);
    p = fma(p, x, 0.5);
    p = fma(p, x, 1.0);
    p = fma(p, x, 1.0);
    return ldexp(p, -(int)k);
}

// 2^(j/64), j = 0..63, each the double nearest the exact value.
static __device__ const double kExp2Tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// e^{-a} for 0 <= a < 700 with a 64-entry table (tab = kExp2Tab or a copy
// of it): -a = k ln2/64 - f with k = rint(-a 64/ln2) = 64 e + j (e = k >> 6,
// j = k & 63, so 0 <= j < 64) and |f| <= ln2/128 (ln2/64 in two parts, from
// the hi/lo split of ln 2 in fdlibm's e_exp.c); e^{-f} by its Taylor series
// to degree 5 (truncation < 4e-17 relative, a third of an ulp), times 2^{j/64} from the table,
// scaled by 2^e.  (2^{j/64} = 2 * 2^{-(64-j)/64} exactly in doubles, so this
// rounds exactly as the same scheme with a table of 2^{-j/64} would.)
__device__ __forceinline__ double exp_neg_tab(double a, const double *tab)
{
    const double inv = -92.332482616893656;           // -64 / ln 2
    const double ln2_hi = 6.93147180369123816490e-01 / 64.0;
    const double ln2_lo = 1.90821492927058770002e-10 / 64.0;
    const double k = rint(a * inv);   // <= 0
    double f = fma(k, ln2_hi, a);
    f = fma(k, ln2_lo, f);
    const double x = -f;
    double q = 1.0 / 120.0;
    q = fma(q, x, 1.0 / 24.0);
    q = fma(q, x, 1.0 / 6.0);
    q = fma(q, x, 0.5);
    q = fma(q, x, 1.0);
    q = fma(q, x, 1.0);
    const int ki = (int)k;
    return ldexp(q * tab[ki & 63], ki >> 6);
}

// (size/r)^n for the LJ module, n = ljNum: from x = (size/r)^2 for the common
// even exponents (powBinaryDecomp of size * rinv otherwise, utilities.cpp:
// 144-154); LJN > 0 fixes n at compile time (no branch on n).
template <int LJN = 0>
__device__ __forceinline__ double lj_power_x(double x, int n, double size, double rinv)
{
    switch (LJN > 0 ? LJN : n) {
    case 6: return x * x * x;
    case 12: { double x2 = x * x; double x4 = x2 * x2; return x4 * x2; }
    case 18: { double x2 = x * x; double x4 = x2 * x2; double x8 = x4 * x4; return x8 * x; }
    case 24: { double x2 = x * x; double x4 = x2 * x2; double x8 = x4 * x4; return x8 * x4; }
    default: return pow_int(size * rinv, n);
    }
}

// fNet / r, the pair force magnitude (positive = attractive) summed over the
// enabled modules and divided by r: the kernels need only the force vector
// (fNet / r) (x_j - x_i).  r^-2 is shared between the terms and the division
// by r folded into each term.  rinv = 1/r; size = r_i + r_j.  TAB: the Yukawa
// exponential through the 2^(j/64) table at etab (kernels that keep a copy
// in LDS).  LJN: ljNum fixed at compile time (0: p.lj_num).  Sets *nan_ao
// when the AO term is NaN (AOPotential.cpp:105-109).
template <int MASK, bool TAB = false, int LJN = 0>
__device__ __forceinline__ double pair_fr(const Pot &p, double rinv, double r, double size,
                                          bool *nan_ao, const double *etab = nullptr)
{
    const double r2inv = rinv * rinv;
    // the first enabled module's term starts the sum (no 0 + x)
    double f = 0.0;
    if (MASK & F_CAL) {
        // Calibration.cpp:75-81: fNet = -36 rInv^37
        f = -(36.0 * pow_const<19>(r2inv));
    }
    if (MASK & F_LJ) {
        // LJPotential.cpp:97-109: fNet = (-kT eps) (attract + repel) with
        // attract = (2 lj - 1) 4 n lj / r, repel = A e^{-r/lambda} (lambda + r)
        // / r^2; the factor (-kT eps) folded into each term's constant, and
        // (lambda + r) / r = lambda / r + 1 taken as one fma
        const double lj = lj_power_x<LJN>((size * size) * r2inv, p.lj_num, size, rinv);
        double t = fma(2.0, lj, -1.0) * (p.lj_att * lj);
        if (p.yuk_a != 0.0) {
            const double ye = TAB ? exp_neg_tab(r * p.debye_inv, etab)
                                  : exp_neg(r * p.debye_inv);
            t = fma(ye, fma(p.lj_rep_lam, rinv, p.lj_rep), t);
        }
        t *= r2inv;
        f = (MASK & F_CAL) ? f + t : t;
    }
    if (MASK & F_AO) {
        // AOPotential.cpp:95-102: fNet = -(36 r^-38 + C1 / r + C2 r)
        const double r38 = pow_const<19>(r2inv);
        const double fa = -(fma(36.0 * r38, rinv, fma(p.ao_c1, r2inv, p.ao_c2)));
        if (isnan(fa)) *nan_ao = true;
        f = (MASK & (F_CAL | F_LJ)) ? f + fa : fa;
    }
    return f;
}

// Per-cell rank through the count array.  Adjacent active lanes with equal
// keys form a run; the run's first lane reserves the run length with one
// atomicAdd and the lanes take consecutive ranks.  (The particles of a cell
// sit in consecutive slots, so per-lane atomics would serialise on a few
// addresses; a key split over several runs simply reserves several ranges.)
__device__ __forceinline__ int wave_aggregated_rank(int *count, int key)
{
    const int lane = (int)__lane_id();
    const unsigned long long active = __ballot(1);
    const int prev_key = __shfl_up(key, 1);
    const bool prev_active = lane > 0 && ((active >> (lane - 1)) & 1ull);
    const bool start = !prev_active || prev_key != key;
    const unsigned long long starts = __ballot(start);
    // this lane's run leader: highest start at or below the lane
    const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int leader = 63 - __clzll((long long)(starts & le));
    // run end: lowest start above the lane, else past the last active lane
    const unsigned long long above = starts & ~le;
    const int end = above ? (__ffsll((long long)above) - 1) : 64;
    int base = 0;
    if (start) {
        const unsigned long long span = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) &
                                        ~((1ull << lane) - 1ull);
        base = atomicAdd(&count[key], __popcll(active & span));
    }
    base = __shfl(base, leader);
    return base + __popcll(active & le & ~((1ull << leader) - 1ull)) - 1;
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

// One Box-Muller pair q (0, 1, 2) of philox_normals6: normals 2q and 2q + 1.
__device__ inline void philox_pair(unsigned long long seed, unsigned long long step, uint32_t id,
                                   int q, double &z0, double &z1)
{
    uint32_t c[4] = {id, (uint32_t)step, (uint32_t)(step >> 32), q < 2 ? 0u : 1u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint32_t ur = (q == 1) ? c[2] : c[0], ua = (q == 1) ? c[3] : c[1];
    const double rr = sqrt(-2.0 * log(u01(ur)));
    const double th = 6.283185307179586 * u01(ua);
    z0 = rr * cos(th);
    z1 = rr * sin(th);
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

// Sum over the SPLIT consecutive lanes that share a particle (a fixed
// butterfly: every lane of the group gets the same bits).
template <int SPLIT>
__device__ __forceinline__ double split_sum(double v)
{
#pragma unroll
    for (int d = 1; d < SPLIT; d <<= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Workgroups are handed out round-robin over the 8 XCDs.  With a grid of
// 8 * per workgroups (xcd_grid), workgroup b takes slab position
// (b mod 8) * per + b / 8: each XCD walks one contiguous range of slots, and
// the neighbour cells those slots read stay in that XCD's L2.
__device__ __forceinline__ int xcd_block()
{
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Grid size for xcd_block: nb workgroups rounded up to a multiple of 8 (the
// extra workgroups find no slot and return).
inline unsigned xcd_grid(unsigned nb) { return 8u * ((nb + 7u) / 8u); }

}  // namespace psim
