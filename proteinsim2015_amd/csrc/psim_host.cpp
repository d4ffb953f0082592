// psim_host.cpp -- host-side system initialisation.
//
// psim_init_rsa replaces system::initParticles + initCheck
// (systemInit.cpp:31-160), which places particles uniformly at random and
// then re-draws any particle closer than r_i + r_j to another one by an O(N^2)
// scan.  Here: random sequential addition, O(N) expected, with a cell grid:
//   * candidate k is a pure function of (seed, k) (SplitMix64 of the counter),
//   * candidates are processed in rounds: every candidate of a round is
//     tested in parallel against the points accepted in earlier rounds, then
//     the survivors are accepted in candidate order with a second test
//     against the points accepted earlier in the same round.
// The result is exactly sequential addition over the candidate stream, so it
// does not depend on the thread count.
#include <cmath>
#include <cstdint>
#include <vector>

#include "psim_gpu.h"

namespace {

inline uint64_t splitmix(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline double unit(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

inline double min_image(double d, double L)
{
    if (d > 0.5 * L) return d - L;
    if (d < -0.5 * L) return d + L;
    return d;
}

struct Grid {
    int G = 1;
    int cap = 0;
    double edge = 0, box = 0, d2 = 0;
    std::vector<int> count;
    std::vector<double> xyz;   // [cell][cap][3]

    int cell_of(double x) const
    {
        int c = (int)(x / edge);
        return c < 0 ? 0 : (c >= G ? G - 1 : c);
    }
    size_t cell_index(const double p[3]) const
    {
        return ((size_t)cell_of(p[2]) * G + cell_of(p[1])) * G + cell_of(p[0]);
    }
    bool free_at(const double p[3]) const
    {
        int cx = cell_of(p[0]), cy = cell_of(p[1]), cz = cell_of(p[2]);
        for (int oz = -1; oz <= 1; ++oz) {
            size_t rz = (size_t)((cz + oz + G) % G);
            for (int oy = -1; oy <= 1; ++oy) {
                size_t ry = (size_t)((cy + oy + G) % G);
                for (int ox = -1; ox <= 1; ++ox) {
                    size_t c = (rz * G + ry) * G + (size_t)((cx + ox + G) % G);
                    const double *q = &xyz[c * cap * 3];
                    for (int k = 0; k < count[c]; ++k) {
                        double dx = min_image(p[0] - q[3 * k], box);
                        double dy = min_image(p[1] - q[3 * k + 1], box);
                        double dz = min_image(p[2] - q[3 * k + 2], box);
                        if (dx * dx + dy * dy + dz * dz < d2) return false;
                    }
                }
            }
        }
        return true;
    }
    bool insert(const double p[3])
    {
        size_t c = cell_index(p);
        if (count[c] >= cap) return false;
        double *q = &xyz[(c * cap + count[c]) * 3];
        q[0] = p[0];
        q[1] = p[1];
        q[2] = p[2];
        ++count[c];
        return true;
    }
};

void candidate(uint64_t seed, uint64_t k, double box, double p[3])
{
    uint64_t base = splitmix(seed ^ splitmix(k));
    for (int d = 0; d < 3; ++d) {
        double v = unit(splitmix(base + (uint64_t)d)) * box;
        p[d] = v < box ? v : std::nextafter(box, 0.0);
    }
}

int rsa_brute(int n, double box, double dmin, uint64_t seed, double *out)
{
    const double d2 = dmin * dmin;
    const long long max_cand = 2000ll * n + 1000000ll;
    int placed = 0;
    for (long long k = 0; placed < n; ++k) {
        if (k > max_cand) return PSIM_ERR_INIT;
        double p[3];
        candidate(seed, (uint64_t)k, box, p);
        bool ok = true;
        for (int j = 0; j < placed && ok; ++j) {
            double dx = min_image(p[0] - out[3 * (size_t)j], box);
            double dy = min_image(p[1] - out[3 * (size_t)j + 1], box);
            double dz = min_image(p[2] - out[3 * (size_t)j + 2], box);
            ok = dx * dx + dy * dy + dz * dz >= d2;
        }
        if (!ok) continue;
        for (int d = 0; d < 3; ++d) out[3 * (size_t)placed + d] = p[d];
        ++placed;
    }
    return 0;
}

}  // namespace

extern "C" int psim_init_rsa(int n, double box, double dmin, unsigned long long seed,
                             double *out)
{
    if (n <= 0 || !(box > 0) || !(dmin > 0) || !out) return PSIM_E_ARG;
    int G = (int)std::floor(box / dmin);
    if (G < 3) return rsa_brute(n, box, dmin, seed, out);
    Grid g;
    g.G = G;
    g.edge = box / G;
    g.box = box;
    g.d2 = dmin * dmin;
    // points per cell of edge e with pair distance >= d is bounded by the
    // packing density: ((e + d)/d)^3 * 0.7405 / (pi/6); e < 1.1 d for G >= 10
    g.cap = (G >= 10) ? 16 : 32;
    const size_t ncell = (size_t)G * G * G;
    g.count.assign(ncell, 0);
    g.xyz.assign(ncell * g.cap * 3, 0.0);
    const long long max_cand = 4000ll * n + 1000000ll;
    long long next_cand = 0;
    int placed = 0;
    std::vector<double> cand;
    std::vector<unsigned char> ok;
    while (placed < n) {
        if (next_cand > max_cand) return PSIM_ERR_INIT;   // error.cpp:27-32
        long long remaining = n - placed;
        long long m = remaining < 16384 ? 16384 : (remaining > (1 << 20) ? (1 << 20) : remaining);
        cand.resize((size_t)m * 3);
        ok.assign((size_t)m, 0);
#pragma omp parallel for schedule(static)
        for (long long k = 0; k < m; ++k) {
            double *p = &cand[(size_t)k * 3];
            candidate(seed, (uint64_t)(next_cand + k), box, p);
            ok[(size_t)k] = g.free_at(p) ? 1 : 0;
        }
        for (long long k = 0; k < m && placed < n; ++k) {
            if (!ok[(size_t)k]) continue;
            const double *p = &cand[(size_t)k * 3];
            if (!g.free_at(p)) continue;           // a candidate of this batch took the spot
            if (!g.insert(p)) return PSIM_E_UNSUPPORTED;
            for (int d = 0; d < 3; ++d) out[3 * (size_t)placed + d] = p[d];
            ++placed;
        }
        next_cand += m;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// The reference's own initial configuration
// ---------------------------------------------------------------------------
// psim_init_reference replays system::initParticles, initCheck and
// maxwellVelocityInit (systemInit.cpp:31-267): the same generator
// (std::mt19937 seeded with the config seed), the same
// std::uniform_real_distribution<double>(0, 1) draws in the same order, the
// particle setters' single wrap and previous-position image
// (particle.cpp:61-99, utilities.cpp:26-50), and the same libm calls, so with
// the same C++ library it reproduces the reference's positions and velocities
// bit for bit.  initCheck's O(N^2) scan (each particle against every other
// one in index order, re-drawn on the first overlap found and the scan
// continued with the new position) is replayed with a cell grid that yields
// the overlapping particles in index order, which keeps the draw sequence.
#include <algorithm>
#include <random>

namespace {

struct RefParticles {
    int n = 0;
    double box = 0;
    std::vector<double> x, x0;   // AoS3

    // particle::setX/Y/Z (particle.cpp:61-92)
    void set(int i, int d, double val)
    {
        const double old = x[3 * (size_t)i + d];
        double v = val;
        if (v < 0) v = v + box;            // util::safeMod
        else if (v >= box) v = v - box;
        x[3 * (size_t)i + d] = v;
        const double dx = v - old;          // util::safeMod0
        x0[3 * (size_t)i + d] = (std::fabs(dx) > (box / 2)) ? ((dx < 0) ? old - box : old + box)
                                                             : old;
    }
};

// util::pbcDist (utilities.cpp:79-104)
inline double ref_pbc_dist(const double *a, const double *b, double L)
{
    double dx = std::fabs(a[0] - b[0]), dy = std::fabs(a[1] - b[1]), dz = std::fabs(a[2] - b[2]);
    const double half = L / 2.0;
    if (dx > half) dx -= L;
    if (dy > half) dy -= L;
    if (dz > half) dz -= L;
    return (dx * dx) + (dy * dy) + (dz * dz);
}

struct IndexGrid {
    int G = 0;
    double edge = 0;
    std::vector<std::vector<int>> cells;
    std::vector<int> where;

    int c1(double v) const
    {
        int c = (int)(v / edge);
        return c < 0 ? 0 : (c >= G ? G - 1 : c);
    }
    int cell(const double *p) const { return (c1(p[2]) * G + c1(p[1])) * G + c1(p[0]); }
    void insert(int i, const double *p)
    {
        const int c = cell(p);
        cells[c].push_back(i);
        where[i] = c;
    }
    void remove(int i)
    {
        auto &v = cells[where[i]];
        v.erase(std::find(v.begin(), v.end(), i));
    }
};

}  // namespace

extern "C" int psim_init_reference(int n, int box, double radius, unsigned int seed,
                                   int seed_size, double temp, double *pos_out,
                                   double *pos0_out, double *vel_out)
{
    if (n <= 0 || box <= 0 || !(radius > 0) || seed_size < 0 || !pos_out) return PSIM_E_ARG;
    const int seed_cube = seed_size * seed_size * seed_size;
    if (seed_cube > n) return PSIM_E_ARG;
    RefParticles P;
    P.n = n;
    P.box = box;
    P.x.assign(3 * (size_t)n, 0.0);
    P.x0.assign(3 * (size_t)n, 0.0);
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> distribution(0.0, 1.0);
    // seed cube (systemInit.cpp:45-76)
    if (seed_size > 0) {
        int lower, upper;
        if (seed_size % 2 == 0) {
            lower = -seed_size / 2;
            upper = seed_size / 2;
        } else {
            lower = -seed_size / 2;
            upper = (seed_size / 2) + 1;
        }
        int k = 0;
        for (int x = lower; x < upper; x++)
            for (int y = lower; y < upper; y++)
                for (int z = lower; z < upper; z++) {
                    float center = box / 2.0;
                    P.set(k, 0, center + (2.1 * x * radius));
                    P.set(k, 1, center + (2.1 * y * radius));
                    P.set(k, 2, center + (2.1 * z * radius));
                    k++;
                }
    }
    for (int i = seed_cube; i < n; i++) {   // :80-89
        P.set(i, 0, distribution(gen) * box);
        P.set(i, 1, distribution(gen) * box);
        P.set(i, 2, distribution(gen) * box);
    }
    // initCheck (:99-160): overlap when pbcDist (a squared distance) is below
    // the sum of the radii, i.e. distance < sqrt(2 r)
    const double rsum = radius + radius;
    const double reach = std::sqrt(rsum);
    IndexGrid grid;
    grid.G = (int)std::floor(box / reach);
    const bool use_grid = grid.G >= 3;
    if (use_grid) {
        grid.edge = (double)box / grid.G;
        grid.cells.assign((size_t)grid.G * grid.G * grid.G, {});
        grid.where.assign(n, 0);
        for (int i = 0; i < n; i++) grid.insert(i, &P.x[3 * (size_t)i]);
    }
    // smallest j >= from (j != i) overlapping particle i at its current position
    auto next_overlap = [&](int i, int from) -> int {
        const double *pi = &P.x[3 * (size_t)i];
        int best = n;
        if (!use_grid) {
            for (int j = from; j < n; j++)
                if (j != i && ref_pbc_dist(pi, &P.x[3 * (size_t)j], box) < 1.0 * rsum) return j;
            return n;
        }
        const int G = grid.G;
        const int cx = grid.c1(pi[0]), cy = grid.c1(pi[1]), cz = grid.c1(pi[2]);
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    const int c = (((cz + oz + G) % G) * G + (cy + oy + G) % G) * G + (cx + ox + G) % G;
                    for (int j : grid.cells[c])
                        if (j >= from && j < best && j != i &&
                            ref_pbc_dist(pi, &P.x[3 * (size_t)j], box) < 1.0 * rsum)
                            best = j;
                }
        return best;
    };
    long long counter = 0;
    for (int i = seed_cube; i < n; i++) {
        bool resolution = false;
        while (!resolution) {
            resolution = true;
            int j = next_overlap(i, 0);
            while (j < n) {
                counter++;
                if (counter > 10LL * n) return PSIM_ERR_INIT;   // throwInitializationError
                resolution = false;
                if (use_grid) grid.remove(i);
                P.set(i, 0, distribution(gen) * box);
                P.set(i, 1, distribution(gen) * box);
                P.set(i, 2, distribution(gen) * box);
                if (use_grid) grid.insert(i, &P.x[3 * (size_t)i]);
                j = next_overlap(i, j + 1);
            }
        }
    }
    // maxwellVelocityInit (:161-267); the temperature is truncated to int
    std::vector<double> v(3 * (size_t)n, 0.0);
    const int itemp = (int)temp;
    if (n == 1) {
        for (int d = 0; d < 3; d++) {
            const int dir = (distribution(gen) < 0.5) ? -1 : 1;
            v[d] = dir * std::sqrt(itemp);
        }
    } else {
        for (int d = 0; d < 3; d++)
            for (int i = 0; i < n; i++) {
                const double r1 = distribution(gen);
                const double r2 = distribution(gen);
                v[3 * (size_t)i + d] = std::sqrt(-2.0 * std::log(r1)) * std::cos(8.0 * std::atan(1) * r2);
            }
        for (int d = 0; d < 3; d++) {
            double vsum = 0, vsum2 = 0;
            for (int i = 0; i < n; i++) {
                const double vx = v[3 * (size_t)i + d];
                vsum = vsum + vx;
                vsum2 = vsum2 + (vx * vx);
            }
            vsum = vsum / n;
            vsum2 = vsum2 / n;
            const double sigold = std::sqrt(vsum2 - (vsum * vsum));
            const double vsig = std::sqrt(itemp);
            const double ratio = vsig / sigold;
            for (int i = 0; i < n; i++) v[3 * (size_t)i + d] = ratio * (v[3 * (size_t)i + d] - vsum);
        }
    }
    std::copy(P.x.begin(), P.x.end(), pos_out);
    if (pos0_out) std::copy(P.x0.begin(), P.x0.end(), pos0_out);
    if (vel_out) std::copy(v.begin(), v.end(), vel_out);
    return 0;
}
