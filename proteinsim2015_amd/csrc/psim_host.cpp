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
            if (!g.free_at(p)) continue;           // accepted earlier this round
            if (!g.insert(p)) return PSIM_E_UNSUPPORTED;
            for (int d = 0; d < 3; ++d) out[3 * (size_t)placed + d] = p[d];
            ++placed;
        }
        next_cand += m;
    }
    return 0;
}
