// psim_state.hip -- full-precision state file of a context.
//
// No reference counterpart: the reference's recovery file
// (analysisOutput.cpp:55-74) keeps 6 significant digits and drops the
// velocities, the integrator memory, the time and the displacement tracker,
// so its resume (RecoverySystem.cpp:27-43) restarts from the first-step
// branch.  This file keeps every piece of state bit for bit, including the
// slot order of the particles: the cell build keeps the previous order within
// each cell, and the force sums follow that order, so a resumed run that
// stages the particles in the saved order continues exactly as the
// uninterrupted one would.
//
// Layout (native endianness): header, then AoS3 doubles by original index
// for positions, previous positions, velocities, last forces and integrator
// memory, then the n original indices by slot, then (if the tracker was
// started) AoS3 start positions and AoS3 int crossing counters.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "psim_ctx.hpp"

namespace {

constexpr char kMagic[8] = {'P', 'S', 'I', 'M', 'S', 'T', '0', '1'};

struct StateHeader {
    char magic[8];
    int32_t n;
    int32_t has_tracker;
    double box;
    double time;
    uint64_t step_index;
    int32_t vel_counter;
    int32_t reserved;
};

struct File {
    std::FILE *f = nullptr;
    explicit File(std::FILE *g) : f(g) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int psim_gpu_save(psim_ctx *c, const char *path)
{
    if (!c || !path || !c->have_state) return PSIM_E_ARG;
    const int n = c->n;
    const size_t m = 3 * (size_t)n;
    std::vector<double> arr(5 * m);
    StateHeader h{};
    std::memcpy(h.magic, kMagic, sizeof kMagic);
    h.n = n;
    h.box = c->prm.box;
    unsigned long long si = 0;
    int vc = 0;
    int rc = psim_gpu_get_state(c, arr.data(), arr.data() + m, arr.data() + 2 * m,
                                arr.data() + 3 * m, nullptr, arr.data() + 4 * m, &h.time, &vc,
                                &si);
    if (rc) return rc;
    h.step_index = si;
    h.vel_counter = vc;
    std::vector<int> order(n);
    CK(hipMemcpyAsync(order.data(), c->id, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    std::vector<double> start(m);
    std::vector<int> cross(m);
    rc = psim_gpu_get_tracker(c, start.data(), cross.data());
    if (rc && rc != PSIM_E_UNSUPPORTED) return rc;
    h.has_tracker = rc ? 0 : 1;
    File out(std::fopen(path, "wb"));
    if (!out.f) return PSIM_E_IO;
    bool ok = std::fwrite(&h, sizeof h, 1, out.f) == 1 &&
              std::fwrite(arr.data(), sizeof(double), arr.size(), out.f) == arr.size() &&
              std::fwrite(order.data(), sizeof(int), (size_t)n, out.f) == (size_t)n;
    if (ok && h.has_tracker)
        ok = std::fwrite(start.data(), sizeof(double), m, out.f) == m &&
             std::fwrite(cross.data(), sizeof(int), m, out.f) == m;
    const bool closed = std::fclose(out.f) == 0;
    out.f = nullptr;
    return (ok && closed) ? 0 : PSIM_E_IO;
}

int psim_gpu_load(psim_ctx *c, const char *path)
{
    if (!c || !path) return PSIM_E_ARG;
    File in(std::fopen(path, "rb"));
    if (!in.f) return PSIM_E_IO;
    StateHeader h{};
    if (std::fread(&h, sizeof h, 1, in.f) != 1 || std::memcmp(h.magic, kMagic, sizeof kMagic))
        return PSIM_E_IO;
    if (h.n != c->n || h.box != c->prm.box) return PSIM_E_ARG;
    const size_t n = (size_t)h.n, m = 3 * n;
    std::vector<double> arr(5 * m);
    std::vector<int> order(n);
    if (std::fread(arr.data(), sizeof(double), arr.size(), in.f) != arr.size() ||
        std::fread(order.data(), sizeof(int), n, in.f) != n)
        return PSIM_E_IO;
    std::vector<double> start;
    std::vector<int> cross;
    if (h.has_tracker) {
        start.resize(m);
        cross.resize(m);
        if (std::fread(start.data(), sizeof(double), m, in.f) != m ||
            std::fread(cross.data(), sizeof(int), m, in.f) != m)
            return PSIM_E_IO;
    }
    int rc = psim::eng::set_state(c, arr.data(), arr.data() + m, arr.data() + 2 * m,
                                  arr.data() + 3 * m, arr.data() + 4 * m, h.time, h.vel_counter,
                                  h.step_index, order.data());
    if (rc) return rc;
    if (h.has_tracker) rc = psim_gpu_set_tracker(c, start.data(), cross.data());
    return rc;
}

}  // extern "C"
