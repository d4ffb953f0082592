// psim_host_demo.cpp -- a minimal PSim host for force plugins.
//
// Mirrors how the reference host finds and drives a force plugin:
//   settings file (key = value)          config.cpp:27-69
//   geometry from nParticles/conc/scale  system.cpp:50-64
//   dlopen(plugin) + dlsym("getForce")   runSim.cpp:29-61 (exit 100 on failure)
//   cell build: hash, sort, start/end    systemHandling.cpp:33-85
//   per-particle getAcceleration loop    defaultForceManager.cpp:55-65
// and, with mode "batch", the plugin's psim_force_batch entry instead.
//
//   psim_host_demo <plugin.so> <settings.cfg> <pos.f64> <force.f64> [batch|each] [evals]
//
// pos.f64 holds nParticles x 3 doubles (original order); force.f64 receives
// particleForce (nParticles x 3 doubles, original order).  With evals > 1
// the evaluation is repeated at currentTime = k * timeStep, k = 0..evals-1
// (one force evaluation per step, as system::run makes), and the mean wall
// time of the evaluations after the first is printed.
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "psim/IForce.h"

namespace PSim {
template <>
int config::getParam<int>(std::string key, int def)
{
    return containsKey(key) ? std::stoi(options[key], nullptr) : def;
}
template <>
double config::getParam<double>(std::string key, double def)
{
    return containsKey(key) ? std::stod(options[key], nullptr) : def;
}
template <>
std::string config::getParam<std::string>(std::string key, std::string def)
{
    return containsKey(key) ? options[key] : def;
}
}  // namespace PSim

typedef int batch_fn(void *, const double *, double *, const int *, const int *,
                     const PSim::systemState *);

static std::string trim(const std::string &s)
{
    size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s plugin.so settings.cfg pos.f64 force.f64 [batch|each]\n",
                     argv[0]);
        return 2;
    }
    const std::string mode = argc > 5 ? argv[5] : "each";
    const int evals = argc > 6 ? std::max(1, std::atoi(argv[6])) : 1;
    PSim::config cfg;
    std::ifstream in(argv[2]);
    for (std::string line; std::getline(in, line);) {
        if (line.empty() || line[0] == '#') continue;
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        cfg.set(trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
    PSim::systemState st{};
    st.nParticles = cfg.getParam<int>("nParticles", 1000);
    const double conc = cfg.getParam<double>("conc", 0.01);
    const int scale = cfg.getParam<int>("scale", 4);
    const double r = cfg.getParam<double>("radius", 0.5);
    const double vP = st.nParticles * (4.0 / 3.0) * std::atan(1.0) * 4.0 * r * r * r;
    st.boxSize = (int)std::cbrt(vP / conc);
    st.cellSize = st.boxSize / scale;
    st.boxSize = st.cellSize * scale;
    st.cellScale = scale;
    st.dTime = cfg.getParam<double>("timeStep", 0.001);
    const int n = st.nParticles;

    std::vector<double> pos(3 * (size_t)n);
    std::ifstream pf(argv[3], std::ios::binary);
    if (!pf.read(reinterpret_cast<char *>(pos.data()), sizeof(double) * pos.size())) {
        std::fprintf(stderr, "cannot read %d positions from %s\n", n, argv[3]);
        return 2;
    }
    // cell build (stable sort: the within-cell order is the original order)
    const int S = st.cellScale;
    std::vector<std::tuple<int, int>> phi(n);
    for (int i = 0; i < n; ++i) {
        int c[3];
        for (int d = 0; d < 3; ++d) {
            c[d] = (int)std::floor(pos[3 * (size_t)i + d] / st.cellSize);
            c[d] = std::min(std::max(c[d], 0), S - 1);
        }
        phi[i] = std::make_tuple(c[0] + S * c[1] + S * S * c[2], i);
    }
    std::stable_sort(phi.begin(), phi.end(), [](const std::tuple<int, int> &a,
                                                const std::tuple<int, int> &b) {
        return std::get<0>(a) < std::get<0>(b);
    });
    std::vector<std::tuple<int, int>> cse((size_t)S * S * S, std::make_tuple(-1, -1));
    std::vector<double> sorted(4 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        int h = std::get<0>(phi[i]);
        if (i == 0 || std::get<0>(phi[i - 1]) != h) std::get<0>(cse[h]) = i;
        if (i == n - 1 || std::get<0>(phi[i + 1]) != h) std::get<1>(cse[h]) = i + 1;
        int o = std::get<1>(phi[i]);
        for (int d = 0; d < 3; ++d) sorted[4 * (size_t)i + d] = pos[3 * (size_t)o + d];
        sorted[4 * (size_t)i + 3] = r;
    }

    void *lib = dlopen(argv[1], RTLD_LAZY);
    if (!lib) {
        std::fprintf(stderr, "Error loading in force library: %s\n", dlerror());
        return 100;
    }
    dlerror();
    PSim::create_Force *factory = (PSim::create_Force *)dlsym(lib, "getForce");
    if (dlerror()) {
        std::fprintf(stderr, "Could not find symbol: getForce\n");
        return 100;
    }
    PSim::IForce *force = factory(&cfg);
    std::vector<double> out(3 * (size_t)n, 0.0);
    batch_fn *batch = nullptr;
    std::vector<int> hi, ce;
    if (mode == "batch") {
        batch = (batch_fn *)dlsym(lib, "psim_force_batch");
        if (!batch) {
            std::fprintf(stderr, "plugin has no psim_force_batch\n");
            return 100;
        }
        hi.resize(2 * (size_t)n);
        ce.resize(2 * cse.size());
        for (int i = 0; i < n; ++i) {
            hi[2 * i] = std::get<0>(phi[i]);
            hi[2 * i + 1] = std::get<1>(phi[i]);
        }
        for (size_t h = 0; h < cse.size(); ++h) {
            ce[2 * h] = std::get<0>(cse[h]);
            ce[2 * h + 1] = std::get<1>(cse[h]);
        }
    }
    double later = 0.0;
    for (int k = 0; k < evals; ++k) {
        st.currentTime = k * st.dTime;
        const auto t0 = std::chrono::steady_clock::now();
        if (batch) {
            int rc = batch(force, sorted.data(), out.data(), hi.data(), ce.data(), &st);
            if (rc) {
                std::fprintf(stderr, "psim_force_batch returned %d\n", rc);
                return rc > 0 ? rc : 3;
            }
        } else {
#pragma omp parallel for
            for (int i = 0; i < n; ++i)
                force->getAcceleration(i, sorted.data(), out.data(), &phi, &cse, &st);
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                              .count();
        if (k > 0) later += dt;
    }
    if (evals > 1) std::printf("seconds per evaluation: %.6e\n", later / (evals - 1));
    std::ofstream of(argv[4], std::ios::binary);
    of.write(reinterpret_cast<const char *>(out.data()), sizeof(double) * out.size());
    std::printf("%s: %d particles, L=%d, cell=%d, S=%d, mode=%s\n", force->getName().c_str(), n,
                st.boxSize, st.cellSize, S, mode.c_str());
    return 0;
}
