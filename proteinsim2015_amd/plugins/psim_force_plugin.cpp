// psim_force_plugin.cpp -- GPU force plugins for a PSim host.
//
// Built three times (LJ.so, AO.so, Cal.so) with PSIM_PLUGIN_MASK selecting the
// module.  A host that loads the reference's plugins with
//   dlopen("./" + force + ".so") ; dlsym("getForce") ; factory(cfg)
// (ClusteredUI/src/runSim.cpp:29-61) loads these unchanged.
//
// IForce::getAcceleration is called once per sorted particle, concurrently
// from the host's OpenMP threads (defaultForceManager.cpp:55-65).  The first
// call of a force evaluation computes every particle's force on the GPU from
// the host's own sortedParticles / particleHashIndex / cellStartEnd (through
// psim_gpu_force_table) and caches it; the remaining calls copy the cached
// force of their particle.  The cache is keyed on the state's time, N and the
// sortedParticles contents, so a second evaluation at the same time with
// moved particles recomputes.
//
// A new host can skip the per-particle calls: dlsym("psim_force_batch") and
// call it once per step.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "psim/IForce.h"
#include "psim_gpu.h"

#ifndef PSIM_PLUGIN_MASK
#error "PSIM_PLUGIN_MASK must name the force module"
#endif

namespace PSim {
// Typed parameter lookup (config.cpp:97-175): std::stoi / std::stof /
// std::stod read the leading number of the value.
template <>
int config::getParam<int>(std::string key, int def)
{
    return containsKey(key) ? std::stoi(options[key], nullptr) : def;
}
template <>
float config::getParam<float>(std::string key, float def)
{
    return containsKey(key) ? std::stof(options[key], nullptr) : def;
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

namespace {

const char *status_text(int st) { return psim_gpu_status_string(st); }

class GpuForce : public PSim::IForce {
public:
    explicit GpuForce(PSim::config *cfg)
    {
        // parameter keys and defaults of the reference plugin constructors
        std::memset(&prm_, 0, sizeof(prm_));
        prm_.force_mask = PSIM_PLUGIN_MASK;
        prm_.radius = cfg->getParam<double>("radius", 0.5);
        prm_.mass = cfg->getParam<double>("mass", 1.0);
        prm_.kT = cfg->getParam<double>("kT", 1.0);
        prm_.gamma = cfg->getParam<double>("gamma", 0.5);
        prm_.dt = cfg->getParam<double>("timeStep", 0.001);
        prm_.vel_freq = cfg->getParam<int>("velFreq", 1000);
        prm_.seed = (unsigned long long)cfg->getParam<int>("seed", 90210);
        prm_.lj_num = 18;
        prm_.lj_well_depth = 10.0;
        prm_.yukawa_strength = 8.0;
        prm_.debye_length = 0.5;
        prm_.ao_well_depth = 0.261;
        if (PSIM_PLUGIN_MASK == PSIM_FORCE_LJ) {
            // LJPotential.cpp:34-55
            name = "Lennard Jones (MI355X)";
            prm_.lj_well_depth = cfg->getParam<double>("wellDepth", 10.0);
            prm_.yukawa_strength = cfg->getParam<double>("yukawaStrength", 8.0);
            prm_.lj_num = cfg->getParam<int>("ljNum", 18);
            prm_.cutoff = cfg->getParam<double>("cutOff", 2.5);
            prm_.debye_length = cfg->getParam<double>("debyeLength", 0.5);
        } else if (PSIM_PLUGIN_MASK == PSIM_FORCE_AO) {
            // AOPotential.cpp:40-47
            name = "AOPotential (MI355X)";
            prm_.ao_well_depth = cfg->getParam<double>("wellDepth", 0.261);
            prm_.cutoff = cfg->getParam<double>("cutOff", 1.1);
        } else {
            // Calibration.cpp:36
            name = "Calibration (MI355X)";
            prm_.cutoff = cfg->getParam<double>("cutOff", 1.1);
        }
    }

    ~GpuForce() override
    {
        if (ctx_) psim_gpu_destroy(ctx_);
    }

    void getAcceleration(int index, double *sortedParticles, double *particleForce,
                         std::vector<std::tuple<int, int>> *particleHashIndex,
                         std::vector<std::tuple<int, int>> *cellStartEnd,
                         PSim::systemState *state) override
    {
        {
            std::lock_guard<std::mutex> lock(mu_);
            if (!valid(sortedParticles, state)) {
                int st = evaluate(sortedParticles, *particleHashIndex, *cellStartEnd, state);
                if (st > 0) {
                    // the reference reports the error and exits with its code
                    // (error.cpp:47-62)
                    std::fprintf(stderr, "%s: %s\n", name.c_str(), status_text(st));
                    std::exit(st);
                }
                if (st < 0) {
                    std::fprintf(stderr, "%s: GPU evaluation failed: %s\n", name.c_str(),
                                 status_text(st));
                    std::exit(100);
                }
            }
        }
        const int o = std::get<1>((*particleHashIndex)[index]);
        particleForce[3 * (size_t)o] = force_[3 * (size_t)o];
        particleForce[3 * (size_t)o + 1] = force_[3 * (size_t)o + 1];
        particleForce[3 * (size_t)o + 2] = force_[3 * (size_t)o + 2];
    }

    bool isTimeDependent() override { return false; }

    int batch(const double *sorted4, double *force, const int *hash_index,
              const int *cell_start_end, const PSim::systemState *state)
    {
        std::lock_guard<std::mutex> lock(mu_);
        int rc = ensure_ctx(state);
        if (rc) return rc;
        rc = psim_gpu_force_table(ctx_, sorted4, hash_index, cell_start_end, force);
        if (rc >= 0 && ctx_) psim_gpu_clear_error(ctx_);
        return rc;
    }

private:
    bool valid(const double *sorted, const PSim::systemState *st) const
    {
        if (!have_ || st->nParticles != n_ || st->currentTime != time_) return false;
        return std::memcmp(sorted, sorted_copy_.data(), sizeof(double) * 4 * (size_t)n_) == 0;
    }

    int ensure_ctx(const PSim::systemState *st)
    {
        if (ctx_ && st->nParticles == n_ && st->boxSize == box_ && st->cellSize == cell_ &&
            st->cellScale == scale_)
            return 0;
        if (ctx_) psim_gpu_destroy(ctx_);
        ctx_ = nullptr;
        psim_params p = prm_;
        p.n_particles = st->nParticles;
        p.cells_per_dim = st->cellScale;
        p.box = (double)st->boxSize;          // integer geometry (systemState.h:10-12)
        p.cell_size = (double)st->cellSize;
        p.half_uv = 0.0;                      // integer L/2, utilities.cpp:124
        int rc = psim_gpu_create(&p, &ctx_);
        if (rc) return rc;
        n_ = st->nParticles;
        box_ = st->boxSize;
        cell_ = st->cellSize;
        scale_ = st->cellScale;
        return 0;
    }

    int evaluate(const double *sorted, const std::vector<std::tuple<int, int>> &phi,
                 const std::vector<std::tuple<int, int>> &cse, const PSim::systemState *st)
    {
        int rc = ensure_ctx(st);
        if (rc) return rc;
        const size_t n = (size_t)st->nParticles;
        hi_.resize(2 * n);
        for (size_t i = 0; i < n; ++i) {
            hi_[2 * i] = std::get<0>(phi[i]);
            hi_[2 * i + 1] = std::get<1>(phi[i]);
        }
        cse_.resize(2 * cse.size());
        for (size_t h = 0; h < cse.size(); ++h) {
            cse_[2 * h] = std::get<0>(cse[h]);
            cse_[2 * h + 1] = std::get<1>(cse[h]);
        }
        force_.resize(3 * n);
        rc = psim_gpu_force_table(ctx_, sorted, hi_.data(), cse_.data(), force_.data());
        if (rc) return rc;
        sorted_copy_.assign(sorted, sorted + 4 * n);
        time_ = st->currentTime;
        have_ = true;
        return 0;
    }

    psim_params prm_;
    psim_ctx *ctx_ = nullptr;
    std::mutex mu_;
    bool have_ = false;
    int n_ = 0, box_ = 0, cell_ = 0, scale_ = 0;
    double time_ = 0.0;
    std::vector<double> sorted_copy_, force_;
    std::vector<int> hi_, cse_;
};

}  // namespace

// The plugin factory (LJPotential.h:74-77 / AOPotential.h:77-80 /
// Calibration.h:65-68).
extern "C" PSim::IForce *getForce(PSim::config *cfg) { return new GpuForce(cfg); }

// Whole-system force evaluation in one call: the reference's arrays with
// particleHashIndex and cellStartEnd flattened to int pairs.  Returns 0, the
// reference's error code (7704/7705) or a negative PSIM_E_* status.
extern "C" int psim_force_batch(void *self, const double *sorted4, double *force,
                                const int *hash_index, const int *cell_start_end,
                                const PSim::systemState *state)
{
    if (!self || !sorted4 || !force || !hash_index || !cell_start_end || !state)
        return PSIM_E_ARG;
    return static_cast<GpuForce *>(static_cast<PSim::IForce *>(self))
        ->batch(sorted4, force, hash_index, cell_start_end, state);
}
