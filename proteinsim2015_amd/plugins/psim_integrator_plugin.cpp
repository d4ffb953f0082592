// psim_integrator_plugin.cpp -- a GPU G&B integrator for a PSim host
// (GBIntegrator.so).
//
// The reference host constructs its integrator itself
// (ClusteredUI/src/runSim.cpp:127, new brownianIntegrator(cfg)), hands it to
// the system (system.cpp:69) and calls
//   integrator->nextSystem(particles, &state)
// once per step after the force push (system.cpp:196).  This library gives a
// host the same object from the GPU:
//   extern "C" PSim::IIntegrator *getIntegrator(PSim::config *cfg)
// reads the constructor's keys with their defaults (nParticles, velFreq, kT,
// gamma, timeStep, seed; brownianIntegrator.cpp:33-83) and returns an
// IIntegrator whose nextSystem gathers the particles' positions, previous
// positions, velocities, forces and masses, advances them on the GPU
// (psim_int_next: firstStep / normalStep / velocityStep with the reference's
// operation order and, by default, the reference's own per-particle noise
// streams) and stores the results back into the particles, as
// particle::setPos and setV* leave them (particle.cpp:61-99).
//
// Extra key: psimNoise = reference (default) | philox.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "psim/IIntegrator.h"
#include "psim/particle.h"
#include "psim_gpu.h"

namespace PSim {
// Typed parameter lookup (config.cpp:97-175): std::stoi / std::stod read the
// leading number of the value.
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

// The members of a host particle the integrator reads and writes.
struct psim_particle_access {
    static void gather(const PSim::particle *p, double *x, double *x0, double *v, double *f,
                       double *f0, double *m)
    {
        x[0] = p->pos.x;   x[1] = p->pos.y;   x[2] = p->pos.z;
        x0[0] = p->pos0.x; x0[1] = p->pos0.y; x0[2] = p->pos0.z;
        v[0] = p->vel.x;   v[1] = p->vel.y;   v[2] = p->vel.z;
        f[0] = p->frc.x;   f[1] = p->frc.y;   f[2] = p->frc.z;
        f0[0] = p->frc0.x; f0[1] = p->frc0.y; f0[2] = p->frc0.z;
        *m = p->m;
    }
    static void store(PSim::particle *p, const double *x, const double *x0, const double *v)
    {
        p->pos.x = x[0];   p->pos.y = x[1];   p->pos.z = x[2];
        p->pos0.x = x0[0]; p->pos0.y = x0[1]; p->pos0.z = x0[2];
        p->vel.x = v[0];   p->vel.y = v[1];   p->vel.z = v[2];
    }
    static int name(const PSim::particle *p) { return p->name; }
};

namespace {

class GpuBrownianIntegrator : public PSim::IIntegrator {
public:
    explicit GpuBrownianIntegrator(PSim::config *cfg)
    {
        name = "brownianIntegrator (MI355X)";
        std::memset(&prm_, 0, sizeof(prm_));
        prm_.n_particles = cfg->getParam<int>("nParticles", 1000);
        prm_.vel_freq = cfg->getParam<int>("velFreq", 1000);
        prm_.kT = cfg->getParam<double>("kT", 1.0);
        prm_.mass = cfg->getParam<double>("mass", 1.0);
        prm_.gamma = cfg->getParam<double>("gamma", 0.5);
        prm_.dt = cfg->getParam<double>("timeStep", 0.001);
        prm_.seed = (unsigned long long)(unsigned)cfg->getParam<int>("seed", 90210);
        const std::string noise = cfg->getParam<std::string>("psimNoise", "reference");
        noise_ = noise == "philox" ? PSIM_NOISE_PHILOX : PSIM_NOISE_REFERENCE;
        const int rc = psim_int_create(&prm_, noise_, &g_);
        if (rc) {
            std::fprintf(stderr, "%s: GPU integrator could not be created: %s\n", name.c_str(),
                         psim_gpu_status_string(rc));
            std::exit(100);
        }
        const size_t n3 = 3 * (size_t)prm_.n_particles;
        x_.resize(n3);
        x0_.resize(n3);
        v_.resize(n3);
        f_.resize(n3);
        f0_.resize(n3);
        m_.resize((size_t)prm_.n_particles);
    }

    ~GpuBrownianIntegrator() override { psim_int_destroy(g_); }

    int nextSystem(PSim::particle **items, PSim::systemState *state) override
    {
        const int n = state->nParticles;
        if (n != prm_.n_particles) {
            std::fprintf(stderr, "%s: the system has %d particles, the integrator %d\n",
                         name.c_str(), n, prm_.n_particles);
            std::exit(100);
        }
#pragma omp parallel for
        for (int i = 0; i < n; ++i) {
            const size_t b = 3 * (size_t)i;
            psim_particle_access::gather(items[i], &x_[b], &x0_[b], &v_[b], &f_[b], &f0_[b],
                                         &m_[i]);
        }
        int bad = -1;
        const int rc = psim_int_next(g_, state->currentTime, state->boxSize, x_.data(),
                                     x0_.data(), v_.data(), f_.data(), f0_.data(), m_.data(),
                                     &bad);
        if (rc < 0 && rc != PSIM_E_UNSUPPORTED) {
            std::fprintf(stderr, "%s: GPU step failed: %s\n", name.c_str(),
                         psim_gpu_status_string(rc));
            std::exit(100);
        }
#pragma omp parallel for
        for (int i = 0; i < n; ++i) {
            const size_t b = 3 * (size_t)i;
            psim_particle_access::store(items[i], &x_[b], &x0_[b], &v_[b]);
        }
        if (rc == PSIM_ERR_BOUNDS) {
            // particle::setX's bounds check exits with the reference's code
            // (error.cpp:40-46)
            std::fprintf(stderr, "%s: particle %d out of bounds\n", name.c_str(),
                         bad >= 0 && bad < n ? psim_particle_access::name(items[bad]) : bad);
            std::exit(PSIM_ERR_BOUNDS);
        }
        if (rc == PSIM_E_UNSUPPORTED) {
            std::fprintf(stderr,
                         "%s: particle %d: the reference's noise stream seeded with "
                         "seed*(i+1) = a negative int re-seeds on every draw and yields no "
                         "deviate (the reference loops forever here); use a smaller seed or "
                         "psimNoise = philox\n",
                         name.c_str(), bad);
            std::exit(100);
        }
        return 0;
    }

private:
    psim_params prm_;
    int noise_ = PSIM_NOISE_REFERENCE;
    psim_integrator *g_ = nullptr;
    std::vector<double> x_, x0_, v_, f_, f0_, m_;
};

}  // namespace

// The integrator factory (no counterpart in the reference, which constructs
// brownianIntegrator directly, runSim.cpp:127; the name follows getForce).
extern "C" PSim::IIntegrator *getIntegrator(PSim::config *cfg)
{
    return new GpuBrownianIntegrator(cfg);
}
