// ref_harness.cpp -- drives the reference's own compiled code through its
// public API so that the oracle restatement can be pinned against it.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/ref/Makefile from the reference
// sources where they lie (read-only); outputs go to oracle/_ref/ only.
//
// What runs reference code here:
//   * force plugins: dlopen("<plugin>.so") + dlsym("getForce") + factory(cfg)
//     exactly as loadForces does (ClusteredUI/src/runSim.cpp:29-61), then
//     defaultForceManager::getAcceleration (defaultForceManager.cpp:55-65);
//   * particle objects and their setters (particle.cpp:61-131);
//   * brownianIntegrator::nextSystem (brownianIntegrator.cpp:181-356) with its
//     internal fastRandom streams (fastRandom.cpp:40-124);
//   * fastRandom::g250 directly.
// What is restated here (system's members are private, system.h:20): the
// cell rebuild hash/sort/clear/reorder of systemHandling.cpp:33-85, using the
// reference comparator util::sortParticleTuple and __gnu_parallel::sort as
// the reference does.
#include "forceManager.h"
#include "integrator.h"
#include "utilities.h"
#include "fastRandom.h"
#include "RecoverySystem.h"
#include "system.h"

#include <dlfcn.h>
#include <parallel/algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

using namespace PSim;

namespace {

IForce *load_force(const char *so_path, config *cfg)
{
    // runSim.cpp:35-61
    static std::map<std::string, void *> libs;
    void *lib = libs[so_path];
    if (!lib) {
        lib = dlopen(so_path, RTLD_LAZY);
        if (!lib) {
            std::fprintf(stderr, "ref_harness: dlopen(%s) failed: %s\n", so_path, dlerror());
            return nullptr;
        }
        libs[so_path] = lib;
    }
    dlerror();
    create_Force *factory = (create_Force *)dlsym(lib, "getForce");
    if (dlerror()) return nullptr;
    return factory(cfg);
}

struct Cells {
    vector<tuple<int, int>> hash_index;
    vector<tuple<int, int>> start_end;
    vector<double> sorted;
};

// systemHandling.cpp:33-85 (hashParticles, sortParticles, clearCells,
// reorderParticles) on the harness side.
void rebuild(particle **parts, const systemState &st, Cells &c)
{
    int n = st.nParticles;
    int scale = st.cellScale;
    int ncell = scale * scale * scale;
    c.hash_index.assign(n, tuple<int, int>());
    c.start_end.assign(ncell, tuple<int, int>(0xffffffff, 0xffffffff));
    c.sorted.assign(4 * (size_t)n, 0.0);
    int sq = scale * scale;
    // the reference runs the hash and the reorder as OpenMP loops
    // (systemHandling.cpp:35, 55)
#pragma omp parallel for
    for (int i = 0; i < n; i++) {
        int cx = floor(parts[i]->getX() / st.cellSize);
        int cy = floor(parts[i]->getY() / st.cellSize);
        int cz = floor(parts[i]->getZ() / st.cellSize);
        get<0>(c.hash_index[i]) = cx + (scale * cy) + (sq * cz);
        get<1>(c.hash_index[i]) = i;
    }
    __gnu_parallel::sort(c.hash_index.begin(), c.hash_index.end(), util::sortParticleTuple);
#pragma omp parallel for
    for (int i = 0; i < n; i++) {
        int h = get<0>(c.hash_index[i]);
        if (i == 0) get<0>(c.start_end[h]) = 0;
        if (i > 0) {
            int prev = get<0>(c.hash_index[i - 1]);
            if (prev != h) {
                get<0>(c.start_end[h]) = i;
                get<1>(c.start_end[prev]) = i;
            }
        }
        if (i == n - 1) get<1>(c.start_end[h]) = n;
        int idx = get<1>(c.hash_index[i]);
        c.sorted[4 * i + 0] = parts[idx]->getX();
        c.sorted[4 * i + 1] = parts[idx]->getY();
        c.sorted[4 * i + 2] = parts[idx]->getZ();
        c.sorted[4 * i + 3] = parts[idx]->getRadius();
    }
}

particle **make_particles(int n, const double *pos, const double *vel,
                          double radius, double mass, int box)
{
    // the same setter sequence as system::initParticles (systemInit.cpp:78-88)
    particle **p = new particle *[n];
    for (int i = 0; i < n; i++) {
        p[i] = new particle(i);
        p[i]->setX(pos[3 * i + 0], box);
        p[i]->setY(pos[3 * i + 1], box);
        p[i]->setZ(pos[3 * i + 2], box);
        p[i]->setRadius(radius);
        p[i]->setMass(mass);
        if (vel) {
            p[i]->setVX(vel[3 * i + 0]);
            p[i]->setVY(vel[3 * i + 1]);
            p[i]->setVZ(vel[3 * i + 2]);
        }
    }
    return p;
}

void record(particle **p, int n, double *pos, double *pos0, double *vel,
            double *frc, double *frc0)
{
    for (int i = 0; i < n; i++) {
        pos[3 * i + 0] = p[i]->getX();  pos[3 * i + 1] = p[i]->getY();  pos[3 * i + 2] = p[i]->getZ();
        pos0[3 * i + 0] = p[i]->getX0(); pos0[3 * i + 1] = p[i]->getY0(); pos0[3 * i + 2] = p[i]->getZ0();
        vel[3 * i + 0] = p[i]->getVX(); vel[3 * i + 1] = p[i]->getVY(); vel[3 * i + 2] = p[i]->getVZ();
        if (frc) {
            frc[3 * i + 0] = p[i]->getFX(); frc[3 * i + 1] = p[i]->getFY(); frc[3 * i + 2] = p[i]->getFZ();
        }
        if (frc0) {
            frc0[3 * i + 0] = p[i]->getFX0(); frc0[3 * i + 1] = p[i]->getFY0(); frc0[3 * i + 2] = p[i]->getFZ0();
        }
    }
}

systemState make_state(int n, int box, int cell, int scale, double dt, int seed)
{
    systemState st = systemState();
    st.nParticles = n;
    st.boxSize = box;
    st.cellSize = cell;
    st.cellScale = scale;
    st.currentTime = 0;
    st.dTime = dt;
    st.seed = seed;
    return st;
}

// An integrator from a library that exports getIntegrator(config*) (a host
// that loads its integrator by name, as runSim.cpp:35-61 loads forces).
IIntegrator *load_integrator(const char *so_path, config *cfg)
{
    void *lib = dlopen(so_path, RTLD_LAZY);
    if (!lib) {
        std::fprintf(stderr, "ref_harness: dlopen(%s) failed: %s\n", so_path, dlerror());
        return nullptr;
    }
    dlerror();
    typedef IIntegrator *create_Integrator(config *);
    create_Integrator *factory = (create_Integrator *)dlsym(lib, "getIntegrator");
    if (dlerror()) return nullptr;
    return factory(cfg);
}

}  // namespace

extern "C" {

// fastRandom::g250 streams with threadSeed = seed*(i+1)
// (brownianIntegrator.cpp:204,236).
int ref_fastrandom(int seed, int n, int ndraws, double *out)
{
    for (int i = 0; i < n; i++) {
        fastRandom g;
        int thread_seed = seed * (i + 1);
        for (int k = 0; k < ndraws; k++) out[(size_t)i * ndraws + k] = g.g250(thread_seed);
    }
    return 0;
}

// One force evaluation through the reference plugin and force manager.
// Outputs: force3 by original index; the harness cell table (perm = orig
// index per sorted slot, start/end per cell) and sortedParticles.
int ref_forces(const char *so_path, const char *cfg_path, int n,
               const double *pos, double radius, int box, int cell, int scale,
               int nthreads, double *force3, int *perm, int *cstart,
               int *cend, double *sorted4)
{
    config cfg(cfg_path);
    IForce *f = load_force(so_path, &cfg);
    if (!f) return -1;
    particle **p = make_particles(n, pos, nullptr, radius, 1.0, box);
    systemState st = make_state(n, box, cell, scale, 0.001, 0);
    Cells c;
    rebuild(p, st, c);
    std::vector<double> force(3 * (size_t)n, 0.0);
    // the manager's destructor is not run (it deletes its own member,
    // defaultForceManager.cpp:37-42)
    defaultForceManager *mgr = new defaultForceManager();
    mgr->addForce(f);
    mgr->setNumThreads(nthreads > 0 ? nthreads : 1);
    mgr->getAcceleration(c.sorted.data(), force.data(), &c.hash_index, &c.start_end, &st);
    std::memcpy(force3, force.data(), sizeof(double) * 3 * n);
    int ncell = scale * scale * scale;
    for (int i = 0; i < n; i++) perm[i] = get<1>(c.hash_index[i]);
    for (int h = 0; h < ncell; h++) {
        cstart[h] = get<0>(c.start_end[h]);
        cend[h] = get<1>(c.start_end[h]);
    }
    std::memcpy(sorted4, c.sorted.data(), sizeof(double) * 4 * n);
    for (int i = 0; i < n; i++) delete p[i];
    delete[] p;
    return 0;
}

// nsteps of the system::run loop (system.cpp:187-211) without the contact
// pass and analysis output.  Trajectory arrays hold (nsteps+1) x 3n doubles;
// slot 0 is the state after construction (tr_pos NULL: no trajectory).
// loop_seconds receives the wall time of the step loop alone, from the end
// of step `skip` on (0: the whole loop).  integ_so NULL: the reference's
// brownianIntegrator; else the integrator that library's
// getIntegrator(config*) returns.
int ref_run(const char *so_path, const char *cfg_path, int n,
            const double *pos_in, const double *vel_in, double radius,
            double mass, int box, int cell, int scale, double dt, int nthreads,
            int nsteps, double *tr_pos, double *tr_pos0, double *tr_vel,
            double *tr_frc, double *tr_frc0, double *loop_seconds, const char *integ_so,
            int skip)
{
    config cfg(cfg_path);
    IForce *f = load_force(so_path, &cfg);
    if (!f) return -1;
    IIntegrator *integ = integ_so ? load_integrator(integ_so, &cfg)
                                  : new brownianIntegrator(&cfg);
    if (!integ) return -2;
    particle **p = make_particles(n, pos_in, vel_in, radius, mass, box);
    systemState st = make_state(n, box, cell, scale, dt, 0);
    defaultForceManager *mgr = new defaultForceManager();
    mgr->addForce(f);
    mgr->setNumThreads(nthreads > 0 ? nthreads : 1);
    std::vector<double> force(3 * (size_t)n, 0.0);
    size_t stride = 3 * (size_t)n;
    if (tr_pos) record(p, n, tr_pos, tr_pos0, tr_vel, tr_frc, tr_frc0);
    Cells c;
    double t0 = omp_get_wtime(), rec = 0.0;
    for (int s = 0; s < nsteps; s++) {
        rebuild(p, st, c);
        mgr->getAcceleration(c.sorted.data(), force.data(), &c.hash_index, &c.start_end, &st);
        // pushParticleForce (systemHandling.cpp:87-92)
#pragma omp parallel for
        for (int i = 0; i < n; i++) p[i]->setForce(&force[3 * (size_t)i]);
        integ->nextSystem(p, &st);
        st.currentTime += st.dTime;
        size_t o = (size_t)(s + 1) * stride;
        double r0 = omp_get_wtime();
        if (tr_pos) record(p, n, tr_pos + o, tr_pos0 + o, tr_vel + o, tr_frc + o, tr_frc0 + o);
        rec += omp_get_wtime() - r0;
        if (s + 1 == skip) {
            t0 = omp_get_wtime();
            rec = 0.0;
        }
    }
    if (loop_seconds) *loop_seconds = omp_get_wtime() - t0 - rec;
    for (int i = 0; i < n; i++) delete p[i];
    delete[] p;
    delete integ;
    return 0;
}

// As ref_run, recording positions, re-imaged previous positions and
// velocities only every `every` steps (slot k = state after k*every steps;
// nsteps/every + 1 slots), for long statistical runs.
int ref_run_sampled(const char *so_path, const char *cfg_path, int n,
                    const double *pos_in, const double *vel_in, double radius,
                    double mass, int box, int cell, int scale, double dt, int nthreads,
                    int nsteps, int every, double *tr_pos, double *tr_pos0,
                    double *tr_vel)
{
    if (every <= 0) return -2;
    config cfg(cfg_path);
    IForce *f = load_force(so_path, &cfg);
    if (!f) return -1;
    brownianIntegrator *integ = new brownianIntegrator(&cfg);
    particle **p = make_particles(n, pos_in, vel_in, radius, mass, box);
    systemState st = make_state(n, box, cell, scale, dt, 0);
    defaultForceManager *mgr = new defaultForceManager();
    mgr->addForce(f);
    mgr->setNumThreads(nthreads > 0 ? nthreads : 1);
    std::vector<double> force(3 * (size_t)n, 0.0);
    size_t stride = 3 * (size_t)n;
    record(p, n, tr_pos, tr_pos0, tr_vel, nullptr, nullptr);
    Cells c;
    for (int s = 0; s < nsteps; s++) {
        rebuild(p, st, c);
        mgr->getAcceleration(c.sorted.data(), force.data(), &c.hash_index, &c.start_end, &st);
        for (int i = 0; i < n; i++) p[i]->setForce(&force[3 * (size_t)i]);
        integ->nextSystem(p, &st);
        st.currentTime += st.dTime;
        if ((s + 1) % every == 0) {
            size_t o = (size_t)((s + 1) / every) * stride;
            record(p, n, tr_pos + o, tr_pos0 + o, tr_vel + o, nullptr, nullptr);
        }
    }
    for (int i = 0; i < n; i++) delete p[i];
    delete[] p;
    delete integ;
    return 0;
}

// The reference's own system construction (system.cpp:108-140): the
// constructor places the particles (systemInit.cpp:31-267) and writes
// <trial_dir>/sysConfig.cfg, <trial_dir>/initialState.txt (recovery format,
// analysisOutput.cpp:53-74) and <trial_dir>/movie/system-initial.xyz.  The
// directories must exist.  The system is not destroyed (its destructor
// deletes the integrator and force manager it was handed, here none).
int ref_init(const char *cfg_path, const char *trial_dir)
{
    config cfg(cfg_path);
    new PSim::system(&cfg, nullptr, nullptr, trial_dir);
    return 0;
}

// RecoverySystem (RecoverySystem.cpp:27-43): reads
// <sys_state>/snapshots/time-<stamp>/recovery.txt (systemRecovery.cpp:43-81)
// and writes the loaded particles to <sys_state>/-rewind-<stamp>/initialState.txt.
int ref_recover(const char *cfg_path, const char *sys_state, const char *stamp)
{
    config cfg(cfg_path);
    new PSim::RecoverySystem(&cfg, sys_state, stamp, nullptr, nullptr);
    return 0;
}

}  // extern "C"
