// psim_sim.cpp -- the simulation program: a settings file in, a trial
// directory of snapshots, movie frames and observable series out.
//
// Mirrors the reference's command line and run loop:
//   main (ClusteredUI/src/main.cpp:47-107)      -r DIR STAMP, -a DIR STAMP tests...
//   runScript (ClusteredUI/src/runSim.cpp:115-159) settings.cfg from the working
//                                              directory, force / integrator /
//                                              system, run to endTime
//   runAnalysis (ClusteredUI/src/analysis.cpp:28-49) + AnalysisSystem
//                                              (AnalysisSystem.cpp:27-47) +
//                                              postAnalysis (analysisPost.cpp:26-52)
//   system::system + writeSystemInit (system.cpp:69-140)
//   system::run (system.cpp:170-211) with analysisManager::writeRunTimeState
//                                              (analysisManager.cpp:56-67) and
//                                              writeSystemState (analysisOutput.cpp:97-145)
//   RecoverySystem (RecoverySystem.cpp:27-43)
// over the C ABI of libpsim_gpu.so: the steps run on the GPU, the files are
// written by the host.
//
// Differences from the reference, on purpose:
//   * no interactive prompts (the reference asks y/n before running and
//     before overwriting a trial directory): the run starts at once and an
//     existing directory is written into;
//   * the force is chosen by name (force = LJ | AO | Cal, or the class names
//     LennardJones | AOPotential | Calibration; several joined by '+' are
//     summed) and computed by the GPU engine rather than dlopen'ed;
//   * every snapshot directory also receives state.psim (full precision);
//     -r resumes from it bit for bit when it is there, and from recovery.txt
//     with the reference's semantics (velocities and integrator memory zero,
//     time restarted at 0, so the first step takes the first-step branch)
//     otherwise or with --from-recovery;
//   * the tracked mean-square displacement is accumulated in FP64 (the
//     reference uses float); the members of a cluster file are listed in
//     index order (the reference lists them in its search order).
#include <sys/stat.h>

#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "psim_gpu.h"

namespace {

// ---- settings (config.cpp:27-69, 97-175) -------------------------------------
struct Settings {
    std::map<std::string, std::string> options;

    static std::string trim(const std::string &s)
    {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace((unsigned char)s[a])) ++a;
        while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
        return s.substr(a, b - a);
    }
    bool load(const std::string &path)
    {
        std::ifstream in(path);
        if (!in) return false;
        for (std::string line; std::getline(in, line);) {
            if (line.empty() || line[0] == '#') continue;
            const size_t eq = line.find('=');
            if (eq == std::string::npos) {
                options[trim(line)] = trim(line);   // config::parseLine keeps the whole line
                continue;
            }
            options[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
        }
        return true;
    }
    bool has(const std::string &k) const { return options.count(k) != 0; }
    int get_int(const std::string &k, int def) const
    {
        return has(k) ? std::stoi(options.at(k)) : def;
    }
    double get_double(const std::string &k, double def) const
    {
        return has(k) ? std::stod(options.at(k)) : def;
    }
    std::string get_string(const std::string &k, const std::string &def) const
    {
        return has(k) ? options.at(k) : def;
    }
};

[[noreturn]] void die(int code, const std::string &msg)
{
    std::fprintf(stderr, "\n%s\n", msg.c_str());
    std::exit(code);
}

void check(int rc, const char *what)
{
    if (rc == 0) return;
    die(rc > 0 ? rc : 1, std::string(what) + ": " + psim_gpu_status_string(rc));
}

std::string fmt_g(double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%g", v);
    return b;
}

void make_dir(const std::string &p)
{
    if (mkdir(p.c_str(), 0777) != 0 && errno != EEXIST)
        die(1, "cannot create directory " + p);
}

bool file_exists(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// ---- the system (system.cpp:27-67, runSim.cpp:29-79) ---------------------------
struct System {
    Settings cfg;
    psim_params prm{};
    psim_ctx *ctx = nullptr;
    std::string trial;
    int n = 0, box = 0, cell = 0, scale = 0, seed = 0, output_freq = 0;
    double conc = 0, temp = 0, dt = 0, radius = 0, mass = 0;
    double cycle_hour = 0;   // (endTime / dt) / 3600, as system::run sets it

    void geometry()
    {
        n = cfg.get_int("nParticles", 1000);
        conc = cfg.get_double("conc", 0.01);
        scale = cfg.get_int("scale", 4);
        radius = cfg.get_double("radius", 0.5);
        mass = cfg.get_double("mass", 1.0);
        dt = cfg.get_double("timeStep", 0.001);
        seed = cfg.get_int("seed", 90210);
        temp = cfg.get_double("kT", 1.0);
        output_freq = cfg.get_int("outputFreq", int(1.0 / dt));
        cycle_hour = ((double)(int)cfg.get_double("endTime", 1000) / dt) / 3600.0;
        const double vP = n * (4.0 / 3.0) * std::atan(1.0) * 4.0 * radius * radius * radius;
        box = (int)std::cbrt(vP / conc);
        if (scale <= 0) die(1, "scale must be positive");
        cell = box / scale;
        box = cell * scale;
        conc = vP / std::pow((double)box, 3.0);
        if (n < 1 || scale < 3 || cell < 1 || output_freq < 1)
            die(1, "settings give nParticles < 1, scale < 3, a zero cell size or outputFreq < 1");
    }

    void forces()
    {
        const std::string name = cfg.get_string("force", "");
        int mask = 0;
        std::stringstream ss(name);
        for (std::string part; std::getline(ss, part, '+');) {
            part = Settings::trim(part);
            if (part == "LJ" || part == "LennardJones") mask |= PSIM_FORCE_LJ;
            else if (part == "AO" || part == "AOPotential") mask |= PSIM_FORCE_AO;
            else if (part == "Cal" || part == "Calibration") mask |= PSIM_FORCE_CAL;
            else die(100, "Error loading in force library: '" + part + "'");   // runSim.cpp:41-45
        }
        if (!mask) die(100, "Error loading in force library: no force given");
        prm.force_mask = mask;
        // the plugin constructors' keys and defaults (LJPotential.cpp:34-55,
        // AOPotential.cpp:40-55, Calibration.cpp:30-39); cutOff is one key
        const double cut_default = (mask & PSIM_FORCE_LJ) ? 2.5 : 1.1;
        prm.cutoff = cfg.get_double("cutOff", cut_default);
        prm.lj_well_depth = cfg.get_double("wellDepth", 10.0);
        prm.ao_well_depth = cfg.get_double("wellDepth", 0.261);
        prm.yukawa_strength = cfg.get_double("yukawaStrength", 8.0);
        prm.lj_num = cfg.get_int("ljNum", 18);
        prm.debye_length = cfg.get_double("debyeLength", 0.5);
    }

    void create(const std::string &kernel_path)
    {
        geometry();
        forces();
        prm.n_particles = n;
        prm.cells_per_dim = scale;
        prm.box = box;
        prm.cell_size = cell;
        prm.half_uv = 0.0;
        prm.radius = radius;
        prm.mass = mass;
        prm.kT = temp;
        prm.gamma = cfg.get_double("gamma", 0.5);
        prm.dt = dt;
        prm.vel_freq = cfg.get_int("velFreq", 1000);
        prm.seed = (unsigned long long)(long long)seed;
        prm.kernel_path = kernel_path == "sparse"  ? 1
                          : kernel_path == "dense" ? 2
                          : kernel_path == "brick" ? 3
                                                   : 0;
        check(psim_gpu_create(&prm, &ctx), "psim_gpu_create");
        std::printf("---Number of Particles: %d\n---Box Size: %d\n---Cell Size: %d\n\n", n, box,
                    cell);
    }

    // system::writeSystemInit (system.cpp:108-140) + writeInitialState
    // (analysisManager.cpp:49-54)
    void write_init(const std::vector<double> &pos, const std::vector<double> &pos0,
                    const std::vector<double> &frc, const std::vector<double> &frc0,
                    const std::string &settings_path)
    {
        check(psim_write_sys_config((trial + "/sysConfig.cfg").c_str(), trial.c_str(), n, conc,
                                    box, cell, scale, temp, dt, output_freq, cycle_hour, seed),
              "sysConfig.cfg");
        check(psim_write_recovery((trial + "/initialState.txt").c_str(), n, pos.data(),
                                  pos0.data(), frc.data(), frc0.data(), mass, radius),
              "initialState.txt");
        check(psim_write_xyz((trial + "/movie/system-initial.xyz").c_str(), n, pos.data(), 0.0),
              "system-initial.xyz");
        std::ifstream in(settings_path, std::ios::binary);
        std::ofstream out(trial + "/settings.cfg", std::ios::binary);
        if (in) out << in.rdbuf();
    }
};

void append(const std::string &path, double t, double value)
{
    std::ofstream f(path, std::ios_base::app | std::ios_base::out);
    f << fmt_g(t) << " " << fmt_g(value) << "\n";
}

// particle::calculatePotential (particle.cpp:133-149), in float as there
float particle_potential(const double *x, const double *x0, const double *f, const double *f0)
{
    float dx = x[0] - x0[0], dy = x[1] - x0[1], dz = x[2] - x0[2];
    float dfx = f[0] - f0[0], dfy = f[1] - f0[1], dfz = f[2] - f0[2];
    float dr = std::sqrt(dx * dx + dy * dy + dz * dz);
    float df = std::sqrt(dfx * dfx + dfy * dfy + dfz * dfz);
    float fz0 = std::sqrt(f0[0] * f0[0] + f0[1] * f0[1] + f0[2] * f0[2]);
    float p = (fz0 * dr) + (0.5 * df * dr);
    return -p;
}

// Components of the contact graph with more than four members, numbered by
// their smallest member (the order analysisCluster.cpp:26-83 finds them in).
std::vector<std::vector<int>> clusters_of(psim_ctx *ctx, int n)
{
    std::vector<int> lab(n);
    int nc = 0, mean = 0;
    check(psim_gpu_clusters(ctx, 1.2, 5, lab.data(), &nc, &mean), "psim_gpu_clusters");
    std::map<int, std::vector<int>> by;
    for (int i = 0; i < n; ++i) by[lab[i]].push_back(i);
    std::vector<std::vector<int>> out;
    for (auto &kv : by)
        if (kv.second.size() > 4) out.push_back(kv.second);
    return out;
}

void histogram_file(const std::string &path, const std::map<int, int> &h)
{
    std::ofstream f(path);
    for (auto &kv : h) f << kv.first << " " << kv.second << "\n";
}

// ---- one snapshot (analysisOutput.cpp:97-145) -----------------------------------
void write_snapshot(System &S, double t, const std::vector<double> &frc0, bool save_state)
{
    const int n = S.n;
    const size_t m = 3 * (size_t)n;
    std::vector<double> pos(m), pos0(m), frc(m);
    check(psim_gpu_get_state(S.ctx, pos.data(), pos0.data(), nullptr, frc.data(), nullptr,
                             nullptr, nullptr, nullptr, nullptr),
          "psim_gpu_get_state");
    const std::string name = std::to_string((int)std::round(t));
    std::printf("Writing: %s.txt\n", name.c_str());
    const std::string dir = S.trial + "/snapshots/time-" + name;
    make_dir(dir);
    check(psim_write_recovery((dir + "/recovery.txt").c_str(), n, pos.data(), pos0.data(),
                              frc.data(), frc0.data(), S.mass, S.radius),
          "recovery.txt");
    if (save_state) check(psim_gpu_save(S.ctx, (dir + "/state.psim").c_str()), "state.psim");
    check(psim_write_xyz((S.trial + "/movie/system-" + name + ".xyz").c_str(), n, pos.data(), t),
          "movie frame");

    psim_observables ob{};
    std::vector<int> coord(n);
    check(psim_gpu_observables(S.ctx, 1.2, &ob, coord.data()), "psim_gpu_observables");
    long long total_coor = 0;
    float total_pot = 0;
    for (int i = 0; i < n; ++i) {
        total_coor += coord[i];
        total_pot += particle_potential(&pos[3 * (size_t)i], &pos0[3 * (size_t)i],
                                        &frc[3 * (size_t)i], &frc0[3 * (size_t)i]);
    }
    const double pot = total_pot / double(n);
    // writeClusters (analysisCluster.cpp:85-130): clustK.xyz files, returns
    // the mean cluster size by integer division
    auto cl = clusters_of(S.ctx, n);
    int total = 0;
    for (size_t k = 0; k < cl.size(); ++k) {
        total += (int)cl[k].size();
        std::string s = std::to_string(cl[k].size()) + "\nCluster: " + std::to_string(k) + "\n";
        for (int i : cl[k])
            s += "H " + fmt_g(pos[3 * (size_t)i]) + " " + fmt_g(pos[3 * (size_t)i + 1]) + " " +
                 fmt_g(pos[3 * (size_t)i + 2]) + "\n";
        std::ofstream f(dir + "/clust" + std::to_string(k) + ".xyz");
        f << s;
    }
    const int avg_size = cl.empty() ? 0 : total / (int)cl.size();
    const double avg_coor = double(total_coor) / double(n);
    std::printf("<Coor>: %g - Total Coor: %lld\n<EAP>: %g\n<N>/Nc: %d\n<R^2>: %g\n"
                "Temperature: %g\n#Clusters: %zu\n",
                avg_coor, total_coor, pot, avg_size, ob.mean_disp2, ob.temperature, cl.size());
    append(S.trial + "/clustGraph.txt", t, avg_size);
    append(S.trial + "/coorGraph.txt", t, avg_coor);
    append(S.trial + "/meanR2Graph.txt", t, ob.mean_disp2);
    append(S.trial + "/meanPot.txt", t, pot);
    append(S.trial + "/trackedMeanR2Graph.txt", t, ob.msd);
    std::fflush(stdout);
}

std::vector<double> forces_now(psim_ctx *ctx, int n)
{
    std::vector<double> f(3 * (size_t)n);
    check(psim_gpu_get_state(ctx, nullptr, nullptr, nullptr, f.data(), nullptr, nullptr,
                             nullptr, nullptr, nullptr),
          "psim_gpu_get_state");
    return f;
}

// system::run (system.cpp:170-211): step, then a snapshot whenever the step
// counter is a multiple of outputFreq, until currentTime reaches endTime.
// The steps between snapshots are issued as one batch.
int run(System &S, double t, long long counter, double end_time, std::vector<double> frc0,
        bool save_state)
{
    const int n = S.n;
    // the loop's own times and counters, accumulated as the reference does
    std::vector<std::pair<long long, double>> outputs;   // (steps done, time)
    long long steps = 0;
    {
        double tt = t;
        long long c = counter;
        while (tt < end_time) {
            ++steps;
            if (c % S.output_freq == 0) outputs.push_back({steps, tt});
            ++c;
            tt += S.dt;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    long long done = 0;
    for (auto &o : outputs) {
        if (o.first - 1 > done) {
            int rc = psim_gpu_step(S.ctx, (int)(o.first - 1 - done), nullptr);
            check(rc, "psim_gpu_step");
            done = o.first - 1;
            frc0 = forces_now(S.ctx, n);   // the force pushed to frc0 by the next step
        }
        check(psim_gpu_step(S.ctx, 1, nullptr), "psim_gpu_step");
        ++done;
        write_snapshot(S, o.second, frc0, save_state);
        frc0 = forces_now(S.ctx, n);
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                              .count();
        std::printf("Average Cycle Time: %g\n", el / (double)done);
    }
    if (steps > done) check(psim_gpu_step(S.ctx, (int)(steps - done), nullptr), "psim_gpu_step");
    check(psim_gpu_sync(S.ctx), "psim_gpu_sync");
    return 0;
}

void usage()
{
    std::fprintf(stderr,
                 "usage: psim_sim [-c settings.cfg] [-r DIR STAMP [--from-recovery]]\n"
                 "                [--kernel-path auto|sparse|dense|brick]\n"
                 "       psim_sim -a DIR STAMP [--coorhist|-CH] [--clusthist|-CLH] "
                 "[--clustcoor|-CLC]\n");
    std::exit(2);
}

// Load DIR/snapshots/time-STAMP/recovery.txt the reference's way.
void load_recovery(System &S, const std::string &dir, const std::string &stamp,
                   std::vector<double> &pos, std::vector<double> &pos0, std::vector<double> &frc,
                   std::vector<double> &frc0)
{
    const size_t m = 3 * (size_t)S.n;
    pos.assign(m, 0.0);
    pos0.assign(m, 0.0);
    frc.assign(m, 0.0);
    frc0.assign(m, 0.0);
    const std::string path = dir + "/snapshots/time-" + stamp + "/recovery.txt";
    std::printf("Reading: %s\n", path.c_str());
    check(psim_read_recovery(path.c_str(), S.n, (double)S.box, pos.data(), pos0.data(),
                             frc.data(), frc0.data(), nullptr, nullptr),
          path.c_str());
}

int analysis_mode(const std::vector<std::string> &args, const std::string &kernel_path)
{
    // runAnalysis (analysis.cpp:28-49): DIR STAMP then at least one test
    if (args.size() < 2) die(1, "No testing parameters defined.");
    if (args.size() < 3) die(1, "No analysis routines defined.");
    System S;
    const std::string dir = args[0], stamp = args[1];
    if (!S.cfg.load(dir + "/settings.cfg")) die(1, "cannot read " + dir + "/settings.cfg");
    S.create(kernel_path);
    S.trial = dir + "/-analysis-" + stamp;
    make_dir(S.trial);
    std::vector<double> pos, pos0, frc, frc0;
    load_recovery(S, dir, stamp, pos, pos0, frc, frc0);
    check(psim_gpu_set_state(S.ctx, pos.data(), pos0.data(), nullptr, frc.data(), nullptr,
                             nullptr, 0.0, 0, 0),
          "psim_gpu_set_state");
    std::vector<int> coord(S.n);
    psim_observables ob{};
    check(psim_gpu_observables(S.ctx, 1.2, &ob, coord.data()), "psim_gpu_observables");
    auto cl = clusters_of(S.ctx, S.n);
    std::printf("Loaded %zu clusters.\n", cl.size());
    // postAnalysis (analysisPost.cpp:26-52) and the histograms
    // (analysisCluster.cpp:132-200)
    for (size_t k = 2; k < args.size(); ++k) {
        const std::string &a = args[k];
        std::map<int, int> h;
        if (a == "--coorhist" || a == "-CH") {
            for (int i = 0; i < S.n; ++i) h[coord[i]]++;
            histogram_file(S.trial + "/coordinationHistogram.txt", h);
        } else if (a == "--clusthist" || a == "-CLH") {
            for (auto &c : cl) h[(int)c.size()]++;
            histogram_file(S.trial + "/clusterSizeHistogram.txt", h);
        } else if (a == "--clustcoor" || a == "-CLC") {
            for (auto &c : cl)
                for (int i : c) h[coord[i]]++;
            histogram_file(S.trial + "/clusterCoorHistogram.txt", h);
        }
    }
    psim_gpu_destroy(S.ctx);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    std::printf("---Particle Simulator 2015 (MI355X)---\n\n");
    std::string settings = "settings.cfg", rewind, stamp, kernel_path = "auto";
    bool analysis = false, from_recovery = false;
    std::vector<std::string> analysis_args;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-r") {
            if (i + 2 >= argc) usage();   // error::throwInputError
            rewind = argv[++i];
            stamp = argv[++i];
        } else if (a == "-a") {
            analysis = true;
            int j = i + 1;
            for (; j < argc && std::string(argv[j]) != "-r"; ++j) analysis_args.push_back(argv[j]);
            i = j - 1;
        } else if (a == "-c" && i + 1 < argc) {
            settings = argv[++i];
        } else if (a == "--kernel-path" && i + 1 < argc) {
            kernel_path = argv[++i];
        } else if (a == "--from-recovery") {
            from_recovery = true;
        } else {
            usage();
        }
    }
    if (analysis) return analysis_mode(analysis_args, kernel_path);

    System S;
    if (!S.cfg.load(settings)) die(1, "cannot read " + settings);
    S.create(kernel_path);
    const size_t m = 3 * (size_t)S.n;
    // frc / frc0 as the particles hold them at the start (initialState.txt);
    // push = the force the first step moves into frc0
    std::vector<double> pos(m), pos0(m), vel(m, 0.0), frc(m, 0.0), frc0(m, 0.0), push(m, 0.0);
    double t = 0.0;
    long long counter = 0;
    bool exact = false;
    if (!rewind.empty()) {
        // RecoverySystem (RecoverySystem.cpp:27-43)
        S.trial = rewind + "/-rewind-" + stamp;
        const std::string state = rewind + "/snapshots/time-" + stamp + "/state.psim";
        exact = !from_recovery && file_exists(state);
        load_recovery(S, rewind, stamp, pos, pos0, frc, frc0);
        if (exact) {
            check(psim_gpu_load(S.ctx, state.c_str()), state.c_str());
            unsigned long long si = 0;
            check(psim_gpu_get_state(S.ctx, pos.data(), pos0.data(), nullptr, frc.data(),
                                     nullptr, nullptr, &t, nullptr, &si),
                  "psim_gpu_get_state");
            counter = (long long)si;
        }
        push = frc;
    } else {
        S.trial = S.cfg.get_string("trialName", "");
        if (S.trial.empty()) die(1, "settings must name trialName");
        // system::initParticles + initCheck + maxwellVelocityInit
        // (systemInit.cpp:31-267), the reference's own draws
        check(psim_init_reference(S.n, S.box, S.radius, (unsigned)S.seed,
                                  S.cfg.get_int("seedSize", 0), S.temp, pos.data(), pos0.data(),
                                  vel.data()),
              "initial configuration");
    }
    make_dir(S.trial);
    make_dir(S.trial + "/snapshots");
    make_dir(S.trial + "/movie");
    if (!exact) {
        check(psim_gpu_set_state(S.ctx, pos.data(), pos0.data(), vel.data(), frc.data(), nullptr,
                                 nullptr, 0.0, 0, 0),
              "psim_gpu_set_state");
        check(psim_gpu_track_start(S.ctx), "psim_gpu_track_start");
    }
    S.write_init(pos, pos0, frc, frc0, settings);
    std::printf("System concentration: %g\nStarting integration.\n", S.conc);
    std::fflush(stdout);
    const int end_time = (int)S.cfg.get_double("endTime", 1000);   // runSim.cpp:154
    run(S, t, counter, (double)end_time, push, true);
    std::printf("\nIntegration complete.\n");
    psim_gpu_destroy(S.ctx);
    return 0;
}
