// psim_io.cpp -- snapshot files: the reference's text formats and a
// full-precision state file for exact resume.
//
//   psim_write_recovery  analysisManager::writeSystem (analysisOutput.cpp:55-74)
//   psim_read_recovery   system::particlesInFile + readParticles
//                        (systemRecovery.cpp:30-81)
//   psim_write_xyz       analysisManager::writeSystemXYZ (analysisOutput.cpp:31-53)
//   psim_write_sys_config system::writeSystemInit's sysConfig.cfg (system.cpp:108-127)
//
// Host code; the full-precision state file (psim_gpu_save / psim_gpu_load)
// is in psim_state.hip.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "psim_gpu.h"

namespace {

// std::ostream's default floating-point format (precision 6, neither fixed
// nor scientific) is printf's %g.
inline void put_g(std::string &s, double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%g", v);
    s += b;
}

struct File {
    std::FILE *f = nullptr;
    explicit File(std::FILE *g) : f(g) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int write_text(const char *path, const std::string &text)
{
    File out(std::fopen(path, "wb"));
    if (!out.f) return PSIM_E_IO;
    if (!text.empty() && std::fwrite(text.data(), 1, text.size(), out.f) != text.size())
        return PSIM_E_IO;
    if (std::fclose(out.f) != 0) {
        out.f = nullptr;
        return PSIM_E_IO;
    }
    out.f = nullptr;
    return 0;
}

// particle::setX/Y/Z (particle.cpp:61-92): single wrap (util::safeMod), the
// previous coordinate re-imaged next to the new one (util::safeMod0,
// utilities.cpp:26-50), 7703 if still outside [0, box).
inline int set_coord(double &x, double &x0, double val, double box)
{
    const double old = x;
    double v = val;
    if (v < 0) v = v + box;
    else if (v >= box) v = v - box;
    x = v;
    const double dx = v - old;
    x0 = (std::fabs(dx) > (box / 2)) ? ((dx < 0) ? old - box : old + box) : old;
    return (x < 0.0 || x >= box) ? PSIM_ERR_BOUNDS : 0;
}

}  // namespace

extern "C" {

int psim_write_recovery(const char *path, int n, const double *pos, const double *pos0,
                        const double *frc, const double *frc0, double mass, double radius)
{
    if (!path || n < 0 || (n > 0 && (!pos || !pos0))) return PSIM_E_ARG;
    std::string s;
    s.reserve((size_t)n * 120);
    for (int i = 0; i < n; ++i) {
        const size_t b = 3 * (size_t)i;
        const double *cols[4] = {pos, pos0, frc, frc0};
        for (const double *a : cols)
            for (int d = 0; d < 3; ++d) {
                put_g(s, a ? a[b + d] : 0.0);
                s += ' ';
            }
        put_g(s, mass);
        s += ' ';
        put_g(s, radius);
        if (i < n - 1) s += '\n';
    }
    return write_text(path, s);
}

int psim_read_recovery(const char *path, int n, double box, double *pos, double *pos0,
                       double *frc, double *frc0, double *mass, double *radius)
{
    if (!path || n < 0 || !pos || !pos0 || !(box > 0)) return PSIM_E_ARG;
    File in(std::fopen(path, "rb"));
    if (!in.f) return PSIM_E_IO;
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, in.f)) > 0) text.append(buf, got);
    // std::getline's lines: a final newline does not start another line
    std::vector<std::string> lines;
    size_t a = 0;
    while (a < text.size()) {
        size_t e = text.find('\n', a);
        if (e == std::string::npos) e = text.size();
        lines.emplace_back(text, a, e - a);
        a = e + 1;
    }
    if ((int)lines.size() != n) return PSIM_E_IO;
    for (int i = 0; i < n; ++i) {
        double v[14];
        const char *p = lines[i].c_str();
        for (int k = 0; k < 14; ++k) {
            char *end = nullptr;
            errno = 0;
            v[k] = std::strtod(p, &end);
            if (end == p) return PSIM_E_IO;
            p = end;
        }
        const size_t b = 3 * (size_t)i;
        // a fresh particle (particle.cpp:31-52) at the origin; setPos(pos0),
        // updateForce(frc0), nextIter, setPos(pos), updateForce(frc)
        for (int d = 0; d < 3; ++d) {
            double x = 0.0, x0 = 0.0;
            int rc = set_coord(x, x0, v[3 + d], box);
            if (!rc) rc = set_coord(x, x0, v[d], box);
            if (rc) return rc;
            pos[b + d] = x;
            pos0[b + d] = x0;
            if (frc) frc[b + d] = v[6 + d];
            if (frc0) frc0[b + d] = v[9 + d];
        }
        // mass and radius are read into floats (systemRecovery.cpp:58)
        if (mass) mass[i] = (double)(float)v[12];
        if (radius) radius[i] = (double)(float)v[13];
    }
    return 0;
}

int psim_write_xyz(const char *path, int n, const double *pos, double current_time)
{
    if (!path || n < 0 || (n > 0 && !pos)) return PSIM_E_ARG;
    std::string s = std::to_string(n) + "\n";
    s += "Current Time: " + std::to_string((int)std::round(current_time)) + "\n";
    for (int i = 0; i < n; ++i) {
        s += "H ";
        for (int d = 0; d < 3; ++d) {
            put_g(s, pos[3 * (size_t)i + d]);
            s += (d < 2) ? ' ' : '\n';
        }
    }
    return write_text(path, s);
}

int psim_write_sys_config(const char *path, const char *trial_name, int n,
                          double concentration, int box, int cell_size, int cell_scale,
                          double temp, double dt, int output_freq, double cycle_hour, int seed)
{
    if (!path || !trial_name) return PSIM_E_ARG;
    std::string s = "trialName = " + std::string(trial_name) + "\n";
    s += "nParticles = " + std::to_string(n) + "\n";
    s += "Concentration = ";
    put_g(s, concentration);
    s += "\nboxSize = " + std::to_string(box) + "\n";
    s += "cellSize = " + std::to_string(cell_size) + "\n";
    s += "cellScale = " + std::to_string(cell_scale) + "\n";
    s += "temp = ";
    put_g(s, temp);
    s += "\ndTime = ";
    put_g(s, dt);
    s += "\noutputFreq = " + std::to_string(output_freq) + "\n";
    s += "cycleHour = ";
    put_g(s, cycle_hour);
    s += "\nseed = " + std::to_string(seed);
    return write_text(path, s);
}

}  // extern "C"
