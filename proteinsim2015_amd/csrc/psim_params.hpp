// psim_params.hpp -- host-side derivation of the device constants from
// psim_params: box/cell geometry, pair-force constants
// (LJPotential.cpp:34-55, AOPotential.cpp:50-55) and the G&B coefficients
// (brownianIntegrator.cpp:53-81, 111-175), all in host FP64 with the
// reference's formulas and operation order.
#pragma once

#include <cmath>

#include "psim_device.hpp"
#include "psim_gpu.h"

namespace psim {

// The G&B coefficients of brownianIntegrator.cpp:53-81, 111-175 (host FP64,
// glibc libm, the reference's formulas and operation order): setupHigh,
// the series of setupLow when gamma < 0.05, setupZero when gamma == 0, then
// the noise widths and their correlation.
struct GBCoeffs {
    double coeff0, coeff1, coeff2, coeff3;
    double sig1, sig2, corr, dev;
    double gn, goy2, goy3, hn;
};

inline GBCoeffs gb_coefficients(double kT, double gamma, double dt)
{
    double y = gamma * dt;
    double coeff0, coeff1, coeff2, coeff3, sig1, sig2, gn, goy2, goy3, hn;
    {
        double ty = 2.0 * y;
        coeff0 = std::exp(-y);
        double aa1 = 1.0 - coeff0;
        double aa2 = 0.5 * y * (1.0 + coeff0) - aa1;
        double aa3 = y - aa1;
        coeff1 = aa1 / y;
        coeff2 = aa2 / (y * y);
        coeff3 = aa3 / (y * y);
        sig1 = 2.0 * y - 3.0 + 4.0 * std::exp(-y) - std::exp(-ty);
        sig2 = -2.0 * y - 3.0 + 4.0 * std::exp(y) - std::exp(ty);
        gn = std::exp(y) - ty - std::exp(-y);
        goy2 = gn / (y * y);
        goy3 = gn / (y * y * y);
        hn = y / (std::exp(y) - std::exp(-y));
    }
    if (gamma < 0.05) {
        double y1 = y, y2 = y1 * y1, y3 = y2 * y1, y4 = y3 * y1, y5 = y4 * y1;
        double y6 = y5 * y1, y7 = y6 * y1, y8 = y7 * y1, y9 = y8 * y1;
        coeff1 = 1.0 - 0.5 * y1 + (1.0 / 6.0) * y2 - (1.0 / 24.0) * y3 + (1.0 / 120.0) * y4;
        coeff2 = (1.0 / 12.0) * y1 - (1.0 / 24.0) * y2 + (1.0 / 80.0) * y3 - (1.0 / 360.0) * y4;
        coeff3 = 0.5 - (1.0 / 6.0) * y1 + (1.0 / 24.0) * y2 - (1.0 / 120.0) * y3;
        sig1 = +(2.0 / 3.0) * y3 - 0.5 * y4 + (7.0 / 30.0) * y5 - (1.0 / 12.0) * y6
             + (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 + (127.0 / 90720.0) * y9;
        sig2 = -(2.0 / 3.0) * y3 - 0.5 * y4 - (7.0 / 30.) * y5 - (1.0 / 12.0) * y6
             - (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 - (127.0 / 90720.0) * y9;
        goy2 = (1.0 / 3.0) * y1 + (1.0 / 60.0) * y3;
        goy3 = 1.0 / 3.0 + (1.0 / 60.0) * y2;
        hn = 0.5 - (1.00 / 12.0) * y2 + (7.0 / 720.0) * y4;
        gn = (1.0 / 3.0) * y3 + (1.0 / 60.0) * y5;
    }
    if (gamma == 0) {
        coeff0 = 1.0;
        coeff1 = 1.0;
        coeff2 = 0.0;
        coeff3 = 0.5;
    }
    double gamma2 = gamma * gamma;
    sig1 = std::sqrt(+kT * sig1 / gamma2);
    sig2 = std::sqrt(-kT * sig2 / gamma2);
    double corr = (kT / (gamma2)) * (gn / (sig1 * sig2));
    double dev = std::sqrt(1.0 - (corr * corr));
    return GBCoeffs{coeff0, coeff1, coeff2, coeff3, sig1, sig2, corr, dev, gn, goy2, goy3, hn};
}

inline void make_constants(const psim_params &P, Geo &g, Pot &p, GB &b)
{
    g.n = P.n_particles;
    g.S = P.cells_per_dim;
    g.S2 = g.S * g.S;
    g.ncell = g.S2 * g.S;
    g.L = P.box;
    g.half_pbc = P.box / 2.0;
    if (P.half_uv > 0)
        g.half_uv = P.half_uv;
    else if (std::floor(P.box) == P.box)
        g.half_uv = (double)((long long)P.box / 2);   // int L / 2 (utilities.cpp:124)
    else
        g.half_uv = P.box / 2.0;
    g.cell = P.cell_size;
    g.exact_mi = (g.S < 4) ? 1 : 0;
    g.c0 = 0;

    p.mask = P.force_mask;
    p.lj_num = P.lj_num;
    const double rc = P.cutoff;
    p.rc2 = rc * rc;
    // guard band: the fast image distance differs from pbcDist's by a few
    // ulps of L on each component
    double band = 1e-12 * std::fmax(1.0, P.box / rc);
    p.rc2_hi = p.rc2 * (1.0 + band);
    p.rc2_lo = p.rc2 * (1.0 - band);
    p.ov_factor = 0.8;
    p.ov_band = band;
    p.lj_neg_kt_eps = -P.kT * P.lj_well_depth;
    p.lj_four_n = 4.0 * P.lj_num;
    p.debye_len = P.debye_length;
    p.debye_inv = 1.0 / P.debye_length;
    p.yuk_a = P.yukawa_strength;
    p.lj_att = p.lj_neg_kt_eps * p.lj_four_n;
    p.lj_rep = p.lj_neg_kt_eps * p.yuk_a;
    p.lj_rep_lam = p.lj_rep * P.debye_length;
    // AOPotential.cpp:50-55
    double q = (rc / (rc - 1.0));
    double a1 = -P.kT * P.ao_well_depth * q * q * q;
    double a2 = -3.0 / (2.0 * rc);
    double a3 = 1.0 / (2.0 * rc * rc * rc);
    p.ao_c1 = -a1 * a2;
    p.ao_c2 = -3.0 * a1 * a3;

    const double dt = P.dt;
    const GBCoeffs k = gb_coefficients(P.kT, P.gamma, P.dt);
    const double coeff0 = k.coeff0, coeff1 = k.coeff1, coeff2 = k.coeff2, coeff3 = k.coeff3;
    const double sig1 = k.sig1, sig2 = k.sig2, corr = k.corr, dev = k.dev;
    const double goy2 = k.goy2, goy3 = k.goy3, hn = k.hn;
    b.coeff0 = coeff0;
    b.coeff1 = coeff1;
    b.coeff3 = coeff3;
    b.c0 = 1.0 + coeff0;
    double m = 1.0 / P.mass;
    double dt2 = dt * dt;
    b.c1 = m * dt2 * coeff1;
    b.c2 = m * dt2 * coeff2;
    b.sig1 = sig1;
    b.sig2 = sig2;
    b.corr = corr;
    b.dev = dev;
    b.m = m;
    b.dt = dt;
    double dt3 = dt * dt2;
    b.g2 = m * dt2 * goy2;
    b.g3 = m * dt3 * goy3;
    b.g4 = hn * (1.0 / dt);
}

}  // namespace psim
