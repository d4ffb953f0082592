/*
 * psim_oracle.c -- CPU restatement of the ProteinSim2015 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built with
 * -ffp-contract=off so every expression rounds exactly as the reference's
 * x86-64 g++ build does (no FMA there).  Each function follows the reference
 * file:line named in its comment, keeping the reference's operation order
 * where rounding depends on it.
 */
#include "psim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Geometry: system.cpp:50-64                                          */
/* ------------------------------------------------------------------ */
void po_geometry(int n, double conc, int scale, double radius,
                 int *box, int *cell_size, double *concentration)
{
    /* vP = N * (4/3) * pi * r^3 with pi = atan(1)*4, evaluated left to right */
    double vP = n * (4.0 / 3.0) * atan(1.0) * 4.0 * radius * radius * radius;
    int b = (int)(cbrt(vP / conc));
    int c = b / scale;           /* integer division, system.cpp:60 */
    b = c * scale;               /* system.cpp:61 */
    *box = b;
    *cell_size = c;
    *concentration = vP / pow((double)b, 3.0);
}

/* ------------------------------------------------------------------ */
/* Cell hash: systemHandling.cpp:33-48                                   */
/* ------------------------------------------------------------------ */
static inline int po_cell_1d(double x, double cell_size, int S)
{
    int c = (int)floor(x / cell_size);
    /* divergence from the reference: clamp (reference indexes out of range
     * when x/cellSize rounds up to S for x just below L) */
    if (c < 0) c = 0;
    if (c > S - 1) c = S - 1;
    return c;
}

void po_hash(int n, const double *pos, double cell_size, int S, int *key)
{
    int S2 = S * S;
    for (int i = 0; i < n; i++) {
        int cx = po_cell_1d(pos[3 * i + 0], cell_size, S);
        int cy = po_cell_1d(pos[3 * i + 1], cell_size, S);
        int cz = po_cell_1d(pos[3 * i + 2], cell_size, S);
        key[i] = cx + (S * cy) + (S2 * cz);
    }
}

/* Stable counting sort + start/end table (systemHandling.cpp:50-75,
 * system.h:166).  Empty cells hold -1 in both start and end. */
void po_build_cells(int n, const int *key, int ncell,
                    int *perm, int *cstart, int *cend)
{
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (int i = 0; i < n; i++) cnt[key[i] + 1]++;
    for (int c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    for (int c = 0; c < ncell; c++) {
        if (cnt[c + 1] > cnt[c]) {
            cstart[c] = cnt[c];
            cend[c] = cnt[c + 1];
        } else {
            cstart[c] = -1;
            cend[c] = -1;
        }
    }
    for (int i = 0; i < n; i++) perm[cnt[key[i]]++] = i;
    free(cnt);
}

/* systemHandling.cpp:77-84 */
void po_reorder(int n, const int *perm, const double *pos, double radius,
                double *sorted4)
{
    for (int i = 0; i < n; i++) {
        int o = perm[i];
        sorted4[4 * i + 0] = pos[3 * o + 0];
        sorted4[4 * i + 1] = pos[3 * o + 1];
        sorted4[4 * i + 2] = pos[3 * o + 2];
        sorted4[4 * i + 3] = radius;
    }
}

/* ------------------------------------------------------------------ */
/* Pair geometry: utilities.cpp:79-104 (pbcDist), :114-142 (unitVectorAdv),
 * :144-154 (powBinaryDecomp)                                            */
/* ------------------------------------------------------------------ */
static inline double po_pbc_dist(double X, double Y, double Z,
                                 double X1, double Y1, double Z1, double L)
{
    double dx = fabs(X - X1);
    double dy = fabs(Y - Y1);
    double dz = fabs(Z - Z1);
    double half = L / 2.0;
    if (dx > half) dx -= L;
    if (dy > half) dy -= L;
    if (dz > half) dz -= L;
    return (dx * dx) + (dy * dy) + (dz * dz);
}

static inline void po_unit_vector(double X, double Y, double Z,
                                  double X1, double Y1, double Z1,
                                  double acc[3], double r, double L,
                                  double half_uv)
{
    double dx = X1 - X, dy = Y1 - Y, dz = Z1 - Z;
    double one_over = 1.0 / r;
    if (fabs(dx) > half_uv) { if (dx < 0) dx += L; else dx -= L; }
    if (fabs(dy) > half_uv) { if (dy < 0) dy += L; else dy -= L; }
    if (fabs(dz) > half_uv) { if (dz < 0) dz += L; else dz -= L; }
    acc[0] = dx * one_over;
    acc[1] = dy * one_over;
    acc[2] = dz * one_over;
}

static inline double po_pow_binary(double base, int e)
{
    double answer = 1;
    while (e) {
        if (e & 1) answer *= base;
        e >>= 1;
        base *= base;
    }
    return answer;
}

/* ------------------------------------------------------------------ */
/* Pair force magnitudes (positive = attractive)                         */
/* ------------------------------------------------------------------ */
typedef struct {
    double rc2;
    double lj_kt_eps;      /* kT * wellDepth, negated when applied */
    double lj_four_n;      /* 4.0 * ljNum */
    double debye_inv;
    double ao_c1, ao_c2;   /* AOPotential.cpp:50-55 */
} po_pair_consts;

/* LJPotential.cpp:97-109 */
static inline double po_f_lj(const po_force_params *fp, const po_pair_consts *k,
                             double r, double size)
{
    double rInv = (1.0 / r);
    double yukExp = exp(-1.0 * (r * k->debye_inv));
    double LJ = po_pow_binary((size / r), fp->lj_num);
    double attract = ((2.0 * LJ) - 1.0);
    attract *= (4.0 * fp->lj_num * rInv * LJ);
    double repel = yukExp;
    repel *= (rInv * rInv * (fp->debye_length + r) * fp->yukawa_strength);
    return -fp->kT * fp->lj_well_depth * (attract + repel);
}

/* AOPotential.cpp:95-102 */
static inline double po_f_ao(const po_pair_consts *k, double r, double r2)
{
    double rInv = 1.0 / r;
    double r_36 = pow(rInv, 36);
    double r_38 = r_36 / r2;
    double fNet = 36.0 * r_38 + k->ao_c1 * rInv + k->ao_c2 * r;
    return -fNet;
}

/* Calibration.cpp:75-81 */
static inline double po_f_cal(double r)
{
    double rInv = 1.0 / r;
    double r_37 = 36.0 * pow(rInv, 37);
    return -r_37;
}

/* iterCells for one module: LJPotential.cpp:64-133 (AO :61-130, Cal :41-102) */
static int po_iter_cell(int module, const po_force_params *fp,
                        const po_pair_consts *k, int index, int hash,
                        const double *sp, const int *cstart, const int *cend,
                        double L, double half_uv, double cell_force[3],
                        int *err_ij)
{
    int start = cstart[hash];
    cell_force[0] = cell_force[1] = cell_force[2] = 0.0;
    if (start == -1) return 0;
    int end = cend[hash];
    for (int i = start; i < end; i++) {
        if (i == index) continue;
        const double *a = sp + 4 * index, *b = sp + 4 * i;
        double rSquared = po_pbc_dist(a[0], a[1], a[2], b[0], b[1], b[2], L);
        if (!(rSquared < k->rc2)) continue;
        double r = sqrt(rSquared);
        double size = (a[3] + b[3]);
        if (r < (0.8 * size)) {
            err_ij[0] = index;
            err_ij[1] = i;
            return 7704;
        }
        double fNet;
        if (module == PO_LJ) {
            fNet = po_f_lj(fp, k, r, size);
        } else if (module == PO_AO) {
            fNet = po_f_ao(k, r, rSquared);
            if (isnan(fNet)) {
                err_ij[0] = index;
                err_ij[1] = i;
                return 7705;
            }
        } else {
            fNet = po_f_cal(r);
        }
        double u[3];
        po_unit_vector(a[0], a[1], a[2], b[0], b[1], b[2], u, r, L, half_uv);
        cell_force[0] += fNet * u[0];
        cell_force[1] += fNet * u[1];
        cell_force[2] += fNet * u[2];
    }
    return 0;
}

static void po_pair_consts_init(const po_force_params *fp, po_pair_consts *k)
{
    double c = fp->cutoff;
    k->rc2 = c * c;
    k->lj_kt_eps = fp->kT * fp->lj_well_depth;
    k->lj_four_n = 4.0 * fp->lj_num;
    k->debye_inv = 1.0 / fp->debye_length;
    /* AOPotential.cpp:50-55 */
    double q = (c / (c - 1.0));
    double a1 = -fp->kT * fp->ao_well_depth * q * q * q;
    double a2 = -3.0 / (2.0 * c);
    double a3 = 1.0 / (2.0 * c * c * c);
    k->ao_c1 = -a1 * a2;
    k->ao_c2 = -3.0 * a1 * a3;
}

/* getAcceleration per module: LJPotential.cpp:139-182 */
int po_forces(const po_force_params *fp, int n, const double *sorted4,
              const int *perm, const int *cstart, const int *cend,
              int S, double cell_size, double box, double half_uv,
              double *force3, int *err_ij, int nthreads)
{
    po_pair_consts k;
    po_pair_consts_init(fp, &k);
    int status = 0;
    int eij[2] = {-1, -1};
    for (int i = 0; i < 3 * n; i++) force3[i] = 0.0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    (void)nthreads;
#endif
    static const int modules[3] = {PO_CAL, PO_LJ, PO_AO};
    for (int m = 0; m < 3; m++) {
        int module = modules[m];
        if (!(fp->mask & module)) continue;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int index = 0; index < n; index++) {
            const double *a = sorted4 + 4 * index;
            int cx = po_cell_1d(a[0], cell_size, S);
            int cy = po_cell_1d(a[1], cell_size, S);
            int cz = po_cell_1d(a[2], cell_size, S);
            int S2 = S * S;
            double net[3] = {0.0, 0.0, 0.0};
            for (int x = -1; x <= 1; x++) {
                for (int y = -1; y <= 1; y++) {
                    for (int z = -1; z <= 1; z++) {
                        int rx = (cx + x) % S, ry = (cy + y) % S, rz = (cz + z) % S;
                        rx = (rx < 0) ? rx + S : rx;
                        ry = (ry < 0) ? ry + S : ry;
                        rz = (rz < 0) ? rz + S : rz;
                        int hash = rx + (S * ry) + (S2 * rz);
                        double cf[3];
                        int local_err[2];
                        int st = po_iter_cell(module, fp, &k, index, hash,
                                              sorted4, cstart, cend, box,
                                              half_uv, cf, local_err);
                        if (st) {
#pragma omp critical(po_err)
                            {
                                if (!status) {
                                    status = st;
                                    eij[0] = local_err[0];
                                    eij[1] = local_err[1];
                                }
                            }
                        }
                        net[0] += cf[0];
                        net[1] += cf[1];
                        net[2] += cf[2];
                    }
                }
            }
            int o = perm[index];
            /* a composite is the sum of the separate module results */
            force3[3 * o + 0] += net[0];
            force3[3 * o + 1] += net[1];
            force3[3 * o + 2] += net[2];
        }
    }
    if (err_ij) { err_ij[0] = eij[0]; err_ij[1] = eij[1]; }
    return status;
}

/* ------------------------------------------------------------------ */
/* Neighbour sets of the force loop: getAcceleration's 27-cell stencil   */
/* (LJPotential.cpp:139-177) and iterCells' test (:64-95: skip i ==      */
/* index, util::pbcDist, strict rSquared < cutOff^2).                    */
/* ------------------------------------------------------------------ */
static inline unsigned long long po_nbr_hash(unsigned o)
{
    /* SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014) */
    unsigned long long z = (unsigned long long)o + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void po_pair_stats(const po_force_params *fp, int n, const double *sorted4,
                   const int *perm, const int *cstart, const int *cend,
                   int S, double cell_size, double box, int *count,
                   unsigned long long *hash, double *abs_sum, int id0, int nthreads)
{
    po_pair_consts k;
    po_pair_consts_init(fp, &k);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int index = 0; index < n; index++) {
        const double *a = sorted4 + 4 * index;
        int cx = po_cell_1d(a[0], cell_size, S);
        int cy = po_cell_1d(a[1], cell_size, S);
        int cz = po_cell_1d(a[2], cell_size, S);
        int S2 = S * S;
        int cnt = 0;
        unsigned long long hs = 0;
        double fsum = 0.0;
        for (int x = -1; x <= 1; x++) {
            for (int y = -1; y <= 1; y++) {
                for (int z = -1; z <= 1; z++) {
                    int rx = (cx + x) % S, ry = (cy + y) % S, rz = (cz + z) % S;
                    rx = (rx < 0) ? rx + S : rx;
                    ry = (ry < 0) ? ry + S : ry;
                    rz = (rz < 0) ? rz + S : rz;
                    int hsh = rx + (S * ry) + (S2 * rz);
                    int start = cstart[hsh];
                    if (start == -1) continue;
                    for (int i = start; i < cend[hsh]; i++) {
                        if (i == index) continue;
                        const double *b = sorted4 + 4 * i;
                        double r2 = po_pbc_dist(a[0], a[1], a[2], b[0], b[1], b[2], box);
                        if (!(r2 < k.rc2)) continue;
                        cnt++;
                        hs += po_nbr_hash((unsigned)(perm[i] + id0));
                        if (abs_sum) {
                            double r = sqrt(r2), size = a[3] + b[3], f = 0.0;
                            if (fp->mask & PO_CAL) f += po_f_cal(r);
                            if (fp->mask & PO_LJ) f += po_f_lj(fp, &k, r, size);
                            if (fp->mask & PO_AO) f += po_f_ao(&k, r, r2);
                            fsum += fabs(f);
                        }
                    }
                }
            }
        }
        int o = perm[index];
        count[o] = cnt;
        hash[o] = hs;
        if (abs_sum) abs_sum[o] = fsum;
    }
}

/* ------------------------------------------------------------------ */
/* G&B integrator coefficients: brownianIntegrator.cpp:53-81, 111-175    */
/* ------------------------------------------------------------------ */
void po_gb_coeffs_init(double kT, double mass, double gamma, double dt,
                       po_gb_coeffs *c)
{
    memset(c, 0, sizeof(*c));
    c->kT = kT;
    c->mass = mass;
    c->gamma = gamma;
    c->dt = dt;
    c->dt_inv = 1.0 / dt;
    double y = gamma * dt;
    c->y = y;
    double sig1, sig2;
    /* setupHigh :111-133 */
    {
        double ty = 2.0 * y;
        c->coeff0 = exp(-y);
        double aa1 = 1.0 - c->coeff0;
        double aa2 = 0.5 * y * (1.0 + c->coeff0) - aa1;
        double aa3 = y - aa1;
        c->coeff1 = aa1 / y;
        c->coeff2 = aa2 / (y * y);
        c->coeff3 = aa3 / (y * y);
        sig1 = 2.0 * y - 3.0 + 4.0 * exp(-y) - exp(-ty);
        sig2 = -2.0 * y - 3.0 + 4.0 * exp(y) - exp(ty);
        c->gn = exp(y) - ty - exp(-y);
        c->goy2 = c->gn / (y * y);
        c->goy3 = c->gn / (y * y * y);
        c->hn = y / (exp(y) - exp(-y));
    }
    /* setupLow :135-166 -- selected by gamma (not y) < 0.05, :69 */
    if (gamma < 0.05) {
        double y1 = y, y2 = y1 * y1, y3 = y2 * y1, y4 = y3 * y1, y5 = y4 * y1;
        double y6 = y5 * y1, y7 = y6 * y1, y8 = y7 * y1, y9 = y8 * y1;
        c->coeff1 = 1.0 - 0.5 * y1 + (1.0 / 6.0) * y2 - (1.0 / 24.0) * y3
                  + (1.0 / 120.0) * y4;
        c->coeff2 = (1.0 / 12.0) * y1 - (1.0 / 24.0) * y2 + (1.0 / 80.0) * y3
                  - (1.0 / 360.0) * y4;
        c->coeff3 = 0.5 - (1.0 / 6.0) * y1 + (1.0 / 24.0) * y2 - (1.0 / 120.0) * y3;
        sig1 = +(2.0 / 3.0) * y3 - 0.5 * y4 + (7.0 / 30.0) * y5 - (1.0 / 12.0) * y6
             + (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 + (127.0 / 90720.0) * y9;
        sig2 = -(2.0 / 3.0) * y3 - 0.5 * y4 - (7.0 / 30.) * y5 - (1.0 / 12.0) * y6
             - (31.0 / 1260.0) * y7 - (1.0 / 160.0) * y8 - (127.0 / 90720.0) * y9;
        c->goy2 = (1.0 / 3.0) * y1 + (1.0 / 60.0) * y3;
        c->goy3 = 1.0 / 3.0 + (1.0 / 60.0) * y2;
        c->hn = 0.5 - (1.00 / 12.0) * y2 + (7.0 / 720.0) * y4;
        c->gn = (1.0 / 3.0) * y3 + (1.0 / 60.0) * y5;
    }
    /* setupZero :168-175 */
    if (gamma == 0) {
        c->coeff0 = 1.0;
        c->coeff1 = 1.0;
        c->coeff2 = 0.0;
        c->coeff3 = 0.5;
    }
    double gamma2 = gamma * gamma;
    c->sig1 = sqrt(+kT * sig1 / gamma2);
    c->sig2 = sqrt(-kT * sig2 / gamma2);
    c->corr = (kT / (gamma2)) * (c->gn / (c->sig1 * c->sig2));
    c->dev = sqrt(1.0 - (c->corr * c->corr));
}

/* util::safeMod (utilities.cpp:26-39) */
static inline double po_safe_mod(double val, double base)
{
    if (val < 0) return val + base;
    if (val >= base) return val - base;
    return val;
}

/* util::safeMod0 (utilities.cpp:41-50) */
static inline double po_safe_mod0(double val0, double val, double base)
{
    double dx = val - val0;
    return (fabs(dx) > (base / 2)) ? ((dx < 0) ? val0 - base : val0 + base) : val0;
}

/* particle::setX (particle.cpp:61-70); returns nonzero when out of box */
static inline int po_set_coord(double *pos, double *pos0, double val, double L)
{
    double old = *pos;
    *pos = po_safe_mod(val, L);
    *pos0 = po_safe_mod0(old, *pos, L);
    return (*pos < 0.0) || (*pos >= L);
}

/* brownianIntegrator.cpp:191-223 */
int po_gb_first_step(const po_gb_coeffs *c, int n, double box,
                     double *pos, double *pos0, const double *vel,
                     const double *frc, double *memx, double *memcorr,
                     const double *noise_b, int *err_i)
{
    int status = 0;
    double m = 1.0 / c->mass;
    double dt = c->dt;
    for (int i = 0; i < n; i++) {
        int bad = 0;
        for (int d = 0; d < 3; d++) {
            memcorr[3 * i + d] = 0.0;
            memx[3 * i + d] = noise_b[3 * i + d];
            double xn = pos[3 * i + d] + (vel[3 * i + d] * c->coeff1 * dt)
                      + (frc[3 * i + d] * c->coeff3 * dt * dt * m)
                      + (c->sig1 * memx[3 * i + d]);
            bad |= po_set_coord(&pos[3 * i + d], &pos0[3 * i + d], xn, box);
        }
        if (bad && !status) { status = 7703; if (err_i) *err_i = i; }
    }
    return status;
}

/* brownianIntegrator.cpp:225-296 and velocityStep :298-356 */
int po_gb_normal_step(const po_gb_coeffs *c, int n, double box,
                      double *pos, double *pos0, double *vel,
                      const double *frc, const double *frc0,
                      double *memx, double *memcorr,
                      const double *noise_a, const double *noise_b,
                      int do_vel, int *err_i)
{
    int status = 0;
    double dt = c->dt;
    double dt2 = dt * dt;
    double c0 = 1.0 + c->coeff0;
    double m = 1.0 / c->mass;
    double c1 = m * dt2 * c->coeff1;
    double c2 = m * dt2 * c->coeff2;
    /* velocityStep constants :325-330 */
    double dt3 = dt * dt2;
    double g2 = m * dt2 * c->goy2;
    double g3 = m * dt3 * c->goy3;
    double g4 = c->hn * c->dt_inv;
    for (int i = 0; i < n; i++) {
        double xn[3];
        for (int d = 0; d < 3; d++) {
            int k = 3 * i + d;
            memcorr[k] = c->sig2 * ((c->corr * memx[k]) + (c->dev * noise_a[k]));
        }
        for (int d = 0; d < 3; d++) memx[3 * i + d] = noise_b[3 * i + d];
        for (int d = 0; d < 3; d++) {
            int k = 3 * i + d;
            double p = (c0 * pos[k]);
            p -= (c->coeff0 * pos0[k]);
            p += (c1 * frc[k]);
            p += (c2 * (frc[k] - frc0[k]));
            p += (c->sig1 * memx[k]) + (c->coeff0 * memcorr[k]);
            xn[d] = p;
        }
        if (do_vel) {
            double v[3];
            for (int d = 0; d < 3; d++) {
                int k = 3 * i + d;
                double dx0 = pos[k] - pos0[k];
                double xN = po_safe_mod(xn[d], box);
                double x0 = po_safe_mod0(pos[k], xN, box);
                double dx = xN - x0;
                double vv = dx + dx0;
                vv += (g2 * frc[k]);
                vv -= (g3 * (frc[k] - frc0[k]));
                vv += (memcorr[k] - c->sig1 * memx[k]);
                vv *= g4;
                v[d] = vv;
            }
            for (int d = 0; d < 3; d++) vel[3 * i + d] = v[d];
        }
        int bad = 0;
        for (int d = 0; d < 3; d++)
            bad |= po_set_coord(&pos[3 * i + d], &pos0[3 * i + d], xn[d], box);
        if (bad && !status) { status = 7703; if (err_i) *err_i = i; }
    }
    return status;
}

/* ------------------------------------------------------------------ */
/* fastRandom replay: fastRandom.cpp:28-124                              */
/* ------------------------------------------------------------------ */
typedef struct {
    int istack[58];
    int jrand;
    int init;
    int avail;
    float stored;       /* float storage, fastRandom.h:16 */
} po_fast_random;

/* fastRandom.cpp:40-89 */
static double po_psdrand(po_fast_random *g, int iseed)
{
    const int ndim = 55, m10 = 1000000000, is = 21, ir = 30;
    const double base = 1.0E9;
    int i, j, k, inx;
    if ((!g->init) || (iseed < 0)) {
        iseed = abs(iseed);
        g->istack[ndim] = iseed;
        j = iseed;
        k = 1;
        for (i = 1; i <= (ndim - 1); ++i) {
            inx = i * is - (int)((double)(i * is) / (double)(ndim)) * ndim;
            g->istack[inx] = k;
            k = j - k;
            if (k < 0) k += m10;
            j = g->istack[inx];
        }
        for (j = 1; j <= 3; ++j) {
            for (i = 1; i <= ndim; ++i) {
                inx = i + ir - (int)((double)(i + ir) / (double)(ndim)) * ndim;
                g->istack[i] -= g->istack[inx + 1];
                if (g->istack[i] < 0) g->istack[i] += m10;
            }
        }
        g->jrand = 0;
        g->init = 1;
    }
    g->jrand += 1;
    if (g->jrand > ndim) {
        for (i = 1; i <= ndim; ++i) {
            inx = i + ir - ((int)((double)(i + ir) / (double)(ndim))) * ndim;
            g->istack[i] -= g->istack[inx + 1];
            if (g->istack[i] < 0) g->istack[i] += m10;
        }
        g->jrand = 1;
    }
    return ((double)g->istack[g->jrand]) / base;
}

/* fastRandom.cpp:92-124 (mu = 0, sigma = 1) */
static double po_g250(po_fast_random *g, int seed)
{
    if (!g->avail) {
        double polar, rsquared, var1, var2;
        int tries = 0;
        do {
            var1 = 2.0 * po_psdrand(g, seed) - 1.0;
            var2 = 2.0 * po_psdrand(g, seed) - 1.0;
            rsquared = var1 * var1 + var2 * var2;
            /* a negative seed re-seeds on every call, so a rejected pair
             * repeats for ever in the reference; stop instead */
            if (seed < 0 && ++tries > 4) return NAN;
        } while (rsquared >= 1.0 || rsquared == 0.0);
        polar = sqrt(-2.0 * log(rsquared) / rsquared);
        g->stored = (float)(var1 * polar);
        g->avail = 1;
        return var2 * polar * 1.0 + 0.0;
    }
    g->avail = 0;
    return (double)g->stored * 1.0 + 0.0;
}

int po_fastrandom_streams(int seed, int n, int ndraws, double *out)
{
    int status = 0;
    for (int i = 0; i < n; i++) {
        po_fast_random g;
        memset(&g, 0, sizeof(g));
        /* int threadSeed = seed*(i+1): two's-complement wrap as executed */
        int thread_seed = (int)((unsigned int)seed * (unsigned int)(i + 1));
        if (thread_seed < 0) status = -1;
        for (int k = 0; k < ndraws; k++) out[(size_t)i * ndraws + k] = po_g250(&g, thread_seed);
    }
    return status;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) -> normals          */
/* ------------------------------------------------------------------ */
void po_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2],
                      unsigned int out[4])
{
    unsigned int c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    unsigned int k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        unsigned int hi0 = (unsigned int)(p0 >> 32), lo0 = (unsigned int)p0;
        unsigned int hi1 = (unsigned int)(p1 >> 32), lo1 = (unsigned int)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double po_u01(unsigned int v)
{
    return ((double)v + 0.5) * 2.3283064365386963e-10;   /* 2^-32 */
}

/* counter = (id, step_lo, step_hi, block), key = (seed_lo, seed_hi);
 * two blocks -> six uniforms -> three Box-Muller pairs. */
void po_philox_normals(unsigned long long seed, unsigned long long step,
                       int id0, int n, double *out)
{
    const double two_pi = 6.283185307179586;
    unsigned int key[2] = {(unsigned int)seed, (unsigned int)(seed >> 32)};
    for (int t = 0; t < n; t++) {
        unsigned int id = (unsigned int)(id0 + t);
        unsigned int u[8];
        unsigned int ctr[4] = {id, (unsigned int)step, (unsigned int)(step >> 32), 0u};
        po_philox4x32_10(ctr, key, u);
        ctr[3] = 1u;
        po_philox4x32_10(ctr, key, u + 4);
        double z[6];
        for (int p = 0; p < 3; p++) {
            double u1 = po_u01(u[2 * p]), u2 = po_u01(u[2 * p + 1]);
            double rr = sqrt(-2.0 * log(u1));
            double th = two_pi * u2;
            z[2 * p] = rr * cos(th);
            z[2 * p + 1] = rr * sin(th);
        }
        for (int k = 0; k < 6; k++) out[(size_t)k * n + t] = z[k];
    }
}

/* ------------------------------------------------------------------ */
/* Full step loop: system.cpp:187-211                                    */
/* ------------------------------------------------------------------ */
int po_run(const po_force_params *fp, const po_gb_coeffs *c,
           int n, double box, double half_uv, int S, double cell_size,
           double radius, int vel_freq, int *vel_counter,
           double *current_time,
           double *pos, double *pos0, double *vel, double *frc, double *frc0,
           double *memx, int nsteps, const double *noise,
           unsigned long long seed, unsigned long long *step_index,
           int *err_ij, int nthreads)
{
    return po_run_ids(fp, c, n, box, half_uv, S, cell_size, radius, vel_freq, vel_counter,
                      current_time, pos, pos0, vel, frc, frc0, memx, nsteps, noise, seed,
                      step_index, err_ij, nthreads, 0);
}

int po_run_ids(const po_force_params *fp, const po_gb_coeffs *c,
               int n, double box, double half_uv, int S, double cell_size,
               double radius, int vel_freq, int *vel_counter,
               double *current_time,
               double *pos, double *pos0, double *vel, double *frc, double *frc0,
               double *memx, int nsteps, const double *noise,
               unsigned long long seed, unsigned long long *step_index,
               int *err_ij, int nthreads, int id0)
{
    int ncell = S * S * S;
    int *key = (int *)malloc(sizeof(int) * (size_t)n);
    int *perm = (int *)malloc(sizeof(int) * (size_t)n);
    int *cs = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *ce = (int *)malloc(sizeof(int) * (size_t)ncell);
    double *sorted = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    double *fnew = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *memcorr = (double *)calloc(3 * (size_t)n, sizeof(double));
    double *nz = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    double *na = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *nb = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    int status = 0;
    for (int s = 0; s < nsteps && !status; s++) {
        /* cell rebuild (the reference rebuilds after integration and in the
         * constructor; rebuilding at the top of each step is equivalent) */
        po_hash(n, pos, cell_size, S, key);
        po_build_cells(n, key, ncell, perm, cs, ce);
        po_reorder(n, perm, pos, radius, sorted);
        status = po_forces(fp, n, sorted, perm, cs, ce, S, cell_size, box,
                           half_uv, fnew, err_ij, nthreads);
        if (status) break;
        /* pushParticleForce -> particle::setForce (particle.cpp:115-131) */
        for (int k = 0; k < 3 * n; k++) { frc0[k] = frc[k]; frc[k] = fnew[k]; }
        /* noise for this step, [6][n] by original index */
        const double *src;
        if (noise) {
            src = noise + (size_t)s * 6 * n;
        } else {
            po_philox_normals(seed, *step_index, id0, n, nz);
            src = nz;
        }
        for (int i = 0; i < n; i++) {
            for (int d = 0; d < 3; d++) {
                na[3 * i + d] = src[(size_t)d * n + i];
                nb[3 * i + d] = src[(size_t)(3 + d) * n + i];
            }
        }
        int ei = -1;
        if (*current_time == 0) {
            status = po_gb_first_step(c, n, box, pos, pos0, vel, frc, memx,
                                      memcorr, nb, &ei);
        } else {
            int do_vel = (vel_freq == 0 || *vel_counter == vel_freq);
            status = po_gb_normal_step(c, n, box, pos, pos0, vel, frc, frc0,
                                       memx, memcorr, na, nb, do_vel, &ei);
            /* :292 */
            if (*vel_counter == vel_freq) *vel_counter = 0; else (*vel_counter)++;
        }
        if (status && err_ij) { err_ij[0] = ei; err_ij[1] = -1; }
        *step_index += 1;
        *current_time += c->dt;
    }
    free(key); free(perm); free(cs); free(ce); free(sorted); free(fnew);
    free(memcorr); free(nz); free(na); free(nb);
    return status;
}

/* ------------------------------------------------------------------ */
/* Analysis                                                              */
/* ------------------------------------------------------------------ */
double po_temperature(int n, const double *vel)
{
    double vtot = 0;
    for (int i = 0; i < n; i++) {
        double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        vtot += (vx * vx);
        vtot += (vy * vy);
        vtot += (vz * vz);
    }
    vtot = vtot / 3.0;
    vtot = vtot / n;
    return vtot;
}

void po_pair_histogram(int n, const double *pos, double box, double r_lo,
                       double width, int nbins, double *counts)
{
    for (int b = 0; b < nbins; b++) counts[b] = 0.0;
    double r_hi = r_lo + nbins * width;
    for (int i = 0; i < n; i++) {
        for (int j = i + 1; j < n; j++) {
            double r2 = po_pbc_dist(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
                                    pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], box);
            double r = sqrt(r2);
            if (r >= r_lo && r < r_hi) {
                int b = (int)((r - r_lo) / width);
                if (b >= 0 && b < nbins) counts[b] += 1.0;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Analysis (continued)                                                  */
/* ------------------------------------------------------------------ */

/* analysisManager::updateParticlePBC over all particles
 * (analysisTracker.cpp:38-64); pos0 and pbc are AoS3 by original index. */
void po_track_update(int n, double box, const double *pos0, int *pbc)
{
    for (int k = 0; k < 3 * n; k++) {
        if (pos0[k] < 0) pbc[k] += 1;
        else if (pos0[k] > box) pbc[k] -= 1;
    }
}

/* Contacts r = sqrt(pbcDist) < rc over the 27 cells around each sorted
 * particle (system::updateInteractions / iterateParticleInteractions,
 * systemHandling.cpp:94-161).  coord is by original index; if labels is not
 * NULL it receives the connected components of the contact graph
 * (analysisManager::findClusters, analysisCluster.cpp:26-83) as the smallest
 * original index of each component. */
static int po_find(int *p, int x)
{
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

void po_contacts(int n, const double *sorted4, const int *perm, const int *cstart,
                 const int *cend, int S, double cell_size, double box, double rc,
                 int *coord, int *labels)
{
    int *parent = NULL;
    if (labels) {
        parent = (int *)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; i++) parent[i] = i;
    }
    int S2 = S * S;
    for (int index = 0; index < n; index++) {
        const double *a = sorted4 + 4 * index;
        int cx = (int)floor(a[0] / cell_size);
        int cy = (int)floor(a[1] / cell_size);
        int cz = (int)floor(a[2] / cell_size);
        int cnt = 0;
        for (int x = -1; x <= 1; x++)
            for (int y = -1; y <= 1; y++)
                for (int z = -1; z <= 1; z++) {
                    int rx = (cx + x) % S, ry = (cy + y) % S, rz = (cz + z) % S;
                    rx = (rx < 0) ? rx + S : rx;
                    ry = (ry < 0) ? ry + S : ry;
                    rz = (rz < 0) ? rz + S : rz;
                    int hash = rx + (S * ry) + (S2 * rz);
                    int start = cstart[hash];
                    if (start == -1) continue;
                    int end = cend[hash];
                    for (int i = start; i < end; i++) {
                        if (i == index) continue;
                        const double *b = sorted4 + 4 * i;
                        double r2 = po_pbc_dist(a[0], a[1], a[2], b[0], b[1], b[2], box);
                        double r = sqrt(r2);
                        if (r < rc) {
                            cnt++;
                            if (parent) {
                                int ra = po_find(parent, perm[index]);
                                int rb = po_find(parent, perm[i]);
                                if (ra < rb) parent[rb] = ra;
                                else if (rb < ra) parent[ra] = rb;
                            }
                        }
                    }
                }
        coord[perm[index]] = cnt;
    }
    if (labels) {
        for (int i = 0; i < n; i++) labels[i] = po_find(parent, i);
        free(parent);
    }
}

/* Gaussian-smoothed pair-distance histogram over ordered pairs i != j
 * (minimum image): bin b = [r_lo + b w, r_lo + (b+1) w) receives
 * Phi((e_{b+1} - r)/sigma) - Phi((e_b - r)/sigma); sigma == 0 counts.
 * Brute force O(n^2); pos is AoS3. */
static double po_ncdf(double t) { return 0.5 * erfc(-t / sqrt(2.0)); }
static double po_npdf(double t) { return exp(-0.5 * t * t) / sqrt(2.0 * M_PI); }

void po_pair_histogram_smooth(int n, const double *pos, double box, double r_lo,
                              double width, int nbins, double sigma, double *hist)
{
    for (int b = 0; b < nbins; b++) hist[b] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            double r = sqrt(po_pbc_dist(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
                                        pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], box));
            if (sigma == 0.0) {
                double q = floor((r - r_lo) / width);
                if (q >= 0 && q < nbins) hist[(int)q] += 1.0;
                continue;
            }
            for (int b = 0; b < nbins; b++) {
                double lo = r_lo + width * b, hi = r_lo + width * (b + 1);
                hist[b] += po_ncdf((hi - r) / sigma) - po_ncdf((lo - r) / sigma);
            }
        }
}

/* grad (AoS3) = d(sum_b w_b hist_b)/d pos for sigma > 0 (brute force). */
void po_pair_histogram_grad(int n, const double *pos, double box, double r_lo,
                            double width, int nbins, double sigma, const double *w,
                            double *grad)
{
    double half = box / 2.0;
    for (int k = 0; k < 3 * n; k++) grad[k] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            double d[3];
            for (int k = 0; k < 3; k++) {
                double e = pos[3 * j + k] - pos[3 * i + k];
                if (e > half) e -= box;
                else if (e < -half) e += box;
                d[k] = e;
            }
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            double s = 0.0;   /* sum_b w_b dm_b/dr */
            for (int b = 0; b < nbins; b++) {
                double lo = r_lo + width * b, hi = r_lo + width * (b + 1);
                s += w[b] * (po_npdf((lo - r) / sigma) - po_npdf((hi - r) / sigma)) / sigma;
            }
            /* the ordered pairs (i,j) and (j,i) both move with x_i */
            for (int k = 0; k < 3; k++) grad[3 * i + k] += -2.0 * s * d[k] / r;
        }
}

/* ------------------------------------------------------------------ */
/* Derivatives of the pair forces and the trajectory adjoint              */
/* ------------------------------------------------------------------ */

/* fNet(r), dfNet/dr and dfNet/dtheta summed over the enabled modules, with
 * theta = (AO wellDepth, cutOff through the AO coefficients, LJ wellDepth,
 * yukawaStrength, debyeLength); differentiated from po_f_lj / po_f_ao /
 * po_f_cal (LJPotential.cpp:97-109, AOPotential.cpp:50-55, 95-102,
 * Calibration.cpp:75-81). */
void po_pair_fnet_d(const po_force_params *fp, double r, double size, double *f,
                    double *dfdr, double dth[5])
{
    double F = 0.0, D = 0.0;
    for (int k = 0; k < 5; k++) dth[k] = 0.0;
    if (fp->mask & PO_CAL) {
        F += -36.0 * pow(r, -37);
        D += 36.0 * 37.0 * pow(r, -38);
    }
    if (fp->mask & PO_LJ) {
        double n = fp->lj_num, lam = fp->debye_length, A = fp->yukawa_strength;
        double kte = fp->kT * fp->lj_well_depth;
        double lj = pow(size / r, n);
        double attract = 4.0 * n * (2.0 * lj * lj - lj) / r;
        double dattract = 4.0 * n * ((-2.0 * n * 2.0 * lj * lj + n * lj) / (r * r)
                                     - (2.0 * lj * lj - lj) / (r * r));
        double e = exp(-r / lam);
        double yb = e * (lam + r) / (r * r);
        double dyb = (-e / lam * (lam + r) + e) / (r * r) - 2.0 * e * (lam + r) / (r * r * r);
        F += -kte * (attract + A * yb);
        D += -kte * (dattract + A * dyb);
        dth[2] += -fp->kT * (attract + A * yb);
        dth[3] += -kte * yb;
        dth[4] += -kte * A * (e * r / (lam * lam) * (lam + r) + e) / (r * r);
    }
    if (fp->mask & PO_AO) {
        double c = fp->cutoff, phi = fp->ao_well_depth, kT = fp->kT;
        double Kc = -kT * pow(c / (c - 1.0), 3);
        double a1 = Kc * phi, a2 = -3.0 / (2.0 * c), a3 = 1.0 / (2.0 * c * c * c);
        double C1 = -a1 * a2, C2 = -3.0 * a1 * a3;
        F += -(36.0 * pow(r, -38) + C1 / r + C2 * r);
        D += 36.0 * 38.0 * pow(r, -39) + C1 / (r * r) - C2;
        /* d/dphi */
        dth[0] += -((-Kc * a2) / r + (-3.0 * Kc * a3) * r);
        /* d/dc: a1' = a1 (3/c - 3/(c-1)), a2' = 3/(2c^2), a3' = -3/(2c^4) */
        double da1 = a1 * (3.0 / c - 3.0 / (c - 1.0));
        double da2 = 3.0 / (2.0 * c * c), da3 = -3.0 / (2.0 * c * c * c * c);
        double dC1 = -(da1 * a2 + a1 * da2), dC2 = -3.0 * (da1 * a3 + a1 * da3);
        dth[1] += -(dC1 / r + dC2 * r);
    }
    *f = F;
    *dfdr = D;
}

/* Reverse-mode derivative of nsteps of po_run (Philox noise keyed by
 * seed / step / original index) for a loss with gradient gx = dL/dx_K:
 * dtheta[5] (see po_pair_fnet_d) and, if gx0 is not NULL, dL/dx_0.  Stores
 * every step's positions (test sizes).  pos_out (may be NULL) receives x_K.
 * Recurrence: lambda_n = alpha_n a_{n+1} - c2 a_{n+2},
 * a_n = beta_n a_{n+1} - coeff0 a_{n+2} + J(x_n) lambda_n,
 * dL/dtheta = sum_n lambda_n . dF(x_n)/dtheta.  With cut_window w > 0 the
 * cutOff derivative also holds the pairs crossing the cutoff:
 * dF_i/dc += sum_j 1{|r_ij - c| < w} / (2 w) fNet(r_ij) d_ij / r_ij. */
int po_adjoint(const po_force_params *fp, const po_gb_coeffs *c, int n, double box,
               double half_uv, int S, double cell_size, double radius, int vel_freq,
               int vel_counter, double current_time, const double *pos_in,
               const double *pos0_in, const double *vel_in, const double *frc_in,
               const double *memx_in, int nsteps, unsigned long long seed,
               unsigned long long step_index, const double *gx, double *dtheta, double *gx0,
               double *pos_out, int nthreads, double cut_window)
{
    size_t n3 = 3 * (size_t)n;
    int ncell = S * S * S;
    double *traj = (double *)malloc(sizeof(double) * n3 * (size_t)nsteps);
    int *first = (int *)malloc(sizeof(int) * (size_t)nsteps);
    double *pos = (double *)malloc(sizeof(double) * n3);
    double *pos0 = (double *)malloc(sizeof(double) * n3);
    double *vel = (double *)malloc(sizeof(double) * n3);
    double *frc = (double *)malloc(sizeof(double) * n3);
    double *frc0 = (double *)calloc(n3, sizeof(double));
    double *memx = (double *)malloc(sizeof(double) * n3);
    memcpy(pos, pos_in, sizeof(double) * n3);
    memcpy(pos0, pos0_in, sizeof(double) * n3);
    memcpy(vel, vel_in, sizeof(double) * n3);
    memcpy(frc, frc_in, sizeof(double) * n3);
    memcpy(memx, memx_in, sizeof(double) * n3);
    int status = 0, eij[2];
    for (int s = 0; s < nsteps && !status; s++) {
        memcpy(traj + (size_t)s * n3, pos, sizeof(double) * n3);
        first[s] = (current_time == 0);
        status = po_run(fp, c, n, box, half_uv, S, cell_size, radius, vel_freq, &vel_counter,
                        &current_time, pos, pos0, vel, frc, frc0, memx, 1, NULL, seed,
                        &step_index, eij, nthreads);
    }
    if (pos_out) memcpy(pos_out, pos, sizeof(double) * n3);
    double m = 1.0 / c->mass, dt2 = c->dt * c->dt;
    double c1 = m * dt2 * c->coeff1, c2 = m * dt2 * c->coeff2;
    double alpha_first = c->coeff3 * dt2 * m, alpha_normal = c1 + c2;
    double *a0 = (double *)malloc(sizeof(double) * n3);
    double *a1 = (double *)malloc(sizeof(double) * n3);
    double *a2 = (double *)calloc(n3, sizeof(double));
    double *lam = (double *)malloc(sizeof(double) * n3);
    double *lamp = (double *)malloc(sizeof(double) * n3);
    double *jl = (double *)malloc(sizeof(double) * n3);
    int *key = (int *)malloc(sizeof(int) * (size_t)n);
    int *perm = (int *)malloc(sizeof(int) * (size_t)n);
    int *cs = (int *)malloc(sizeof(int) * (size_t)ncell);
    int *ce = (int *)malloc(sizeof(int) * (size_t)ncell);
    double *sorted = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    double th[5] = {0, 0, 0, 0, 0};
    memcpy(a1, gx, sizeof(double) * n3);
    double aK = first[nsteps - 1] ? alpha_first : alpha_normal;
    for (size_t k = 0; k < n3; k++) lam[k] = aK * a1[k];
    for (int s = nsteps - 1; s >= 0 && !status; s--) {
        const double *x = traj + (size_t)s * n3;
        po_hash(n, x, cell_size, S, key);
        po_build_cells(n, key, ncell, perm, cs, ce);
        po_reorder(n, perm, x, radius, sorted);
        for (size_t k = 0; k < n3; k++) jl[k] = 0.0;
        for (int index = 0; index < n; index++) {
            const double *a = sorted + 4 * index;
            int oi = perm[index];
            int cx = po_cell_1d(a[0], cell_size, S);
            int cy = po_cell_1d(a[1], cell_size, S);
            int cz = po_cell_1d(a[2], cell_size, S);
            for (int ox = -1; ox <= 1; ox++)
                for (int oy = -1; oy <= 1; oy++)
                    for (int oz = -1; oz <= 1; oz++) {
                        int rx = (cx + ox + S) % S, ry = (cy + oy + S) % S, rz = (cz + oz + S) % S;
                        int h = rx + S * ry + S * S * rz;
                        if (cs[h] == -1) continue;
                        for (int i = cs[h]; i < ce[h]; i++) {
                            if (i == index) continue;
                            const double *b = sorted + 4 * i;
                            double r2 = po_pbc_dist(a[0], a[1], a[2], b[0], b[1], b[2], box);
                            double cw = fp->cutoff + cut_window;
                            int inside = r2 < fp->cutoff * fp->cutoff;
                            if (!inside && !(cut_window > 0.0 && r2 < cw * cw)) continue;
                            double r = sqrt(r2);
                            double u[3];
                            po_unit_vector(a[0], a[1], a[2], b[0], b[1], b[2], u, r, box,
                                           half_uv);
                            double d[3] = {u[0] * r, u[1] * r, u[2] * r};
                            double f, dfdr, dth[5];
                            po_pair_fnet_d(fp, r, a[3] + b[3], &f, &dfdr, dth);
                            double g = f / r, gp = (dfdr - g) / r;
                            int oj = perm[i];
                            double e[3], dd = 0.0, ld = 0.0;
                            for (int k = 0; k < 3; k++) {
                                e[k] = lam[3 * oj + k] - lam[3 * oi + k];
                                dd += d[k] * e[k];
                                ld += lam[3 * oi + k] * d[k];
                            }
                            if (fabs(r - fp->cutoff) < cut_window)
                                th[1] += f / (2.0 * cut_window) / r * ld;
                            if (!inside) continue;
                            for (int k = 0; k < 3; k++)
                                jl[3 * oi + k] += g * e[k] + gp / r * dd * d[k];
                            for (int q = 0; q < 5; q++) th[q] += dth[q] / r * ld;
                        }
                    }
        }
        double beta = first[s] ? 1.0 : 1.0 + c->coeff0;
        for (size_t k = 0; k < n3; k++) a0[k] = beta * a1[k] - c->coeff0 * a2[k] + jl[k];
        if (s >= 1) {
            double ap = first[s - 1] ? alpha_first : alpha_normal;
            for (size_t k = 0; k < n3; k++) lamp[k] = ap * a0[k] - c2 * a1[k];
        }
        double *t = a2; a2 = a1; a1 = a0; a0 = t;
        t = lam; lam = lamp; lamp = t;
    }
    for (int q = 0; q < 5; q++) dtheta[q] = th[q];
    if (gx0) memcpy(gx0, a1, sizeof(double) * n3);
    free(traj); free(first); free(pos); free(pos0); free(vel); free(frc); free(frc0);
    free(memx); free(a0); free(a1); free(a2); free(lam); free(lamp); free(jl); free(key);
    free(perm); free(cs); free(ce); free(sorted);
    return status;
}
