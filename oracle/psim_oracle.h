/*
 * psim_oracle.h -- CPU restatement of the ProteinSim2015 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libpsim_gpu.so and the plugin
 * adaptors) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arrays are plain C arrays; "AoS3" means
 * n x 3 doubles (x,y,z per particle) in ORIGINAL particle order, "sorted4"
 * means the reference's sortedParticles layout (x,y,z,r per sorted index).
 *
 * Parity pinning (what checks each function, see tests/):
 *   po_geometry            closed form + reference integer arithmetic   pinned
 *   po_hash/po_build_cells brute-force recount, std::stable_sort order,
 *                          reference harness cell table (set equality)  pinned
 *   po_forces              bit-exact vs compiled reference plugins fed
 *                          the same cell table; two-particle closed forms pinned
 *   po_pair_stats          brute-force O(N^2) minimum-image partner sets
 *                          with pbcDist's arithmetic; abs_sum vs the
 *                          closed-form pair forces                       pinned
 *   po_gb_*                bit-exact vs compiled reference integrator;
 *                          series cross-check at small y                pinned
 *   po_fastrandom_streams  bit-exact vs compiled reference fastRandom    pinned
 *   po_philox_normals      Random123 published known-answer vectors for
 *                          Philox4x32-10; moments of the normals        pinned
 *   po_run                 composition of the above, vs reference
 *                          trajectories (tolerance: sort order)         pinned
 *   analysis (T, MSD, g(r))closed forms (ideal-gas g(r)=1, OU MSD)      pinned
 *   po_contacts            brute-force neighbour count, union-find of
 *                          brute-force contacts                          pinned
 *   po_pair_histogram_smooth  sigma -> 0 equals the plain histogram;
 *                          total mass = pairs in range                   pinned
 *   po_pair_histogram_grad central differences of the histogram          pinned
 *   po_pair_fnet_d         central differences of po_forces on two
 *                          particles (r and each parameter)             pinned
 *   po_adjoint             central differences of the loss of po_run
 *                          trajectories (pair set unchanged)            pinned
 */
#ifndef PSIM_ORACLE_H
#define PSIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Force-module selection bits (a composite is the SUM of separate module
 * passes, because the reference evaluates only flist[0] per run,
 * defaultForceManager.cpp:56). */
enum { PO_CAL = 1, PO_LJ = 2, PO_AO = 4 };

typedef struct po_force_params {
    int mask;               /* PO_CAL | PO_LJ | PO_AO                       */
    int lj_num;             /* ljNum (LJPotential.cpp:47)                    */
    double cutoff;          /* cutOff, shared key (LJPotential.cpp:50,
                               AOPotential.cpp:47, Calibration.cpp:36)       */
    double kT;              /* kT                                            */
    double lj_well_depth;   /* wellDepth for the LJ module (epsilon)        */
    double yukawa_strength; /* yukawaStrength                                */
    double debye_length;    /* debyeLength                                   */
    double ao_well_depth;   /* wellDepth for the AO module (phi_p)          */
} po_force_params;

typedef struct po_gb_coeffs {
    double kT, mass, gamma, dt, dt_inv, y;
    double coeff0, coeff1, coeff2, coeff3;
    double sig1, sig2, corr, dev;
    double goy2, goy3, hn, gn;
} po_gb_coeffs;

/* system.cpp:50-64 (reference integer geometry). */
void po_geometry(int n, double conc, int scale, double radius,
                 int *box, int *cell_size, double *concentration);

/* systemHandling.cpp:33-48; cells clamped to [0,S-1] (the reference has
 * undefined behaviour when x/cellSize rounds up to S). pos is AoS3. */
void po_hash(int n, const double *pos, double cell_size, int S, int *key);

/* Stable counting sort by key (the reference's sort, systemHandling.cpp:50-52,
 * is unstable; the within-cell order is implementation defined there) and the
 * cell start/end table of systemHandling.cpp:54-75 with empty = -1
 * (system.h:166).  perm[sorted] = original index. */
void po_build_cells(int n, const int *key, int ncell,
                    int *perm, int *cstart, int *cend);

/* systemHandling.cpp:77-84: sorted4[4i..4i+3] = (x,y,z,r) of perm[i]. */
void po_reorder(int n, const int *perm, const double *pos, double radius,
                double *sorted4);

/* Forces on every particle, written by ORIGINAL index into force3 (AoS3).
 * Per-module pass restates LJPotential.cpp:64-182 / AOPotential.cpp:61-175 /
 * Calibration.cpp:41-147 with util::pbcDist (utilities.cpp:79-104) and
 * util::unitVectorAdv (utilities.cpp:114-142).  half_uv is the half box used
 * by unitVectorAdv (integer L/2 for integer geometry).  Returns 0, 7704
 * (overlap) or 7705 (AO NaN); err_ij receives the (sorted) pair indices.
 * nthreads<=0 uses the OpenMP default. */
int po_forces(const po_force_params *fp, int n, const double *sorted4,
              const int *perm, const int *cstart, const int *cend,
              int S, double cell_size, double box, double half_uv,
              double *force3, int *err_ij, int nthreads);

/* Neighbour sets of po_forces' loop (LJPotential.cpp:64-95, 139-177: the
 * 27-cell stencil, j != index, util::pbcDist, strict r^2 < cutOff^2): per
 * particle (ORIGINAL index) the partner count, the sum mod 2^64 of
 * h(partner's original index) with h the SplitMix64 finaliser of
 * (index + 0x9E3779B97F4A7C15), and (abs_sum, may be NULL) the sum over
 * partners of |fNet| of the enabled modules together.  Partners are named by
 * original index + id0 (a replica's offset in a stacked ensemble). */
void po_pair_stats(const po_force_params *fp, int n, const double *sorted4,
                   const int *perm, const int *cstart, const int *cend,
                   int S, double cell_size, double box, int *count,
                   unsigned long long *hash, double *abs_sum, int id0, int nthreads);

/* brownianIntegrator.cpp:53-81 + setupHigh/Low/Zero :111-175. */
void po_gb_coeffs_init(double kT, double mass, double gamma, double dt,
                       po_gb_coeffs *c);

/* brownianIntegrator.cpp:191-223 (first step) with particle::setPos
 * (particle.cpp:61-99).  noise_b is the 3 draws per particle (AoS3). */
int po_gb_first_step(const po_gb_coeffs *c, int n, double box,
                     double *pos, double *pos0, const double *vel,
                     const double *frc, double *memx, double *memcorr,
                     const double *noise_b, int *err_i);

/* brownianIntegrator.cpp:225-296 (+ velocityStep :298-356 when do_vel). */
int po_gb_normal_step(const po_gb_coeffs *c, int n, double box,
                      double *pos, double *pos0, double *vel,
                      const double *frc, const double *frc0,
                      double *memx, double *memcorr,
                      const double *noise_a, const double *noise_b,
                      int do_vel, int *err_i);

/* Replays fastRandom::g250 with threadSeed = seed*(i+1) for each particle
 * (fastRandom.cpp:40-124, brownianIntegrator.cpp:204,236).  out[i*ndraws+k].
 * Returns 0, or -1 when a stream would re-seed on every call (negative
 * threadSeed, the int overflow of brownianIntegrator.cpp:236). */
int po_fastrandom_streams(int seed, int n, int ndraws, double *out);

/* Philox4x32-10 counter-based normals (not in the reference; the in-kernel
 * noise source of the HIP path).  out is [6][n]: A_x,A_y,A_z,B_x,B_y,B_z for
 * particles with original index id0 .. id0+n-1. */
void po_philox_normals(unsigned long long seed, unsigned long long step,
                       int id0, int n, double *out);
/* Raw Philox4x32-10 block (for known-answer tests). */
void po_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2],
                      unsigned int out[4]);

/* Full steps in the order of system::run (system.cpp:187-211): build cells,
 * forces, push (frc0<-frc, frc<-F), integrate, time += dt.  State arrays are
 * AoS3 in original order.  noise: NULL -> Philox(seed, *step_index, id);
 * else [nsteps][6][n] by original index (first step reads the B slots).
 * Returns 0 or the first error code (7703/7704/7705). */
int po_run(const po_force_params *fp, const po_gb_coeffs *c,
           int n, double box, double half_uv, int S, double cell_size,
           double radius, int vel_freq, int *vel_counter,
           double *current_time,
           double *pos, double *pos0, double *vel, double *frc, double *frc0,
           double *memx, int nsteps, const double *noise,
           unsigned long long seed, unsigned long long *step_index,
           int *err_ij, int nthreads);
/* po_run with the in-kernel noise keyed by original ids id0 .. id0+n-1 */
int po_run_ids(const po_force_params *fp, const po_gb_coeffs *c,
               int n, double box, double half_uv, int S, double cell_size,
               double radius, int vel_freq, int *vel_counter,
               double *current_time,
               double *pos, double *pos0, double *vel, double *frc, double *frc0,
               double *memx, int nsteps, const double *noise,
               unsigned long long seed, unsigned long long *step_index,
               int *err_ij, int nthreads, int id0);

/* analysisGeneral.cpp:26-44: sum v^2 / 3 / N. */
double po_temperature(int n, const double *vel);
/* Radial distribution histogram over min-image pairs with r in
 * [r_lo, r_lo + nbins*width): counts[b] (each unordered pair once). */
void po_pair_histogram(int n, const double *pos, double box, double r_lo,
                       double width, int nbins, double *counts);

/* analysisTracker.cpp:38-64: periodic-crossing counters (AoS3 ints by
 * original index) from the re-imaged previous positions. */
void po_track_update(int n, double box, const double *pos0, int *pbc);
/* systemHandling.cpp:94-161 contact scan (r < rc over 27 cells) on the
 * reference's sorted arrays: coord by original index; labels (may be NULL)
 * = smallest original index of each contact component
 * (analysisCluster.cpp:26-83). */
void po_contacts(int n, const double *sorted4, const int *perm, const int *cstart,
                 const int *cend, int S, double cell_size, double box, double rc,
                 int *coord, int *labels);
/* Gaussian-smoothed pair histogram over ordered pairs (brute force) and
 * the gradient of sum_b w_b hist_b with respect to the positions. */
void po_pair_histogram_smooth(int n, const double *pos, double box, double r_lo,
                              double width, int nbins, double sigma, double *hist);
void po_pair_histogram_grad(int n, const double *pos, double box, double r_lo,
                            double width, int nbins, double sigma, const double *w,
                            double *grad);
/* fNet, dfNet/dr and dfNet/dtheta (theta = AO wellDepth, cutOff via the AO
 * coefficients, LJ wellDepth, yukawaStrength, debyeLength). */
void po_pair_fnet_d(const po_force_params *fp, double r, double size, double *f,
                    double *dfdr, double dth[5]);
/* Trajectory adjoint: nsteps of po_run with Philox noise, loss gradient gx
 * at the last step -> dtheta[5], dL/dx_0 (gx0, may be NULL), x_K (pos_out,
 * may be NULL); cut_window > 0 adds the cutoff-crossing term to dtheta[1]. */
int po_adjoint(const po_force_params *fp, const po_gb_coeffs *c, int n, double box,
               double half_uv, int S, double cell_size, double radius, int vel_freq,
               int vel_counter, double current_time, const double *pos_in,
               const double *pos0_in, const double *vel_in, const double *frc_in,
               const double *memx_in, int nsteps, unsigned long long seed,
               unsigned long long step_index, const double *gx, double *dtheta, double *gx0,
               double *pos_out, int nthreads, double cut_window);

#ifdef __cplusplus
}
#endif
#endif
