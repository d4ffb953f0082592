/*
 * psim_gpu.h -- C ABI of the MI355X (gfx950) ProteinSim2015 hot path.
 *
 * Plain C: pointers, sizes and PODs only.  Host arrays are caller-owned and
 * borrowed for the duration of a call; device memory is owned by the context.
 * Status codes: 0 = OK; the reference's own exit codes for physics errors
 * (7703 particle out of box, 7704 overlap, 7705 non-finite AO force; see
 * ClusteredCore/src/utilities/error.cpp:40-62, where the reference calls
 * exit()); negative values for API/runtime errors (PSIM_E_*).
 *
 * One context per host thread (or per stream); calls on one context are not
 * re-entrant.  AoS3 = n x 3 doubles (x,y,z) in ORIGINAL particle order.
 *
 * Which reference interface each entry point replaces:
 *   psim_gpu_create        PSim::system::setSystemConstants + the force
 *                          plugin / integrator constructors
 *                          (system.cpp:27-67, LJPotential.cpp:29-62,
 *                          AOPotential.cpp:33-59, Calibration.cpp:30-39,
 *                          brownianIntegrator.cpp:27-99)
 *   psim_gpu_set_state     particle setters + the constructor's first cell
 *                          build (particle.cpp:61-99, system.cpp:96-103)
 *   psim_gpu_step          nsteps iterations of system::run
 *                          (system.cpp:187-211): forces
 *                          (defaultForceManager.cpp:55-65 ->
 *                          IForce::getAcceleration, IForce.h:35),
 *                          pushParticleForce (systemHandling.cpp:87-92),
 *                          IIntegrator::nextSystem (IIntegrator.h:44,
 *                          brownianIntegrator.cpp:181-356), cell rebuild
 *                          (systemHandling.cpp:33-85)
 *   psim_gpu_forces        defaultForceManager::getAcceleration alone
 *   psim_gpu_force_table   IForce::getAcceleration over the reference's own
 *                          sortedParticles / particleHashIndex /
 *                          cellStartEnd arrays (IForce.h:35), for the plugin
 *                          adaptor's batch path
 *   psim_gpu_get_state     particle getters (particle.h:82-240)
 *   psim_gpu_last_error    error::throw* (error.cpp:40-62), as a code
 *   psim_write_recovery /  analysisManager::writeSystem / system::readParticles
 *   psim_read_recovery     (analysisOutput.cpp:55-74, systemRecovery.cpp:30-81)
 *   psim_write_xyz         analysisManager::writeSystemXYZ (analysisOutput.cpp:31-53)
 *   psim_write_sys_config  system::writeSystemInit (system.cpp:108-127)
 *   psim_init_rsa          system::initParticles + initCheck
 *                          (systemInit.cpp:31-160), O(N) with a cell grid
 *   psim_init_reference    system::initParticles + initCheck +
 *                          maxwellVelocityInit (systemInit.cpp:31-267),
 *                          the same draws in the same order
 *   psim_gpu_set_force_params  the plugin constructors' parameter reads
 *                          (LJPotential.cpp:34-55, AOPotential.cpp:40-55),
 *                          applied to an existing context
 *   psim_gpu_track_start   analysisManager::initDisplacementTracker
 *                          (analysisTracker.cpp:26-36)
 *   psim_gpu_observables   getTemperature / meanDisplacement
 *                          (analysisGeneral.cpp:26-56), trackedDisplacement
 *                          (analysisTracker.cpp:67-85), coordination from
 *                          updateInteractions (systemHandling.cpp:94-161,
 *                          analysisOutput.cpp:116-126)
 *   psim_gpu_clusters      analysisManager::findClusters / writeClusters
 *                          (analysisCluster.cpp:26-100)
 */
#ifndef PSIM_GPU_H
#define PSIM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PSIM_GPU_ABI_VERSION 1

/* force-module bits; a composite is the sum of the modules' forces */
enum {
    PSIM_FORCE_CAL = 1, /* Calibration soft shell, Calibration.cpp:75-81   */
    PSIM_FORCE_LJ = 2,  /* Lennard-Jones 2n-n + Yukawa, LJPotential.cpp:97-109 */
    PSIM_FORCE_AO = 4   /* Asakura-Oosawa + r^-38 core, AOPotential.cpp:95-109 */
};

enum {
    PSIM_OK = 0,
    PSIM_ERR_INIT = 7701,    /* initial configuration cannot be placed   */
    PSIM_ERR_BOUNDS = 7703,  /* particle out of box after the single wrap */
    PSIM_ERR_OVERLAP = 7704, /* pair closer than 0.8 (r_i + r_j)          */
    PSIM_ERR_NAN = 7705,     /* non-finite AO pair force                  */
    PSIM_E_ARG = -1,
    PSIM_E_HIP = -2,
    PSIM_E_NOMEM = -3,
    PSIM_E_UNSUPPORTED = -4,
    PSIM_E_IO = -5           /* file missing, unreadable or malformed     */
};

typedef struct psim_params {
    int n_particles;        /* nParticles                                   */
    int cells_per_dim;      /* S, the reference's cellScale (>= 3)          */
    double box;             /* L; an integer in the reference's geometry    */
    double cell_size;       /* cell edge used by the hash (cellSize)        */
    double half_uv;         /* half box used by unitVectorAdv; 0 -> integer
                               L/2 when L is integral, else L/2             */
    double radius;          /* particle radius (all particles)              */
    double mass;
    double kT;
    double gamma;
    double dt;              /* timeStep                                     */
    int vel_freq;           /* velFreq                                      */
    int force_mask;         /* PSIM_FORCE_* bits                            */
    int lj_num;             /* ljNum                                        */
    int kernel_path;        /* 0 auto (by density), 1 thread per particle,
                               2 cell per workgroup (needs S >= 4),
                               3 brick of cells per workgroup               */
    double cutoff;          /* cutOff (shared by all modules)               */
    double lj_well_depth;   /* wellDepth of the LJ module (epsilon)         */
    double yukawa_strength; /* yukawaStrength                               */
    double debye_length;    /* debyeLength                                  */
    double ao_well_depth;   /* wellDepth of the AO module (phi_p)           */
    unsigned long long seed;/* in-kernel noise seed                        */
} psim_params;

typedef struct psim_ctx psim_ctx;

/* per-kernel device time accumulated since the last reset (HIP events on
 * the context's stream), in milliseconds, with launch counts */
typedef struct psim_kernel_times {
    double force_ms;     int force_launches;     /* pair kernel (fused G&B on
                                                    the thread-per-particle path) */
    double integrate_ms; int integrate_launches; /* separate G&B kernel (cell-
                                                    per-workgroup path) */
    double build_ms;     int build_launches;     /* scan + scatter + reorder */
    double total_ms;     int steps;
} psim_kernel_times;

int psim_gpu_abi_version(void);
const char *psim_gpu_status_string(int status);

int psim_gpu_create(const psim_params *params, psim_ctx **out);
void psim_gpu_destroy(psim_ctx *ctx);

/* Upload a state.  pos is required; pos0/vel/frc/frc0/memx may be NULL
 * (zeros; pos0 NULL = the reference's fresh-particle image, i.e.
 * safeMod0(0, x)).  current_time == 0 selects the first-step branch. */
int psim_gpu_set_state(psim_ctx *ctx, const double *pos, const double *pos0,
                       const double *vel, const double *frc, const double *frc0,
                       const double *memx, double current_time, int vel_counter,
                       unsigned long long step_index);
int psim_gpu_get_state(psim_ctx *ctx, double *pos, double *pos0, double *vel,
                       double *frc, double *frc0, double *memx,
                       double *current_time, int *vel_counter,
                       unsigned long long *step_index);

/* Advance nsteps.  noise == NULL: in-kernel Philox4x32-10 normals keyed by
 * (seed, step index, original id).  Otherwise host array [nsteps][6][n] by
 * original index (A_x,A_y,A_z,B_x,B_y,B_z; the first step reads B). */
int psim_gpu_step(psim_ctx *ctx, int nsteps, const double *noise);
/* As psim_gpu_step with noise already resident in device memory. */
int psim_gpu_step_device_noise(psim_ctx *ctx, int nsteps, const double *d_noise);

/* steps_per_graph > 0 (even): psim_gpu_step_device_noise and psim_gpu_step
 * without injected noise run whole blocks of that many steps as one replay
 * of a captured hipGraph (captured on first use, again after set_state,
 * load, set_force_params or track_start), the rest eagerly; 0: every step
 * eagerly.  Results are the same bits either way; kernel timing
 * (psim_gpu_set_profiling) runs eagerly.  No reference counterpart (the
 * reference runs one CPU loop per step, system.cpp:187-211). */
int psim_gpu_set_graph(psim_ctx *ctx, int steps_per_graph);

/* The parameters the context was created with (after psim_gpu_set_force_params). */
int psim_gpu_get_params(psim_ctx *ctx, psim_params *out);

/* The pair kernel the context runs: *path = 1 thread per particle, 2 cell
 * per workgroup, 3 brick of cells per workgroup; brick (may be NULL) gets
 * the brick's home cells along x, y, z (0 unless path 3) and lanes per
 * particle.  No reference counterpart (the reference has one CPU loop). */
int psim_gpu_kernel_info(psim_ctx *ctx, int *path, int brick[4]);

/* Forces at the current positions (AoS3, original order); no state change. */
int psim_gpu_forces(psim_ctx *ctx, double *force_out);

/* Forces from the reference's own host arrays: sorted4 = sortedParticles
 * (x,y,z,r per sorted index), hash_index = particleHashIndex as 2n ints
 * (hash, original index), cell_start_end = cellStartEnd as 2*S^3 ints
 * (start, end; -1 empty).  force_out is particleForce (AoS3, original
 * index).  n must equal the context's n_particles. */
int psim_gpu_force_table(psim_ctx *ctx, const double *sorted4,
                         const int *hash_index, const int *cell_start_end,
                         double *force_out);

/* Neighbour sets of the current positions as the pair kernels see them: for
 * each particle (original index) the number of partners its pair test
 * accepts (the strict r^2 < cutOff^2 of LJPotential.cpp:82 over the 27-cell
 * stencil of LJPotential.cpp:139-177 with util::pbcDist,
 * utilities.cpp:79-104) and the sum mod 2^64 of h(partner's original
 * index), h = the SplitMix64 finaliser of (index + 0x9E3779B97F4A7C15).
 * which = PSIM_PAIRS_STEP runs the acceptance code and launch shape of the
 * pair kernel the context steps with; PSIM_PAIRS_ADJOINT that of the
 * trajectory adjoint's backward step.  No state change.  Replaces no
 * reference call: it exposes what iterCells (LJPotential.cpp:64-133)
 * computes internally, so the neighbour sets can be compared exactly. */
enum { PSIM_PAIRS_STEP = 0, PSIM_PAIRS_ADJOINT = 1 };
int psim_gpu_pair_stats(psim_ctx *ctx, int which, int *count_by_orig,
                        unsigned long long *hash_by_orig);

/* Forces at caller-held positions with NO device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  params as for psim_gpu_create; d_pos and d_force are
 * DEVICE arrays of n x 3 doubles by original index; stream is a hipStream_t
 * (NULL: the null stream).  The cell build's working arrays live in a
 * static device array of PSIM_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends), so n and S^3 are bounded: about 40 n + 8 S^3 bytes
 * must fit (else PSIM_E_NOMEM and a message on stderr).  Calls that overlap
 * on one GPU race on that array.  Returns after the forces are written: 0,
 * 7704 or 7705 as psim_gpu_forces, or PSIM_E_*.  The forces equal
 * psim_gpu_forces' on a context with the same params and positions. */
#define PSIM_NOALLOC_SCRATCH_BYTES (64u << 20)
int psim_gpu_forces_noalloc(const psim_params *params, const double *d_pos, double *d_force,
                            void *stream);

/* Cell data of the current positions for inspection: key_by_orig[n]
 * (cell hash per original index), cell_start[S^3+1] (slot offsets),
 * orig_by_slot[n] (original index stored at each slot). */
int psim_gpu_get_cells(psim_ctx *ctx, int *key_by_orig, int *cell_start,
                       int *orig_by_slot);

/* First recorded physics error since the last clear: code (0 if none) and
 * the original indices of the particle(s) involved. */
int psim_gpu_last_error(psim_ctx *ctx, int *code, int *i, int *j);
int psim_gpu_clear_error(psim_ctx *ctx);

int psim_gpu_sync(psim_ctx *ctx);
int psim_gpu_set_profiling(psim_ctx *ctx, int enabled);
int psim_gpu_kernel_times(psim_ctx *ctx, psim_kernel_times *out, int reset);
/* The context's HIP stream (hipStream_t), for callers that order work
 * against it. */
void *psim_gpu_stream(psim_ctx *ctx);

/* ---- replica ensemble ---------------------------------------------------
 * Many independent systems of uneven N, box and force strength advanced
 * together: one launch per kernel stage per step for all replicas, and the
 * step sequence can be captured once in a hipGraph and replayed.  Replicas
 * share the integrator parameters and force modules of `common` (kT, gamma,
 * mass, dt, velFreq, seed, radius, force_mask; kernel_path 3 starts with
 * bricks of cells per workgroup instead of the thread-per-particle pair
 * kernel, with identical results); geometry and force strengths
 * are per replica.  Arrays are AoS3 concatenated in replica order (replica r
 * owns original ids [sum_{q<r} n_q, sum_{q<=r} n_q)).  No reference
 * counterpart: the reference runs one system per process. */
typedef struct psim_replica {
    int n_particles;
    int cells_per_dim;
    double box;
    double cell_size;
    double cutoff;
    double lj_well_depth;
    double yukawa_strength;
    double debye_length;
    double ao_well_depth;
    int lj_num;
    int reserved;
} psim_replica;

typedef struct psim_ens psim_ens;

int psim_ens_create(const psim_params *common, const psim_replica *reps, int nrep,
                    psim_ens **out);
void psim_ens_destroy(psim_ens *ens);
int psim_ens_total_particles(psim_ens *ens);
/* vel may be NULL; current_time == 0 makes the next step a first step */
int psim_ens_set_state(psim_ens *ens, const double *pos, const double *vel,
                       double current_time);
int psim_ens_get_state(psim_ens *ens, double *pos, double *pos0, double *vel, double *frc);
/* steps_per_graph > 0: psim_ens_step replays a captured graph of that many
 * steps (captured on first use); 0: eager launches */
int psim_ens_set_graph(psim_ens *ens, int steps_per_graph);
/* nsteps with in-kernel Philox noise; asynchronous (see psim_ens_sync) */
int psim_ens_step(psim_ens *ens, int nsteps);
int psim_ens_sync(psim_ens *ens);
int psim_ens_last_error(psim_ens *ens, int *code, int *i, int *j);
/* Neighbour sets of every replica as the ensemble's current pair kernel sees
 * them (count and partner-hash sum per particle, by concatenated original
 * index, partners named by their concatenated original index; as
 * psim_gpu_pair_stats). */
int psim_ens_pair_stats(psim_ens *ens, int *count_by_orig, unsigned long long *hash_by_orig);
/* measured choice of the pair kernel, made separately for each occupancy
 * class of replicas (particles per cell < 0.25, < 0.5, < 1, >= 1; see
 * psim_ens_class_info): kernel 1 = thread per particle with workgroup size
 * tile (64/128/256), 3 = bricks of cells per workgroup with about tile home
 * particles per brick (96/192/384); the same is timed for all replicas in
 * one launch, and whole steps decide between one launch per class and one
 * for all (psim_ens_merged).  *kernel and *tile report the class holding the
 * most particles, *ms_per_step the chosen configuration's whole steps; those
 * steps advance the ensemble, and no choice changes a result */
int psim_ens_autotune(psim_ens *ens, int trial_steps, int *kernel, int *tile,
                      float *ms_per_step);
/* occupancy class cls (0..PSIM_ENS_CLASSES-1): its number of replicas and the
 * pair kernel and tile that run them (as psim_ens_autotune) */
enum { PSIM_ENS_CLASSES = 4 };
int psim_ens_class_info(psim_ens *ens, int cls, int *nrep, int *kernel, int *tile);
/* 1 when the autotune found one pair launch for all replicas faster than one
 * per class (every class then reports that launch's kernel and tile), else 0 */
int psim_ens_merged(psim_ens *ens);

/* ---- G&B integrator over a host's particle arrays -------------------------
 * The GPU side of the IIntegrator adaptor (include/psim/IIntegrator.h):
 * brownianIntegrator's constructor (brownianIntegrator.cpp:27-99) and
 * nextSystem (:181-356) for a host that owns its particles.  params gives
 * n_particles, kT, gamma, dt, vel_freq and seed (the reference's int seed);
 * mass is per particle in psim_int_next.  noise PSIM_NOISE_REFERENCE draws
 * the reference's own per-particle fastRandom streams (threadSeed =
 * seed * (i + 1), fastRandom.cpp:40-124) on the device; PSIM_NOISE_PHILOX
 * draws Philox4x32-10 normals keyed by (seed, call count, particle index).
 * Device memory: about 340 bytes per particle. */
enum { PSIM_NOISE_REFERENCE = 0, PSIM_NOISE_PHILOX = 1 };
typedef struct psim_integrator psim_integrator;
int psim_int_create(const psim_params *params, int noise, psim_integrator **out);
void psim_int_destroy(psim_integrator *integ);
/* One IIntegrator::nextSystem call (IIntegrator.h:44): current_time == 0
 * takes firstStep (:191-223), else normalStep (:225-296) with velocityStep
 * (:298-356) when velFreq == 0 or the call counter reaches velFreq.  Host
 * arrays in the host's particle order: pos, pos0 (updated as
 * particle::setPos leaves them, particle.cpp:61-99) and vel (updated on
 * velocity steps) are n x 3 in/out; frc, frc0 (n x 3) and mass (n) are
 * read.  Returns 0, PSIM_ERR_BOUNDS (7703, *err_index = the particle) or
 * PSIM_E_UNSUPPORTED when a reference stream re-seeds on every draw and
 * cannot produce a deviate (a negative threadSeed: the reference loops
 * forever there). */
int psim_int_next(psim_integrator *integ, double current_time, int box, double *pos,
                  double *pos0, double *vel, const double *frc, const double *frc0,
                  const double *mass, int *err_index);

/* ---- analysis of the current state --------------------------------------- */

/* Replace the force-module parameters (force_mask, lj_num, cutoff,
 * lj_well_depth, yukawa_strength, debye_length, ao_well_depth) of a context;
 * geometry, integrator and state are unchanged.  Other fields of params are
 * ignored. */
int psim_gpu_set_force_params(psim_ctx *ctx, const psim_params *params);

/* Start the displacement tracker at the current positions: zero the
 * periodic-crossing counters, which every later step updates. */
int psim_gpu_track_start(psim_ctx *ctx);

/* The tracker's start positions and periodic-crossing counters (AoS3,
 * original order; either may be NULL).  PSIM_E_UNSUPPORTED before
 * psim_gpu_track_start.  set restores both (a resumed run). */
int psim_gpu_get_tracker(psim_ctx *ctx, double *start, int *crossings);
int psim_gpu_set_tracker(psim_ctx *ctx, const double *start, const int *crossings);

typedef struct psim_observables {
    double temperature;       /* sum v^2 / 3 / N                             */
    double mean_disp2;        /* sum |x - x0|^2 / N                          */
    double msd;               /* sum |x + L*crossings - x_start|^2 / N (FP64);
                                 NaN before psim_gpu_track_start            */
    double mean_coordination; /* contacts / N                                */
    long long contacts;       /* ordered pairs with r < contact_r, over the
                                 27 cells of the force grid                 */
    double time;              /* current time                                */
    unsigned long long step;  /* step index                                  */
} psim_observables;

/* contact_r 0 skips the contact pass (coordination 0); the reference uses
 * 1.2.  coord_by_orig (may be NULL) receives each particle's contacts. */
int psim_gpu_observables(psim_ctx *ctx, double contact_r, psim_observables *out,
                         int *coord_by_orig);

/* Connected components of the contact graph (r < contact_r).  labels_by_orig
 * (may be NULL) receives the smallest original index of each particle's
 * component; n_clusters and mean_size count components of at least
 * min_size members (the reference counts more than 4), mean by integer
 * division. */
int psim_gpu_clusters(psim_ctx *ctx, double contact_r, int min_size, int *labels_by_orig,
                      int *n_clusters, int *mean_size);

/* Pair-distance histogram over ordered pairs i != j (minimum image), bins
 * [r_lo + b*width, r_lo + (b+1)*width), b < nbins (<= 2048).  sigma == 0:
 * counts; sigma > 0: each pair adds the mass of a normal distribution
 * N(r_ij, sigma^2) that falls in each bin (support cut at 8 sigma).  The
 * reach r_lo + nbins*width + 8*sigma must not exceed half the box.  No
 * reference counterpart (the structure observable of the fitting loss). */
int psim_gpu_pair_histogram(psim_ctx *ctx, double r_lo, double width, int nbins,
                            double sigma, double *hist);
/* grad (AoS3, original order) = d(sum_b w[b] * hist[b]) / d positions,
 * sigma > 0. */
int psim_gpu_pair_histogram_grad(psim_ctx *ctx, double r_lo, double width, int nbins,
                                 double sigma, const double *w, double *grad);

/* ---- trajectory adjoint -------------------------------------------------------
 * Reverse-mode derivative of a trajectory of nsteps G&B steps (in-kernel
 * Philox noise, so the noise is fixed by the seed and step index) with
 * respect to the pair-force parameters.  No reference counterpart: the
 * reference has no derivatives; the forward steps are psim_gpu_step's.
 * By default the derivative is pathwise: the set of pairs inside the cutoff
 * is the one the forward trajectory used.  dtheta[PSIM_DP_*] receives
 * dL/dparameter. */
enum {
    PSIM_DP_AO_WELL = 0,      /* AO wellDepth (polymer fraction phi_p)       */
    PSIM_DP_AO_CUTOFF = 1,    /* cutOff through the AO coefficients (and, with
                                 a cutoff window, through the pairs that cross) */
    PSIM_DP_LJ_WELL = 2,      /* LJ wellDepth (epsilon)                      */
    PSIM_DP_YUKAWA = 3,       /* yukawaStrength                              */
    PSIM_DP_DEBYE = 4,        /* debyeLength                                 */
    PSIM_DP_COUNT = 5
};
typedef struct psim_adjoint psim_adjoint;

/* Takes the context's current state as the trajectory's start; keeps a
 * checkpoint every checkpoint_every steps (device memory about
 * (nsteps/checkpoint_every + 1) * 16 N doubles + checkpoint_every * 4 N). */
int psim_adj_create(psim_ctx *ctx, int nsteps, int checkpoint_every, psim_adjoint **out);
void psim_adj_destroy(psim_adjoint *adj);
/* Cutoff window w >= 0 (default 0, pathwise).  The pair force jumps at the
 * cutoff (fNet(cutOff) != 0), so in a large system every change of cutOff
 * moves pairs across it and the loss's secant slope holds those jumps.  With
 * w > 0, dtheta[PSIM_DP_AO_CUTOFF] also receives the boundary term
 * sum over steps and ordered pairs with |r - cutOff| < w of
 * lambda_i . fNet(r) d_ij / r / (2 w): the derivative linearised over
 * [cutOff - w, cutOff + w], comparable with a central difference of that
 * half-width.  cutOff + w must not exceed the cell edge (else backward
 * returns PSIM_E_ARG). */
int psim_adj_set_cutoff_window(psim_adjoint *adj, double w);
/* Record-all mode (on != 0): the forward pass records every state of the
 * trajectory (positions, ids, cell table: about (32 N + 4 ncell) bytes per
 * step, so K (32 N + 4 ncell) in all), and the backward pass recomputes
 * nothing; off (the default) keeps checkpoint_every-step segments that the
 * backward pass recomputes.  The gradients are bit-identical either way.
 * Takes effect at the next psim_adj_forward; PSIM_E_NOMEM (the object stays
 * in checkpoint mode) when the records do not fit. */
int psim_adj_set_record_all(psim_adjoint *adj, int on);
/* Re-take the start state from the context's current state. */
int psim_adj_set_initial(psim_adjoint *adj);
/* Run the trajectory from the start state with the context's current
 * parameters; the context is left at step nsteps.  Returns 0 or a physics
 * error code. */
int psim_adj_forward(psim_adjoint *adj);
/* gx = dL/dx at the last step (AoS3, original order).  Writes
 * dtheta[PSIM_DP_COUNT] and, if gx0 is not NULL, dL/dx at the start.  The
 * force parameters must be those of the last psim_adj_forward; the context
 * is left at step nsteps. */
int psim_adj_backward(psim_adjoint *adj, const double *gx, double *dtheta, double *gx0);

/* The reference's own initial configuration: n particles of radius r in a
 * periodic cube of integer edge box, std::mt19937(seed) with
 * std::uniform_real_distribution<double>(0, 1) draws in the reference's
 * order (seed cube of seed_size^3 particles at the centre, uniform
 * placement, overlap re-draws while the squared distance is below 2r, then
 * Maxwell velocities at the temperature truncated to an integer).  pos,
 * pos0 (the setters' previous-position image) and vel are AoS3; pos0 and
 * vel may be NULL.  Returns 0, PSIM_ERR_INIT (more than 10 n re-draws) or
 * PSIM_E_ARG.  Host only. */
int psim_init_reference(int n, int box, double radius, unsigned int seed, int seed_size,
                        double temp, double *pos, double *pos0, double *vel);

/* Seeded random sequential addition of n spheres with pair distance >= dmin
 * in a periodic cube [0, box)^3 (host, O(n)).  Returns 0 or PSIM_ERR_INIT. */
int psim_init_rsa(int n, double box, double dmin, unsigned long long seed,
                  double *pos_out);

/* ---- snapshots (host) ------------------------------------------------------ */
/* analysisManager::writeSystem (analysisOutput.cpp:55-74): one line per
 * particle, "x y z x0 y0 z0 fx fy fz fx0 fy0 fz0 m r", each number in the C++
 * stream's default format (6 significant digits), no newline after the last
 * line.  Arrays AoS3 in original order; frc / frc0 may be NULL (zeros). */
int psim_write_recovery(const char *path, int n, const double *pos, const double *pos0,
                        const double *frc, const double *frc0, double mass, double radius);
/* system::particlesInFile + readParticles (systemRecovery.cpp:30-81): each
 * line moves a fresh particle to its previous position and then to its
 * current one through the particle setters (single wrap, previous position
 * re-imaged next to the current one), so pos0 is the reference's; mass and
 * radius are read as float.  The file must hold n lines (else PSIM_E_IO); a
 * position outside [0, box) after the wrap gives PSIM_ERR_BOUNDS.  frc, frc0,
 * mass[n], radius[n] may be NULL. */
int psim_read_recovery(const char *path, int n, double box, double *pos, double *pos0,
                       double *frc, double *frc0, double *mass, double *radius);
/* analysisManager::writeSystemXYZ (analysisOutput.cpp:31-53): "N", "Current
 * Time: <rounded time>", then "H x y z" per particle. */
int psim_write_xyz(const char *path, int n, const double *pos, double current_time);
/* system::writeSystemInit's sysConfig.cfg (system.cpp:108-127): the
 * run's geometry and settings as "key = value" lines. */
int psim_write_sys_config(const char *path, const char *trial_name, int n,
                          double concentration, int box, int cell_size, int cell_scale,
                          double temp, double dt, int output_freq, double cycle_hour, int seed);
/* Full-precision state of a context (positions, previous positions,
 * velocities, last forces, integrator memory, time, velocity counter, step
 * index, the particles' slot order, which fixes the force summation order,
 * and, once started, the displacement tracker), so that a resumed run
 * continues bit for bit.  No reference counterpart (the reference's
 * recovery file drops velocities, memory, time and tracker).  load needs a
 * context created with the same particle count and box (else PSIM_E_ARG). */
int psim_gpu_save(psim_ctx *ctx, const char *path);
int psim_gpu_load(psim_ctx *ctx, const char *path);

#ifdef __cplusplus
}
#endif
#endif
