/*
 * mph.h -- C ABI of the MI355X-native MPH-WC explicit solver hot path.
 *
 * The reference (source/main.cpp of MphExplicit) has no library API: its "boundary" is the
 * sequence of static calls inside main() (source/main.cpp:357-452) over file-scope globals
 * (source/main.cpp:56-156), fed by a .data keyword file and a .grid particle file.  This
 * header is the drop-in for that loop: the host keeps the reference's formats and the
 * per-step call order, and every call below replaces a named piece of main.cpp.
 *
 * Conventions
 *   - plain C types only; vectors are [n][3] double (the reference's storage, main.cpp:78-80)
 *   - every function returns an int status (MPH_OK == 0, negative on error) unless noted;
 *     mph_last_error() gives a message for the calling thread.  The library never exits
 *     the process (the reference exits on malloc/fopen failure, source/errorfunc.cpp:8-31).
 *   - particle types 0,1 are fluid and 2,3 wall (main.cpp:43-47).
 *   - one context = one device stream; calls on one context are not thread safe.
 */
#ifndef MPH_H_INCLUDED
#define MPH_H_INCLUDED

#ifdef __cplusplus
extern "C" {
#endif

#define MPH_ABI_VERSION 1

enum mph_status {
    MPH_OK = 0,
    MPH_ERR_ARG = -1,       /* invalid argument / unsupported configuration */
    MPH_ERR_IO = -2,        /* file could not be opened */
    MPH_ERR_FORMAT = -3,    /* malformed .grid file */
    MPH_ERR_DEVICE = -4,    /* HIP runtime error or no usable GPU */
    MPH_ERR_CAPACITY = -5,  /* a neighbour list outgrew its reserved width and could not be widened */
    MPH_ERR_STATE = -6      /* call not valid in the context's current state */
};

enum mph_precision {
    MPH_FP64 = 0,           /* all pair math and sums in FP64 */
    MPH_MIXED = 1           /* FP32 pair geometry/force sums; FP64 positions, integration, density/pressure */
};

enum mph_order {
    MPH_ORDER_CURRENT = 0,  /* the solver's current (cell-sorted) particle order, as the reference stores it */
    MPH_ORDER_ORIGINAL = 1  /* by ParticleIndex = line order of the .grid file (main.cpp:645-647) */
};

/* Everything the reference reads from .data (main.cpp:502-521) and the .grid header
 * (main.cpp:549-557), plus the simulated dimension (compile-time TWO_DIMENSIONAL in the
 * reference, main.cpp:38) and the arithmetic precision. */
typedef struct mph_params {
    int dim;                          /* 3, or 2 (storage stays 3-component, main.cpp:40) */
    int precision;                    /* enum mph_precision */
    double particle_spacing;          /* l0 */
    double domain_min[3];
    double domain_max[3];
    double time;                      /* start time, .grid line 1 */
    double dt;
    double output_interval;
    double vtk_output_interval;
    double end_time;
    double radius_ratio_a;            /* also used for the gravity-centre kernel (main.cpp:838) */
    double radius_ratio_p;
    double radius_ratio_v;
    double density[4];
    double bulk_modulus[4];
    double bulk_viscosity[4];
    double shear_viscosity[4];
    double surface_tension[4];
    double interaction_ratio[4][4];   /* [type i][type j] */
    double gravity[3];
    double wall_center[4][3];         /* only entries 2,3 are used */
    double wall_velocity[4][3];
    double wall_omega[4][3];
} mph_params;

/* Host output buffers for mph_get_fields; any pointer may be NULL. */
typedef struct mph_fields {
    int *particle_index;              /* [n] original id (ParticleIndex, main.cpp:75) */
    int *type;                        /* [n] Property */
    double *position;                 /* [n][3] */
    double *velocity;                 /* [n][3] */
    double *force;                    /* [n][3] total force of the last step (main.cpp:80) */
    double *density_a;                /* [n] DensityA (main.cpp:131) */
    double *gravity_center;           /* [n][3] GravityCenter (main.cpp:132) */
    double *pressure_a;               /* [n] PressureA */
    double *vol_strain_p;             /* [n] VolStrainP */
    double *divergence_p;             /* [n] DivergenceP */
    double *pressure_p;               /* [n] PressureP */
    double *virial_pressure;          /* [n] VirialPressureAtParticle (valid after mph_compute_virial) */
    int *neighbor_count;              /* [n] NeighborCount: list length incl. self (main.cpp:81) */
} mph_fields;

typedef struct mph_stats {
    long long steps;                  /* steps taken since create */
    long long rebuilds;               /* neighbour-list rebuilds (incl. the initial one) */
    long long list_entries;           /* entries of all cluster lists at the last rebuild */
    int list_width;                   /* reserved entries per cluster list */
    int max_list_len;                 /* longest cluster list since the lists were sized */
    int cluster_size;                 /* particles per cluster tile, picked by the autotuner */
    int clusters;                     /* clusters at the last rebuild */
    long long eval_slots;             /* (member, list entry) evaluations of one pair sweep over
                                         the current lists, padding and non-interacting entries
                                         included */
    int list_grid;                    /* largest grid of the list build (workgroups), picked by
                                         the autotuner for a single case */
} mph_stats;

/* Per-phase device time of mph_step_timed, summed over its steps. */
typedef struct mph_timing {
    double move_ms;                   /* convection, walls, wrap, Verlet check */
    double rebuild_ms;                /* bucketing + lists (near zero on steps without rebuild) */
    double sweep_a_ms;                /* densities, divergence, pressures */
    double sweep_b_ms;                /* forces and velocity update */
    long long pairs;                  /* sum over steps of ordered pairs i != j closer than the largest kernel radius */
    long long rebuilds;               /* rebuilds during these steps (max over cases) */
} mph_timing;

typedef struct mph_ctx mph_ctx;

/* ---- host-side formats (no device needed) ------------------------------------------ */

/* Reference defaults of the .data statics (main.cpp:56-211): zeros, Dt = 1e100. */
void mph_params_default(mph_params *p);

/* readDataFile, main.cpp:489-539.  Unknown lines are ignored (the reference logs them);
 * CRLF line ends, '#' comments and leading blanks are accepted. */
int mph_read_data_file(const char *path, mph_params *p);

/* readGridFile header, main.cpp:549-557: Time, count, spacing, domain. */
int mph_read_grid_header(const char *path, mph_params *p, int *count);

/* readGridFile body, main.cpp:632-639: type, position, velocity of `count` particles. */
int mph_read_grid_particles(const char *path, int count, int *type, double *x, double *v);

/* writeProfFile, main.cpp:672-696 (the .grid layout with %e). */
int mph_write_prof_file(const char *path, const mph_params *p, double time, int n,
                        const int *type, const double *x, const double *v);

/* writeVtkFile, main.cpp:698-834 (legacy ASCII VTK, float casts, same field order). */
int mph_write_vtk_file(const char *path, int n, const int *type, const double *x,
                       const double *v, const double *density_a, const double *vol_strain_p,
                       const double *virial_pressure, const int *neighbor_count,
                       const double *force);

/* ---- device path -------------------------------------------------------------------- */

/* Number of visible GPUs (0 when none); never fails. */
int mph_device_count(void);

/* Upload one case and run the reference's set-up: initializeWeight/Fluid/Wall/Domain
 * (main.cpp:330-333), then the first bucketing, neighbour build and densities
 * (main.cpp:358-364).  Returns NULL on error (see mph_last_error). */
mph_ctx *mph_create(const mph_params *p, int n, const int *type, const double *x,
                    const double *v);

/* Same for `ncases` independent cases run side by side in every launch (no reference
 * counterpart: the reference runs one case per process).  Particle arrays are the cases'
 * arrays concatenated; counts[c] particles belong to case c. */
mph_ctx *mph_create_batch(int ncases, const mph_params *p, const int *counts, const int *type,
                          const double *x, const double *v);

/* `nsteps` iterations of the reference's loop body (main.cpp:386-431) for every case:
 * convection, walls, periodic wrap, Verlet check + rebuild, densities, pressures,
 * surface tension, viscosity, gravity, acceleration; Time += Dt (main.cpp:449).
 * Asynchronous on the context's stream except for the read of the list-overflow word at the
 * end; a rebuild that outgrew the reserved list width is widened and finished inside the call
 * (the reference sizes NeighborInd at every rebuild, main.cpp:1487-1495). */
int mph_step(mph_ctx *ctx, int nsteps);

/* Capture one step into a HIP graph and replay it for `nsteps` (same results as mph_step). */
int mph_step_graph(mph_ctx *ctx, int nsteps);

/* mph_step with HIP events around each phase on the context's stream (no reference
 * counterpart: the reference only keeps process CPU-time buckets, main.cpp:383-451).
 * Synchronises once at the end. */
int mph_step_timed(mph_ctx *ctx, int nsteps, mph_timing *out);

/* Wait for all queued work of the context. */
int mph_sync(mph_ctx *ctx);

/* Number of particles of case `case_id` (or of all cases for case_id = -1). */
int mph_particle_count(const mph_ctx *ctx, int case_id);

/* Simulated time of a case. */
int mph_get_time(const mph_ctx *ctx, int case_id, double *time);

/* Copy per-particle fields of a case to the host (synchronises).  Replaces the arrays that
 * main.cpp:672-834 reads for .prof/.vtk output.  Where no case has surface tension the step
 * does not sum DensityA (it enters no force then); this call sums it first, on the current
 * positions and lists, which gives the values the step would have produced. */
int mph_get_fields(mph_ctx *ctx, int case_id, int order, mph_fields *out);

/* calculateVirialStressAtParticle, main.cpp:2057-2302, for the current state (the
 * reference calls it before each .vtk write, main.cpp:436-437).  Read the result with
 * mph_get_fields(...virial_pressure...). */
int mph_compute_virial(mph_ctx *ctx);

/* Neighbour lists of the last rebuild in CSR form keyed by original particle id:
 * ptr[n+1], ind[total] with each row ascending and self included, i.e. the reference's
 * NeighborPtr/NeighborInd (main.cpp:1329-1596) mapped through ParticleIndex.  The rows
 * are recovered on the host from the cluster lists with the reference's exact distance
 * test at the rebuild positions.  Call with ind == NULL to get only ptr (ptr[n] = total). */
int mph_get_neighbors(mph_ctx *ctx, int case_id, long long *ptr, int *ind);

/* mph_get_neighbors into device memory, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture elsewhere in the
 * process, serving loops, harnesses that forbid hipMalloc).  ptr: device array of n+1
 * (n = mph_particle_count(ctx, case_id)); ind: device array of ind_capacity entries, or NULL
 * for the row starts only; *total (may be NULL) receives ptr[n].  Same rows as
 * mph_get_neighbors (original ids, each row ascending, self included).  The library's own
 * scratch is a 64 KiB array per GPU: cases of up to 8160 * 4096 (33.4M) particles; larger
 * ones are rejected with MPH_ERR_ARG, and so is an ind too small for the lists (the needed
 * size is reported in *total).  Calls on one GPU share that scratch: calls that overlap
 * (from several threads or streams) race.  Synchronises the context's stream. */
int mph_get_neighbors_noalloc(mph_ctx *ctx, int case_id, long long *ptr, int *ind, long long ind_capacity,
                              long long *total);

int mph_get_stats(const mph_ctx *ctx, mph_stats *out);

void mph_destroy(mph_ctx *ctx);

/* Change the time step of one case (the reference reads Dt once from the .data file,
 * main.cpp:502, and keeps it for the run).  Takes effect at the next step; captured
 * graphs stay valid (they read the case constants from device memory). */
int mph_set_time_step(mph_ctx *ctx, int case_id, double dt);

/* Message of the last failed call on this thread ("" if none). */
const char *mph_last_error(void);

/* ------------------------------------------------------------------------------------------
 * Learned part: a pairwise-message surrogate.  No counterpart in the reference.
 *
 * For each fluid particle i and each neighbour j within the largest interaction radius h
 * (the reference's neighbour set, main.cpp:1408), a 17-input MLP (three hidden layers of 64
 * SiLU units, 3 outputs c0..c2) gives the coefficients of an antisymmetric message
 *   m_ij = c0 e + c1 u + c2 (u.e) e,   e = x_ij / r,   u = (v_j - v_i) / u_scale,
 * and the predicted mean acceleration over two steps is a_i = a_scale (rho_0/rho_i) sum_j m_ij.
 * All inputs are symmetric under i <-> j, so m_ji = -m_ij: pairwise momentum exchange is
 * built in.  Weights are BF16 on the matrix cores with FP32 accumulation; master weights and
 * Adam state are FP32.
 *
 * Inputs (x_ij = periodic x_j - x_i, r = |x_ij|, eps = dens_scale e^P / N0p with the
 * volumetric strain VolStrainP = sum w_P - N0p, main.cpp:1850-1856):
 *   0 r/h   1 (r/h)^2   2 u.e   3 |u|^2   4-9 one-hot of the unordered class pair
 *   (00, 01, 02, 11, 12, 22; class = type 0, type 1, wall)   10 eps_i + eps_j
 *   11 (eps_i - eps_j)^2   12 eps_i eps_j   13-15 zero   16 the constant 1 (bias of layer 1)
 *
 * A surrogate step advances 2 dt: the plain move to x_{n+1}, sweep A's densities at x_{n+1},
 * the MLP at (x_{n+1}, v_n), v_{n+1} = v_n + dt (a + g), the plain move to x_{n+2} and
 * v_{n+2} = v_{n+1} + dt (a + g): two plain steps whose force evaluations are replaced by
 * one prediction.  Training targets come from the plain trajectory:
 * a = (v_{n+2} - v_n) / (2 dt) - g at the state (x_{n+1}, v_n).
 *
 * Parameter layout (MPH_SUR_NPARAMS floats, row-major, [in][out]):
 *   W1[17][64] (row 16 = bias of layer 1), W2[64][64], b2[64], W3[64][64], b3[64],
 *   W4[64][3], b4[3].
 * ---------------------------------------------------------------------------------------- */

#define MPH_SUR_NPARAMS 9603

typedef struct mph_surrogate mph_surrogate;

typedef struct mph_sur_params {
    double u_scale;          /* velocity scale of the relative-velocity features [m/s] */
    double a_scale;          /* acceleration scale of the messages [m/s^2] */
    double dens_scale;       /* factor on the density deviation e^P / N0p */
    double lr, beta1, beta2, eps;   /* Adam */
    int train_stride;        /* training uses fluid particles with sorted index % stride == 0 */
    int pad0;
    unsigned long long seed; /* weight initialisation */
    double radius_ratio;     /* message radius / particle spacing; 0: the largest kernel radius
                                (the neighbour set of main.cpp:1408); at most that radius */
} mph_sur_params;

typedef struct mph_sur_stats {
    long long blocks;        /* judged blocks */
    long long rollbacks;     /* blocks replayed with plain steps */
    long long sur_steps;     /* accepted surrogate steps (2 dt each) */
    long long plain_steps;   /* plain steps (judges and replays) */
    double max_err_v;        /* largest relative L2 velocity error seen by a judge */
    double max_err_n;        /* largest relative L2 number-density error seen by a judge */
    double last_err_v, last_err_n;
    double sim_time;         /* simulated time advanced by the call */
} mph_sur_stats;

void mph_sur_params_default(mph_sur_params *p);

/* Create a surrogate bound to a single-case FP64 context with static walls.  The weights are
 * initialised from p->seed.  Returns NULL on error (mph_last_error). */
mph_surrogate *mph_sur_create(mph_ctx *ctx, const mph_sur_params *p);

int mph_sur_get_weights(mph_surrogate *s, float *w);        /* MPH_SUR_NPARAMS floats */
int mph_sur_set_weights(mph_surrogate *s, const float *w);

/* Acceleration the MLP predicts at the context's current state (positions, velocities and
 * VolStrainP as mph_get_fields reports them), by original particle id, [n][3]; 0 for walls. */
int mph_sur_predict(mph_surrogate *s, double *accel);

/* Loss and gradient at the context's current state (as mph_sur_predict) for the given
 * per-particle targets [n][3] by original id, over the training particles:
 * loss = mean over those fluid particles of |a - target|^2 / a_scale^2.
 * grad: MPH_SUR_NPARAMS floats (may be NULL). */
int mph_sur_gradient(mph_surrogate *s, const double *target, double *loss, float *grad);

/* Train on the context's own plain trajectory: each sample advances the context by two
 * plain steps and then takes `iters` Adam steps on it.  losses: samples*iters values
 * (may be NULL). */
int mph_sur_train(mph_surrogate *s, int samples, int iters, double *losses);

/* Surrogate steps (2 dt each) without a judge. */
int mph_sur_step(mph_surrogate *s, int nsteps);

/* Judged run: `blocks` blocks of k surrogate steps.  After each block one surrogate step
 * and two plain steps start from the same state; if the relative L2 difference of the fluid
 * velocities or number densities exceeds tol, the block is rolled back and replayed with
 * 2k plain steps; otherwise the plain result is kept.  Accumulates into *st. */
int mph_sur_advance(mph_surrogate *s, int blocks, int k, double tol, mph_sur_stats *st);

void mph_sur_destroy(mph_surrogate *s);

#ifdef __cplusplus
}
#endif

#endif /* MPH_H_INCLUDED */
