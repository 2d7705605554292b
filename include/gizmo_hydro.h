/*
 * gizmo_hydro.h -- C ABI of the MI355X-native MFM hydro neighbour-loop pipeline.
 *
 * The library replaces the three hot-path entry points of GIZMO's timestep loop,
 *   void density(void);              src/hydro/density_loop.cc:764
 *   void hydro_gradient_calc(void);  src/hydro/gradients_loop.cc:458
 *   void hydro_force(void);          src/hydro/hydro_force_loop.cc:423
 * which the reference calls from compute_hydro_densities_and_forces()
 * (src/core/accel.cc:80,143,169) and which read and write the P[]/CellP[] particle
 * fields and the All parameters.  Here the particle state lives on the GPU inside a
 * context; gz_load()/gz_store() move the P[]/CellP[] hot-path fields in and out (in the
 * caller's particle order), and gz_density()/gz_hydro_gradient_calc()/gz_hydro_force()
 * run the three stages on the resident state.  Plain pointers and sizes only.
 *
 * Error convention: every entry point returns 0 on success, or a nonzero code
 * (GZ_E_* below, or the reference's own stop codes where a stage hits one: 1234 when no
 * Riemann solution is found (src/hydro/hydro_core_meshless_functions.h:424), 1155 when the
 * kernel-radius iteration does not converge in MAXITER rounds, 90001009 for an empty
 * radius bracket (src/hydro/density_loop.h:893)).  gz_last_error() gives the message.
 * The reference maps such codes to endrun(); a drop-in shim does the same.
 */
#ifndef GIZMO_HYDRO_H
#define GIZMO_HYDRO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GZ_OK 0
#define GZ_E_ARG 1          /* bad argument (null pointer, size, parameter out of range) */
#define GZ_E_HIP 2          /* a HIP runtime call failed */
#define GZ_E_STATE 3        /* call out of order (e.g. gradients before density) */
#define GZ_E_MEMORY 4       /* device allocation failed */
#define GZ_STOP_RIEMANN 1234
#define GZ_STOP_MAXITER 1155
#define GZ_STOP_BRACKET 90001009

/* arithmetic forms of the pair physics */
#define GZ_FP64 0           /* plain FP64, the reference's own precision (MyDouble = double) */
#define GZ_FP32_MIXED 1     /* FP32 pair math on relative coordinates; FP64 positions, 3x3 inverse, h test */

/* Run parameters: the All fields the hot path reads (src/declarations/global_data_all_struct.h)
   plus the compile-time choices of the reference build (KERNEL_FUNCTION, GAMMA_DEFAULT). */
typedef struct {
    double box_size;          /* All.BoxSize (periodic cube) */
    int kernel;               /* KERNEL_FUNCTION: 3 cubic spline, 6 Wendland C2 */
    double gamma;             /* GAMMA_DEFAULT (adiabatic index, > 1) */
    double des_num_ngb;       /* All.DesNumNgb */
    double max_num_ngb_dev;   /* All.MaxNumNgbDeviation */
    double min_kernel_radius; /* All.MinKernelRadius */
    double max_kernel_radius; /* All.MaxKernelRadius */
    double time;              /* All.Time */
    double time_begin;        /* All.TimeBegin */
    int precision;            /* GZ_FP64 or GZ_FP32_MIXED */
} gz_params;

/* Host arrays of the P[]/CellP[] hot-path fields, n entries each in the caller's order;
   vectors are interleaved (n x 3, n x 6, n x 9).  A NULL pointer means "not provided"
   (gz_load) or "not wanted" (gz_store).  Names follow the reference fields. */
typedef struct {
    double *pos;       /* P.Pos[3] */
    double *vel;       /* CellP.VelPred[3] */
    double *mass;      /* P.Mass */
    double *h;         /* P.KernelRadius */
    double *u;         /* CellP.InternalEnergyPred */
    double *numngb;    /* P.NumNgb (stored as its cube root, as the reference does) */
    double *drkern;    /* P.DrkernNgbFactor */
    double *divvel;    /* P.Particle_DivVel */
    double *rho;       /* CellP.Density */
    double *press;     /* CellP.Pressure */
    double *nvt;       /* CellP.NV_T (inverted, [xx,yy,zz,xy,yz,xz]) */
    double *nvt_fw;    /* CellP.NV_T_face_weights[3] */
    double *cn;        /* CellP.ConditionNumber */
    double *fce;       /* CellP.FaceClosureError */
    double *grad_rho;  /* CellP.Gradients.Density[3] */
    double *grad_p;    /* CellP.Gradients.Pressure[3] */
    double *grad_v;    /* CellP.Gradients.Velocity[3][3] ([component][direction]) */
    double *accel;     /* CellP.HydroAccel[3] (dv/dt after hydro_force) */
    double *dtu;       /* CellP.DtInternalEnergy (du/dt after hydro_force) */
    double *maxsig;    /* CellP.MaxSignalVel */
    int *timebin;      /* P.TimeBin */
    int *wakeup;       /* P.wakeup */
    /* timestep-loop fields (gz_run_*): a drop-in that keeps its own loop leaves them NULL */
    double *vel_kick;  /* P.Vel[3] (the kicked velocity; vel above is the predicted one) */
    double *u_kick;    /* CellP.InternalEnergy (u above is InternalEnergyPred) */
    double *soft;      /* P.ForceSoftening (kernel radius cached for the acceleration criterion) */
    long long *ti_begstep; /* P.Ti_begstep */
    long long *dt_step;    /* P.dt_step */
    long long *ti_current; /* P.Ti_current (the time the particle's predicted fields stand at: the run
                              loop drifts a particle only when it is active or read, see gz_drift_all) */
} gz_fields;

typedef struct gz_context gz_context;

/* Create a context on HIP device `device` for up to `capacity` particles. NULL on failure. */
gz_context *gz_create(const gz_params *params, long capacity, int device);
void gz_destroy(gz_context *ctx);
const char *gz_last_error(const gz_context *ctx);

/* Replace the run parameters (All.Time changes every step; the rest normally not). */
int gz_set_params(gz_context *ctx, const gz_params *params);

/* Host -> device: n particles.  pos, vel, mass, u and h are required; the other fields are
   optional (the state an inactive particle keeps between its active steps).  The library
   orders the particles along a space-filling curve on the device. */
int gz_load(gz_context *ctx, long n, const gz_fields *host);
/* Device -> host, in the caller's order, for every non-NULL pointer of `host`. */
int gz_store(gz_context *ctx, gz_fields *host);

/* The active set (ActiveParticleList): one byte per particle in the caller's order, nonzero =
   active; NULL = every particle active. */
int gz_set_active(gz_context *ctx, const unsigned char *active);

/* The three stages (density -> gradients -> force must run in that order after a gz_load or
   after positions change). */
int gz_density(gz_context *ctx);              /* replaces density(), src/hydro/density_loop.cc:764 */
int gz_hydro_gradient_calc(gz_context *ctx);  /* replaces hydro_gradient_calc(), src/hydro/gradients_loop.cc:458 */
int gz_hydro_force(gz_context *ctx);          /* replaces hydro_force(), src/hydro/hydro_force_loop.cc:423 */

/* Wait for the context's stream. */
int gz_synchronize(gz_context *ctx);

/* Diagnostics.  gz_permutation: perm[k] = caller index of the k-th particle in device order.
   gz_stat: named counters of the last calls ("density_iterations", "pairs_density",
   "pairs_symmetric", "ngb_rebuilds", "n_active", "ell_widen"), and totals since the last
   gz_timing_reset ("active_updates": active particles through the hot path, "resorts",
   "force_pairs_total": neighbour pairs of every force launch, "force_rare_rows": rows of
   lane-per-row force passes that a face past the Riemann ladder's common case sent to the
   pass's second kernel).  With the option
   "riemann_stats" set to 1, the force passes of a single box also count, per face, the rungs of
   the Riemann ladder it took (each at most once per face) since the option was last set:
   "riemann_faces" (faces that reached the Riemann solve), "riemann_hllc", "riemann_vacuum",
   "riemann_roe", "riemann_pvrs" (the estimate get_wavespeeds_and_pressure_star took,
   src/hydro/reimann.h:485-549), "riemann_kt" (Kurganov-Tadmor, :445-477), "riemann_exact" and
   "riemann_exact_vacuum" (Riemann_solver_exact, :625-668), "riemann_unphysical" (:284-290),
   "riemann_retry_first_order", "riemann_retry_zero_velocity", "riemann_fail" (the pair body's
   retries, src/hydro/hydro_core_meshless_functions.h:351-427), "entropic_considered" and
   "entropic_applied" (the face entropic switch, :515-573). */
int gz_permutation(gz_context *ctx, long *perm);
int gz_stat(gz_context *ctx, const char *name, double *value);

/* Kernel timing on the context's stream (HIP events around each launch).  gz_timing
   returns the launch count and total milliseconds of the named kernel family since the last
   gz_timing_reset ("ngb_build", "density", "gradients", "force", "sort"; in the timestep loop
   also "drift" -- the drift of every particle -- and "timeline" -- timesteps, bins, wake-ups,
   kicks and the active list). */
int gz_timing_enable(gz_context *ctx, int on);
int gz_timing_reset(gz_context *ctx);
int gz_timing(gz_context *ctx, const char *name, long *launches, double *total_ms);

/* ---- the timestep loop around the hot path (src/core/run.cc:220-507) ----
   A gas-only, gravity-free, non-comoving run on the resident state: power-of-two time bins on
   the integer timeline (TIMEBINS = 60, src/declarations/typedefs.h:8-9), kick-drift-kick with
   the reference's per-particle steps, wake-ups, and the reference's lazy drift: at a sync point
   only the particles of the active bins are drifted (src/core/run.cc:979-994,
   src/core/predict.cc:175-210), every other particle when a pass is about to read it
   (src/mesh/gpu_neighbor_list.cc:2443-2510: here the particles of the grid cells around the
   active rows), and every particle at a step whose active bins hold more than
   domain_build_active_fraction of them (src/core/run.cc:360,380-390) and before the particles
   are ordered again.  The state stays on the device between steps; gz_store reads it back at any
   time (P.Ti_current tells where each particle's predicted fields stand). */
#define GZ_TIMEBINS 60
typedef struct {
    double courant_fac;          /* All.CourantFac */
    double err_tol_int_accuracy; /* All.ErrTolIntAccuracy */
    double max_size_timestep;    /* All.MaxSizeTimestep */
    double min_size_timestep;    /* All.MinSizeTimestep */
    double time_begin;           /* All.TimeBegin */
    double time_max;             /* All.TimeMax (sets All.Timebase_interval) */
    double min_egy_spec;         /* All.MinEgySpec (0 = no floor) */
    double domain_build_active_fraction; /* All.DomainBuild_ActiveFraction: a step whose active bins hold more
                                            than this fraction of the particles drifts every particle
                                            (src/core/run.cc:360,380-390); 0 = every step */
} gz_integ_params;

typedef struct {
    long long ti_current;              /* All.Ti_Current */
    double time;                       /* All.Time */
    double timebase_interval;          /* All.Timebase_interval */
    long long num_steps;               /* All.NumCurrentTiStep */
    long long n_active;                /* active particles of the last step (hot-path updates) */
    long long n_woken;                 /* particles the last wake-up pass moved to a shorter step */
    long long bin_count[GZ_TIMEBINS];  /* TimeBinCount */
    unsigned long long bin_active;     /* TimeBinActive as a bit mask */
    int stop_code;                     /* first controlled-stop code raised, else 0 */
    int full_drift;                    /* 1: the last step drifted every particle; 0: only the active ones
                                          and the particles its passes read */
} gz_timeline;

/* init() (src/core/init.cc:108-113,583,1051,1061) and run()'s first
   compute_hydro_densities_and_forces() (src/core/run.cc:128-150) on the loaded state: every
   particle in bin 0 and active, predicted fields = current ones (vel_kick/u_kick are taken
   from vel/u when gz_load was not given them), density, softening cache, then
   density -> gradients -> force. */
int gz_run_begin(gz_context *ctx, const gz_integ_params *ip);
/* One iteration of run()'s main loop (src/core/run.cc:220-507): find_timesteps() with
   process_wake_ups() (src/core/timestep.cc:39,1565), do_first_halfstep_kick()
   (src/core/kicks.cc:34), find_next_sync_point_and_drift() (src/core/run.cc:857),
   make_list_of_active_particles() (:997), compute_hydro_densities_and_forces()
   (src/core/accel.cc:68) and do_second_halfstep_kick() (src/core/kicks.cc:99). */
int gz_run_step(gz_context *ctx);
/* nsteps iterations of that loop in one call, as run()'s while loop takes them (src/core/run.cc:220-507,
   one sync point per pass), with no return to the caller between them; n_active[k] and n_woken[k]
   (either may be NULL) receive step k's active-particle count and woken count.  Stops at the
   first step that fails and returns its code; the steps before it stand (gz_timeline_get tells
   how many were taken). */
int gz_run_steps(gz_context *ctx, int nsteps, long long *n_active, long long *n_woken);
int gz_timeline_get(gz_context *ctx, gz_timeline *out);
/* gizmo_full_drift_to(All.Ti_Current) (src/core/predict.cc:148-171): every particle that stands
   behind the current sync point drifted to it, as the reference does before a snapshot
   (src/core/run.cc:911).  No effect outside a run. */
int gz_drift_all(gz_context *ctx);
/* Tuning knobs with no counterpart in the reference: "resort_fraction" (default 0.125) -- the
   particles are put back in space-filling-curve order once a bound on how far any of them has
   moved since they were last ordered (the sum of each step's longest drift along an axis)
   exceeds this fraction of the search radius (the search grid widens its cells by twice that
   bound meanwhile); "few_rows" (default 8192) -- passes over at most this
   many active particles run one workgroup per particle (its lanes share the particle's
   candidates or list entries) instead of one lane per particle; "wide_search" (default 0) --
   1: every particle searches at the grid level of its own kernel radius (cell tables of every
   level), for radii that span orders of magnitude, where one grid sized by the largest radius
   would walk far too many candidates (the symmetric list of more than few_rows rows is then
   built one wavefront per cell of the rows' own levels, the density rounds one workgroup per
   row); "row_schedule" (default 2048; 0 or a power of two from 64 to 2048) -- the
   one-lane-per-particle force pass takes the rows of each block of this many longest list
   first (a wavefront lasts as long as its longest list; 0: plain order); "speculate" (default
   1) -- a run-loop step over at most few_rows rows launches the symmetric list, the gradients
   and the force pass behind its first density round without waiting for that round's counters
   (the force kernels, whose wake-up flags and pair counts a second run would not undo, wait on
   a device flag: the round converged every row and no radius outgrew the list's grid); one
   read of the counters after the force pass then settles the step, and when the flag says no,
   the rounds go on and the three passes run again; "force_split" (default -1) -- the force pass
   over more than few_rows rows runs in two kernels, the first evaluating every face through the
   Riemann ladder's common case and leaving rows with a face that needs a later rung to the
   second (1), in one kernel with the whole ladder (0), or (-1) in two unless it is the FP32 form
   and more than a quarter of the rows took the second kernel in the last two-kernel pass;
   "force_waves" (default 0; 2 or 3) -- the FP64 form's first force kernel of those two runs at
   this many waves per SIMD (0: three with wide_search, else two).
   None of these changes a neighbour set or a
   sum, and none has a counterpart in the reference (tuning knobs of this library).
   One option selects physics, the reference's compile-time HYDRO_PARTITION_UNITY_IMPROVE_FD
   (src/hydro/density_loop.cc:622-633, src/hydro/density_loop.h:1281-1286): "fd_correction"
   (default 0) -- 1: the density pass also sums GradH_numer and GradH_denom, and the final
   density is divided by 1 + 4 alpha |grad H|^2 (grad H = -GradH_numer / GradH_denom, alpha =
   the kernel's KERNEL_AWPMHD_FD_ALPHA, src/mesh/kernel.h:185-206) before the pressure is
   taken from it. */
int gz_set_option(gz_context *ctx, const char *name, double value);

/* The Riemann functions of the force pass on n given face states, on the context's device and
   in its precision (a diagnostic for parity tests; the reference's counterparts are internal
   functions of src/hydro/reimann.h and the pair body).  Case k reads 24 doubles at in + 24k:
   left state (rho, p, v[3]), right state (rho, p, v[3]), face normal n[3], the pressure limiter,
   then the first-order states of the ladder's retries (p_i, p_j, rho_i, rho_j, v_i[3], v_j[3]);
   it writes 8 doubles at out + 8k: P_M, S_M, momentum flux[3], energy flux, 1 or 0 (0: the
   ladder found no positive pressure), 0; and rung[k] = the rung bits taken, in the order of the
   "riemann_*"/"entropic_*" stat names (bit 0 hllc ... bit 10 fail).  mode: 0 the ladder of the
   pair body with its two retries (src/hydro/hydro_core_meshless_functions.h:351-427),
   1 Riemann_solver (src/hydro/reimann.h:282-370), 2 get_wavespeeds_and_pressure_star (:485-549),
   3 Riemann_solver_KurganovTadmor_PWK (:445-477), 4 Riemann_solver_exact (:625-668). */
int gz_riemann_probe(gz_context *ctx, int mode, long n, const double *in, double *out, unsigned *rung);

/* Neighbour lists of the resident state, in the caller's indexing, for parity tests.
   symmetric = 0: rows r2 < h_i^2 (self included); 1: r2 < max(h_i,h_j)^2 (self excluded).
   Row order is ascending device order.  Call with nbr = NULL to get the total count. */
long gz_neighbors(gz_context *ctx, int symmetric, long *offsets, int *nbr, long cap);

/* One density() -> hydro_gradient_calc() -> hydro_force() pass over device buffers that
   allocates nothing.  No counterpart in the reference: it is for callers that cannot allocate
   device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  in: device pointers in the caller's order with gz_load's layouts (pos, vel, mass,
   u, h required; the optional inputs as gz_load takes them); out: device pointers with
   gz_store's layouts (NULL fields are skipped).  The results are those of gz_create + gz_load +
   the three stages + gz_store with the same parameters, bit for bit.  The pass runs on the
   device current for the calling thread, on `stream` (a hipStream_t; NULL: the null stream),
   and returns after the stream has finished it.  Every working buffer is carved from one static
   device array of GZ_NOALLOC_SCRATCH_BYTES, which each device the library runs on holds until
   the process ends; n is at most GZ_NOALLOC_MAX_N (a larger n returns GZ_E_ARG, a pass whose
   lists outgrow the array GZ_E_MEMORY).  Calls that overlap on one device race: they share its
   array.  err (may be NULL): receives the message of a failed call. */
#define GZ_NOALLOC_MAX_N 16384L
#define GZ_NOALLOC_SCRATCH_BYTES (48UL << 20)
int gz_hydro_pass_noalloc(const gz_params *params, long n, const gz_fields *in, gz_fields *out, void *stream, char *err,
                          size_t err_len);

#ifdef __cplusplus
}
#endif
#endif /* GIZMO_HYDRO_H */
