/*
 * gz_oracle.h -- interface of the CPU FP64 oracle (test infrastructure only; see gz_oracle.c).
 * Arrays are per particle, in the caller's index order; vectors are interleaved (n x 3).
 */
#ifndef GZ_ORACLE_H
#define GZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double box;               /* periodic cubic box side */
    int kernel;               /* 3 = cubic spline, 6 = Wendland C2 */
    double gamma;             /* adiabatic index */
    double des_num_ngb;       /* All.DesNumNgb */
    double max_num_ngb_dev;   /* All.MaxNumNgbDeviation */
    double min_kernel_radius; /* All.MinKernelRadius */
    double max_kernel_radius; /* All.MaxKernelRadius */
    double time, time_begin;  /* All.Time, All.TimeBegin */
    int fd_correction;        /* HYDRO_PARTITION_UNITY_IMPROVE_FD (src/hydro/density_loop.cc:622-633): 0 off, 1 on */
} orc_params;

typedef struct {
    long n;
    /* P[] fields */
    double *pos;      /* n*3 Pos */
    double *mass;     /* Mass */
    double *h;        /* KernelRadius */
    double *numngb;   /* NumNgb (cube root after density) */
    double *drkern;   /* DrkernNgbFactor */
    double *divvel;   /* Particle_DivVel */
    int *timebin;     /* TimeBin (may be NULL) */
    int *wakeup;      /* wakeup (may be NULL) */
    /* CellP[] fields */
    double *vel;      /* n*3 VelPred */
    double *u;        /* InternalEnergyPred */
    double *rho;      /* Density */
    double *press;    /* Pressure */
    double *nvt;      /* n*6 NV_T (inverted; xx,yy,zz,xy,yz,xz) */
    double *nvt_fw;   /* n*3 NV_T_face_weights */
    double *cn;       /* ConditionNumber */
    double *fce;      /* FaceClosureError */
    double *grad_rho; /* n*3 Gradients.Density */
    double *grad_p;   /* n*3 Gradients.Pressure */
    double *grad_v;   /* n*9 Gradients.Velocity[component][direction] */
    double *accel;    /* n*3 HydroAccel (dv/dt after force) */
    double *dtu;      /* DtInternalEnergy (du/dt after force) */
    double *maxsig;   /* MaxSignalVel */
    /* timestep-loop fields (gz_oracle_time.c; may be NULL for the hydro stages alone) */
    double *vel_kick;         /* n*3 P[].Vel (kicked velocity; vel above is VelPred) */
    double *u_kick;           /* CellP[].InternalEnergy (u above is InternalEnergyPred) */
    double *soft;             /* P[].ForceSoftening: the kernel radius cached for the acceleration step */
    long long *ti_begstep;    /* P[].Ti_begstep */
    long long *dt_step;       /* P[].dt_step */
    long long *ti_current;    /* P[].Ti_current: the time the particle's predicted fields stand at */
} orc_state;

#define ORC_TIMEBINS 60

typedef struct {              /* All.* parameters of the timestep loop */
    double courant_fac;          /* All.CourantFac */
    double err_tol_int_accuracy; /* All.ErrTolIntAccuracy */
    double max_size_timestep;    /* All.MaxSizeTimestep */
    double min_size_timestep;    /* All.MinSizeTimestep */
    double time_begin, time_max; /* All.TimeBegin, All.TimeMax */
    double min_egy_spec;         /* All.MinEgySpec */
    double domain_build_active_fraction; /* All.DomainBuild_ActiveFraction (full drift above it; 0: every step) */
} orc_integ;

typedef struct {              /* the integer timeline and the time bins */
    long long ti_current;                    /* All.Ti_Current */
    double time;                             /* All.Time */
    double timebase_interval;                /* All.Timebase_interval */
    long long num_steps;                     /* All.NumCurrentTiStep */
    long long bin_count[ORC_TIMEBINS];       /* TimeBinCount */
    unsigned char bin_active[ORC_TIMEBINS];  /* TimeBinActive */
    int need_wakeup;                         /* NeedToWakeupParticles_local */
    int stop_code;                           /* first controlled-stop code raised, else 0 */
    long long n_active;                      /* length of the active list of the last step */
    long long n_woken;                       /* particles moved to a shorter step by the last wake-up pass */
    int full_drift;                          /* 1: the last step drifted every particle */
} orc_timeline;

/* The timestep loop's lazy drift (gz_oracle_time.c) as the passes see it: every particle's
   position and kernel radius predicted at the current time, and the drift of one particle to it,
   called before a pass reads that particle's fields (thread-safe, once per particle and step).
   Set for the passes of a lazy step, NULL otherwise. */
typedef struct {
    const double *pos, *h;
    void (*ensure)(long j);
} orc_lazy;
void orc_set_lazy(const orc_lazy *view);

int orc_kernel(int kernel, double r, double h, double *wk, double *dwk);
double orc_fd_alpha(int kernel);
long orc_neighbors(const orc_params *par, long n, const double *pos, const double *h, int symmetric,
                   const unsigned char *rowmask, long *offsets, int *nbr, long cap);
double orc_matrix_invert(const double *T, double *Tinv);
int orc_density(const orc_params *par, orc_state *s, const unsigned char *active, int *iters_out);
int orc_density_fixed_h(const orc_params *par, orc_state *s, const unsigned char *active, double *raw);
void orc_slopelimiter(double *g, double valmax, double valmin, double alim, double h,
                      double shoot_tol, int pos_preserve, double d_max, double val_cen);
int orc_gradients(const orc_params *par, orc_state *s, const unsigned char *active);
void orc_reconstruct(double Q_i, const double *Gi, double Q_j, const double *Gj,
                     const double *di, const double *dj, double *Q_L, double *Q_R, int mode);
int orc_riemann(double gamma, const double *left, const double *right, const double *n,
                double ptl, double *P_M, double *S_M, double *fv, double *fp);
int orc_riemann_exact(double gamma, const double *left, const double *right, double *P_M, double *S_M);
int orc_riemann_probe(int mode, double gamma, long n, const double *in, double *out, unsigned *rung);
void orc_rung_counts(long long *out, int reset);
int orc_force(const orc_params *par, orc_state *s, const unsigned char *active);
int orc_force_ex(const orc_params *par, orc_state *s, const unsigned char *active, int *need_wakeup);
long long orc_get_timestep(const orc_params *par, const orc_integ *ip, orc_timeline *tl, const orc_state *s, long i);
int orc_run_begin(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s);
int orc_run_step(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s);
void orc_drift_all(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s);
void orc_process_wake_ups(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s);
void orc_kick(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s, long i,
              long long tstart, long long tend, int mode);
void orc_drift(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s, long i,
               long long time0, long long time1);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
