/*
 * gz_oracle_time.c -- CPU FP64 restatement of the timestep loop around the hydro hot path:
 * find_timesteps() with get_timestep() and the power-of-two time bins, process_wake_ups(),
 * the two half-step kicks, find_next_sync_point_and_drift() with the particle drift, and
 * the active list.  One call of orc_run_step is one iteration of run()'s main loop
 * (src/core/run.cc:220-507) for a gas-only, gravity-free, non-comoving MFM run.
 *
 * TEST INFRASTRUCTURE ONLY, like gz_oracle.c: the correctness oracle and the CPU baseline
 * for the device timestep loop.  The product path never links, loads or calls it.
 *
 * Restated configuration: HYDRO_MESHLESS_FINITE_MASS, SELFGRAVITY_OFF (so the acceleration
 * criterion uses the gas kernel radius as its softening, src/gravity/forcetree.cc:85-91),
 * BOX_PERIODIC, no cooling, no comoving integration (cf_atime = cf_a2inv = cf_hubble_a = 1,
 * dt_displacement = MaxSizeTimestep, src/core/timestep.cc:1425-1433), no timestep dilation.
 *
 * Drift policy, the reference's lazy drift: at a sync point only the particles of the active
 * bins are drifted (find_next_sync_point_and_drift, src/core/run.cc:979-994; move_particles,
 * src/core/predict.cc:175-210); every other particle stays at its own Ti_current until a pass is
 * about to read it, and is then drifted to the current time in one piece
 * (src/mesh/gpu_neighbor_list.cc:2443-2510).  The passes find their pairs on the positions and
 * radii every particle would have at the current time (the pair test the reference applies
 * after its drift), and a particle is drifted the first time a pass accepts it as a partner.
 * (The reference drifts every member its walk's envelope found, a superset of the accepted
 * partners that depends on its tile bounds; a particle drifted in one piece rather than two
 * differs only in rounding unless the 0.3 log cap on density and radius or the energy
 * positivity catch binds within the interval.)  A step whose active bins hold more than
 * All.DomainBuild_ActiveFraction of the particles drifts every particle first, as the
 * reference does before its domain decomposition (src/core/run.cc:360,380-390); 0 drifts every
 * particle at every step.  After each drift the position is wrapped into [0, box) once per
 * axis, which the reference does at its decompositions (do_box_wrapping); the pair test takes
 * nearest images either way.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "gz_oracle.h"

#define OT_TIMEBASE (((long long)1) << ORC_TIMEBINS) /* src/declarations/typedefs.h:8-9 */
#define OT_WAKEUP 4.1                                 /* src/declarations/constants.h:13 */

/* GET_INTEGERTIME_FROM_TIMEBIN, src/declarations/macros.h:7 */
static long long ot_bin_len(int bin) { return bin ? (((long long)1) << bin) : 0; }

/* get_timestep_bin, src/core/timestep.cc:1525-1546 (a step of 1 is the stop 90001006 there) */
static int ot_timestep_bin(long long ti_step)
{
    int bin = -1;
    if (ti_step == 0) return 0;
    if (ti_step == 1) return 0;
    while (ti_step) {
        bin++;
        ti_step >>= 1;
    }
    return bin;
}

/* KERNEL_CORE_SIZE, src/mesh/kernel.h:59 (cubic spline) and :113 (Wendland C2) */
static double ot_core_size(int kernel) { return kernel == 3 ? (1.0 / 2.0) : (2.0 / 5.0); }

/* Get_Particle_Size in 3D, src/declarations/particle_data.h:461-467 (NumNgb holds the cube root) */
static double ot_particle_size(const orc_state *s, long i) { return 1.61199 * s->h[i] / s->numngb[i]; }

/* internal_energy_positivity_catch with no material floor, src/core/proto.h:417-421 */
static double ot_positivity(double u_old, double u_new) { return (u_new < 0.5 * u_old) ? 0.5 * u_old : u_new; }

/* kernel_radius_drift_log_change, src/core/timestep_functions.h:312,322-328 */
static double ot_drift_log_change(double x)
{
    if (x > +0.3) return +0.3;
    if (x < -0.3) return -0.3;
    return x;
}

static void ot_stop(orc_timeline *tl, int code)
{
    if (!tl->stop_code) tl->stop_code = code;
}

/* get_timestep, src/core/timestep.cc:357-1423, for a gas particle of this configuration:
   the acceleration criterion (:417-440), the Courant criterion on MaxSignalVel (:711-728),
   the velocity-divergence criterion (:1092-1098), then the caps and the integer conversion
   (:1360-1417). */
long long orc_get_timestep(const orc_params *par, const orc_integ *ip, orc_timeline *tl, const orc_state *s, long i)
{
    double ax = 0.0, ay = 0.0, az = 0.0; /* cf_a2inv * GravAccel: no gravity */
    ax += s->accel[3 * i + 0];
    ay += s->accel[3 * i + 1];
    az += s->accel[3 * i + 2];
    double ac = sqrt(ax * ax + ay * ay + az * az);
    if (ac == 0) ac = 1.0e-30;
    double dt;
    {
        double h_for_accel_dt = ot_core_size(par->kernel) * s->soft[i];
        if (h_for_accel_dt <= 0) h_for_accel_dt = ot_particle_size(s, i) * 1.0;
        dt = sqrt(2 * ip->err_tol_int_accuracy * 1.0 * h_for_accel_dt / ac);
    }
    if (s->mass[i] > 0) {
        double csnd = 0.5 * s->maxsig[i];
        double L_particle = ot_particle_size(s, i);
        double dt_courant = ip->courant_fac * (L_particle * 1.0) / csnd;
        if (dt_courant < dt) dt = dt_courant;
        double divVel = s->divvel[i];
        if (divVel != 0) {
            double dt_divv = 1.5 / fabs(1.0 * divVel);
            if (dt_divv < dt) dt = dt_divv;
        }
    }
    dt *= 1.0; /* cf_hubble_a */
    if (dt >= ip->max_size_timestep) dt = ip->max_size_timestep;
    if (dt >= ip->max_size_timestep) dt = ip->max_size_timestep; /* dt_displacement */
    if ((dt < ip->min_size_timestep) || !(dt / tl->timebase_interval >= 2.0)) dt = ip->min_size_timestep;
    if (dt > 0.5 * OT_TIMEBASE * tl->timebase_interval) dt = 0.5 * OT_TIMEBASE * tl->timebase_interval;
    long long ti_step = (long long)(dt / tl->timebase_interval);
    if (ti_step <= 1) ti_step = 2;
    if (!(ti_step > 0 && ti_step < OT_TIMEBASE)) {
        ot_stop(tl, 818); /* endrun(818) of :1398-1411, which hands back 2 */
        ti_step = 2;
    }
    return ti_step;
}

/* do_the_kick, src/core/kicks.cc:284-557, for an MFM gas particle without gravity: the
   energy kick with the positivity catch and temperature floor, the momentum kick, and at the
   end of a full step (mode 1) the predicted velocity and energy reset to the kicked ones.
   The energy/entropy switch never fires here: with no gravitational work its condition
   0.01*e_potential > e_thermal cannot hold (:356-376). */
void orc_kick(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s, long i,
              long long tstart, long long tend, int mode)
{
    (void)par;
    if (!tl->bin_active[s->timebin[i]]) return;
    const double mass = s->mass[i];
    const double dt_hydrokick = (tend - tstart) * tl->timebase_interval;
    const double du_tot = s->dtu[i] * dt_hydrokick + 0.0;
    const double dEnt = s->u_kick[i] + du_tot;
    s->u_kick[i] = ot_positivity(s->u_kick[i], dEnt);
    if (ip->min_egy_spec) {
        if (s->u_kick[i] < ip->min_egy_spec) {
            s->u_kick[i] = ip->min_egy_spec;
            s->dtu[i] = 0;
        }
    }
    const double fac = mass * 1.0 * dt_hydrokick;
    const double inv = 1.0 / mass;
    for (int k = 0; k < 3; k++) {
        double dp = 0.0;
        dp += s->accel[3 * i + k] * fac;
        s->vel_kick[3 * i + k] += dp * inv;
    }
    if (mode == 1) {
        for (int k = 0; k < 3; k++) s->vel[3 * i + k] = s->vel_kick[3 * i + k];
        s->u[i] = s->u_kick[i];
    }
}

/* drift_particle_impl, src/core/drift_particle_functions.h:86-250, for an MFM gas particle:
   position step on the kicked velocity, predicted velocity, density and energy, the kernel
   radius (HYDRO_FIX_MESH_MOTION 5 moves it with the flow, src/declarations/precompiler_logic.h:78)
   and the pressure (set_eos_pressure_impl, src/eos/eos_functions.h:830). */
void orc_drift(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s, long i,
               long long time0, long long time1)
{
    if (time1 == time0) return;
    const double dt_drift = (time1 - time0) * tl->timebase_interval * 1.0;
    for (int k = 0; k < 3; k++) s->pos[3 * i + k] += s->vel_kick[3 * i + k] * dt_drift;
    const double divv_fac = ot_drift_log_change(s->divvel[i] * dt_drift);
    const double h_drift_fac = exp(ot_drift_log_change(s->divvel[i] * dt_drift) / ((double)3));
    if (s->mass[i] > 0) {
        const double dt_hydrokick = (time1 - time0) * tl->timebase_interval;
        for (int k = 0; k < 3; k++) s->vel[3 * i + k] += 0.0 + s->accel[3 * i + k] * (dt_hydrokick * 1.0);
        s->rho[i] *= exp(-divv_fac);
        const double etmp = s->u[i] + s->dtu[i] * dt_hydrokick;
        s->u[i] = ot_positivity(s->u[i], etmp);
        if (s->u[i] < ip->min_egy_spec) s->u[i] = ip->min_egy_spec;
        s->h[i] *= h_drift_fac;
        if (s->h[i] < par->min_kernel_radius) s->h[i] = par->min_kernel_radius;
        if (s->h[i] > par->max_kernel_radius) s->h[i] = par->max_kernel_radius;
        s->press[i] = (par->gamma - 1) * s->u[i] * s->rho[i];
    }
    for (int k = 0; k < 3; k++) {
        double x = s->pos[3 * i + k];
        if (x >= par->box) x -= par->box;
        else if (x < 0) x += par->box;
        s->pos[3 * i + k] = x;
    }
}

static int ot_is_active(const orc_timeline *tl, const orc_state *s, long i)
{
    return tl->bin_active[s->timebin[i]] && (s->mass[i] > 0);
}

/* process_wake_ups, src/core/timestep.cc:1565-1774: an inactive particle flagged by a
   neighbour whose signal speed it cannot follow moves to the waker's bin minus
   log2(WAKEUP) rounded up, floored at the lowest occupied active bin and capped at the
   highest bin active at the next kick.  Its old step is completed from its midpoint and the
   new step's first half-kick applied when the new bin is active now. */
static void ot_process_wake_ups(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s)
{
    tl->n_woken = 0;
    long long ti_next_kick = OT_TIMEBASE;
    for (int n = 0; n < ORC_TIMEBINS; n++) {
        if (tl->bin_count[n]) {
            long long ti_next_for_bin;
            if (n > 0) {
                long long dt_bin = ot_bin_len(n);
                ti_next_for_bin = (tl->ti_current / dt_bin) * dt_bin + dt_bin;
            } else {
                ti_next_for_bin = tl->ti_current;
            }
            if (ti_next_for_bin < ti_next_kick) ti_next_kick = ti_next_for_bin;
        }
    }
    int max_time_bin_active = 0;
    for (int n = 0; n < ORC_TIMEBINS; n++) {
        long long dt_bin = ((long long)1) << n;
        if ((ti_next_kick % dt_bin) == 0) max_time_bin_active = n;
    }
    int wakeup_bin_offset = 0;
    while (((long long)1 << wakeup_bin_offset) < (long long)OT_WAKEUP) wakeup_bin_offset++;
    if (tl->need_wakeup) {
        int lowest_occupied_active_bin = ORC_TIMEBINS;
        for (int nb = 0; nb < ORC_TIMEBINS; nb++) {
            if (tl->bin_active[nb] && tl->bin_count[nb] > 0) {
                lowest_occupied_active_bin = nb;
                break;
            }
        }
        for (long i = 0; i < s->n; i++) {
            if (!s->wakeup[i]) continue;
            if (s->mass[i] <= 0) continue;
            int binold = s->timebin[i];
            if (tl->bin_active[binold]) continue;
            int bin;
            if (s->wakeup[i] > 0) {
                int waker_bin = s->wakeup[i] - 1; /* wakeup_timebin_from_flag, src/core/proto.h:182 */
                bin = waker_bin - wakeup_bin_offset;
                if (bin < 0) bin = 0;
                if (bin < lowest_occupied_active_bin) bin = lowest_occupied_active_bin;
            } else {
                bin = max_time_bin_active;
            }
            if (bin > max_time_bin_active) bin = max_time_bin_active;
            if (bin >= binold) bin = binold;
            if (bin == binold) continue;
            const long long ti_current_i = s->ti_current[i];
            const long long old_midpoint = s->ti_begstep[i] + s->dt_step[i] / 2;
            const int new_step_starts_now = (ti_current_i == tl->ti_current) && tl->bin_active[bin] &&
                                            (s->dt_step[i] == ot_bin_len(binold));
            tl->bin_count[binold]--;
            tl->bin_count[bin]++;
            s->timebin[i] = bin;
            tl->n_woken++;
            if (new_step_starts_now) orc_kick(par, ip, tl, s, i, old_midpoint, tl->ti_current, 1);
            {
                long long dt_truncated = tl->ti_current - ti_current_i;
                if (dt_truncated > 0) {
                    s->ti_begstep[i] = ti_current_i;
                    s->dt_step[i] = dt_truncated;
                } else {
                    s->ti_begstep[i] = tl->ti_current;
                    s->dt_step[i] = ot_bin_len(bin);
                }
            }
            if (ti_current_i < tl->ti_current) s->ti_current[i] = tl->ti_current;
            if (new_step_starts_now)
                orc_kick(par, ip, tl, s, i, s->ti_begstep[i], s->ti_begstep[i] + s->dt_step[i] / 2, 0);
        }
    }
    tl->need_wakeup = 0;
}

/* find_timesteps, src/core/timestep.cc:39-355: a new power-of-two step for every particle of
   the active list, a larger bin only where it is synchronised now, then the wake-ups. */
static void ot_find_timesteps(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s,
                              const unsigned char *act)
{
    for (long i = 0; i < s->n; i++) {
        if (!act[i]) continue;
        long long ti_step = orc_get_timestep(par, ip, tl, s, i);
        ti_step = (long long)(((double)ti_step) / 1.0); /* the (unit) dilation factor */
        long long ti_min = OT_TIMEBASE;
        while (ti_min > ti_step) ti_min >>= 1;
        ti_step = ti_min;
        int bin = ot_timestep_bin(ti_step);
        int binold = s->timebin[i];
        if (bin > binold) {
            while (tl->bin_active[bin] == 0 && bin > binold) bin--;
            ti_step = ot_bin_len(bin);
        }
        if (tl->ti_current >= OT_TIMEBASE) {
            ti_step = 0;
            bin = 0;
        }
        if ((OT_TIMEBASE - tl->ti_current) < ti_step) {
            ot_stop(tl, 90001005);
            ti_step = OT_TIMEBASE - tl->ti_current;
            ti_min = OT_TIMEBASE;
            while (ti_min > ti_step) ti_min >>= 1;
            ti_step = ti_min;
        }
        if (bin != binold) {
            tl->bin_count[binold]--;
            tl->bin_count[bin]++;
            s->timebin[i] = bin;
        }
        s->ti_begstep[i] += s->dt_step[i];
        s->dt_step[i] = ti_step;
    }
    ot_process_wake_ups(par, ip, tl, s);
}

/* process_wake_ups on its own (tests: the bin choice and the step truncation of given states) */
void orc_process_wake_ups(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s)
{
    ot_process_wake_ups(par, ip, tl, s);
}

/* find_next_sync_point_and_drift, src/core/run.cc:857-994: the earliest next kick over the
   occupied bins, the bins active there, and every particle drifted to it. */
static void ot_next_sync_point_and_drift(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s)
{
    long long ti_next_kick = OT_TIMEBASE;
    for (int n = 0; n < ORC_TIMEBINS; n++) {
        if (tl->bin_count[n]) {
            long long ti_next_for_bin;
            if (n > 0) {
                long long dt_bin = ot_bin_len(n);
                ti_next_for_bin = (tl->ti_current / dt_bin) * dt_bin + dt_bin;
            } else {
                ti_next_for_bin = tl->ti_current;
            }
            if (ti_next_for_bin < ti_next_kick) ti_next_kick = ti_next_for_bin;
        }
    }
    tl->num_steps++;
    tl->ti_current = ti_next_kick;
    tl->time = ip->time_begin + tl->ti_current * tl->timebase_interval;
    tl->bin_active[0] = 1;
    for (int n = 1; n < ORC_TIMEBINS; n++) tl->bin_active[n] = ((ti_next_kick % ot_bin_len(n)) == 0);
    /* NumForceUpdate (src/core/run.cc:942-955) against the domain cadence */
    long long nf = 0;
    for (int n = 0; n < ORC_TIMEBINS; n++) if (tl->bin_active[n]) nf += tl->bin_count[n];
    tl->full_drift = (double)nf > ip->domain_build_active_fraction * (double)s->n;
    const long long now = tl->ti_current;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < s->n; i++) {
        if (!tl->full_drift && !tl->bin_active[s->timebin[i]]) continue;
        orc_drift(par, ip, tl, s, i, s->ti_current[i], now);
        s->ti_current[i] = now;
    }
}

/* ------------------------------------------------------------------------- */
/* the lazy drift's hook for the passes of a step (orc_lazy)                    */
/* ------------------------------------------------------------------------- */
static struct {
    const orc_params *par;
    const orc_integ *ip;
    const orc_timeline *tl;
    orc_state *s;
    unsigned char *state;  /* per particle: 0 behind, 1 being drifted, 2 at the current time */
} ot_lazy;

static void ot_ensure(long j)
{
    unsigned char *st = &ot_lazy.state[j];
    if (__atomic_load_n(st, __ATOMIC_ACQUIRE) == 2) return;
    unsigned char expect = 0;
    if (__atomic_compare_exchange_n(st, &expect, 1, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
        orc_drift(ot_lazy.par, ot_lazy.ip, ot_lazy.tl, ot_lazy.s, j, ot_lazy.s->ti_current[j], ot_lazy.tl->ti_current);
        ot_lazy.s->ti_current[j] = ot_lazy.tl->ti_current;
        __atomic_store_n(st, 2, __ATOMIC_RELEASE);
        return;
    }
    while (__atomic_load_n(st, __ATOMIC_ACQUIRE) != 2) {}
}

/* the position and kernel radius orc_drift would give particle i at time1 (the same arithmetic) */
static void ot_predict(const orc_params *par, const orc_timeline *tl, const orc_state *s, long i, long long time1, double *x, double *h)
{
    const long long time0 = s->ti_current[i];
    for (int k = 0; k < 3; k++) x[k] = s->pos[3 * i + k];
    *h = s->h[i];
    if (time1 == time0) return;
    const double dt_drift = (time1 - time0) * tl->timebase_interval * 1.0;
    for (int k = 0; k < 3; k++) x[k] += s->vel_kick[3 * i + k] * dt_drift;
    if (s->mass[i] > 0) {
        const double h_drift_fac = exp(ot_drift_log_change(s->divvel[i] * dt_drift) / ((double)3));
        *h *= h_drift_fac;
        if (*h < par->min_kernel_radius) *h = par->min_kernel_radius;
        if (*h > par->max_kernel_radius) *h = par->max_kernel_radius;
    }
    for (int k = 0; k < 3; k++) {
        if (x[k] >= par->box) x[k] -= par->box;
        else if (x[k] < 0) x[k] += par->box;
    }
}

static int ot_hydro(const orc_params *par0, const orc_integ *ip, const orc_timeline *tl, orc_state *s, const unsigned char *act,
                    int *need_wakeup)
{
    orc_params par = *par0;
    par.time = tl->time;
    const long n = s->n;
    int lazy = 0;
    for (long i = 0; i < n && !lazy; i++) lazy = s->ti_current[i] != tl->ti_current;
    if (!lazy) {
        int rc = orc_density(&par, s, act, NULL);
        if (rc) return rc;
        rc = orc_gradients(&par, s, act);
        if (rc) return rc;
        return orc_force_ex(&par, s, act, need_wakeup);
    }
    /* a lazy step: the passes see every particle where the drift would put it now, and drift a
       particle before they read it */
    double *pos = (double *)malloc((size_t)n * 3 * sizeof(double));
    double *hp = (double *)malloc((size_t)n * sizeof(double));
    unsigned char *state = (unsigned char *)malloc((size_t)n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        ot_predict(&par, tl, s, i, tl->ti_current, &pos[3 * i], &hp[i]);
        state[i] = s->ti_current[i] == tl->ti_current ? 2 : 0;
    }
    ot_lazy.par = &par; ot_lazy.ip = ip; ot_lazy.tl = tl; ot_lazy.s = s; ot_lazy.state = state;
    orc_lazy view = {pos, hp, ot_ensure};
    orc_set_lazy(&view);
    int rc = orc_density(&par, s, act, NULL);
    if (!rc) {
        /* the rows' new radii (the rows stand at the current time) */
        for (long i = 0; i < n; i++) if (state[i] == 2) hp[i] = s->h[i];
        rc = orc_gradients(&par, s, act);
    }
    if (!rc) rc = orc_force_ex(&par, s, act, need_wakeup);
    orc_set_lazy(NULL);
    free(pos); free(hp); free(state);
    return rc;
}

/* gizmo_full_drift_to(All.Ti_Current), src/core/predict.cc:148-171: every particle behind the
   current sync point drifted to it (before a snapshot) */
void orc_drift_all(const orc_params *par, const orc_integ *ip, const orc_timeline *tl, orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < s->n; i++) {
        if (s->ti_current[i] == tl->ti_current) continue;
        orc_drift(par, ip, tl, s, i, s->ti_current[i], tl->ti_current);
        s->ti_current[i] = tl->ti_current;
    }
}

/* Start of a run: the timeline and bins of init() (src/core/init.cc:108-113,258-265,583),
   predicted fields equal to the current ones, the initial density() (:1051) and softening
   cache (:1061), then run()'s first compute_hydro_densities_and_forces()
   (src/core/run.cc:128-150) with every particle active. */
int orc_run_begin(const orc_params *par0, const orc_integ *ip, orc_timeline *tl, orc_state *s)
{
    memset(tl, 0, sizeof *tl);
    tl->full_drift = 1;
    tl->timebase_interval = (ip->time_max - ip->time_begin) / OT_TIMEBASE;
    tl->ti_current = 0;
    tl->time = ip->time_begin;
    for (int b = 0; b < ORC_TIMEBINS; b++) tl->bin_active[b] = 1;
    for (long i = 0; i < s->n; i++) {
        s->timebin[i] = 0;
        s->ti_begstep[i] = 0;
        s->dt_step[i] = 0;
        s->ti_current[i] = 0;
        s->wakeup[i] = 0;
        for (int k = 0; k < 3; k++) s->vel[3 * i + k] = s->vel_kick[3 * i + k];
        s->u[i] = s->u_kick[i];
    }
    tl->bin_count[0] = s->n;
    orc_params par = *par0;
    par.time = tl->time;
    par.time_begin = ip->time_begin;
    int rc = orc_density(&par, s, NULL, NULL);
    if (rc) return rc;
    for (long i = 0; i < s->n; i++) s->soft[i] = s->h[i];
    int nw = 0;
    rc = ot_hydro(&par, ip, tl, s, NULL, &nw);
    tl->need_wakeup |= nw;
    tl->n_active = 0;
    for (long i = 0; i < s->n; i++) tl->n_active += (s->mass[i] > 0);
    return rc;
}

/* One iteration of run()'s main loop (src/core/run.cc:220-507): find_timesteps,
   do_first_halfstep_kick (src/core/kicks.cc:34-97), find_next_sync_point_and_drift,
   make_list_of_active_particles (src/core/run.cc:997-1009), the softening cache of the
   active particles (src/gravity/gravtree.cc:267), compute_hydro_densities_and_forces
   (src/core/accel.cc:68) and do_second_halfstep_kick (src/core/kicks.cc:99-185). */
int orc_run_step(const orc_params *par, const orc_integ *ip, orc_timeline *tl, orc_state *s)
{
    const long n = s->n;
    unsigned char *act = (unsigned char *)malloc((size_t)(n > 0 ? n : 1));
    for (long i = 0; i < n; i++) act[i] = (unsigned char)ot_is_active(tl, s, i);
    ot_find_timesteps(par, ip, tl, s, act);
    for (long i = 0; i < n; i++) {
        if (!act[i]) continue;
        orc_kick(par, ip, tl, s, i, s->ti_begstep[i], s->ti_begstep[i] + s->dt_step[i] / 2, 0);
    }
    ot_next_sync_point_and_drift(par, ip, tl, s);
    long long na = 0;
    for (long i = 0; i < n; i++) {
        act[i] = (unsigned char)ot_is_active(tl, s, i);
        na += act[i];
        if (act[i]) s->soft[i] = s->h[i];
    }
    tl->n_active = na;
    int nw = 0;
    int rc = ot_hydro(par, ip, tl, s, act, &nw);
    tl->need_wakeup |= nw;
    if (!rc) {
        for (long i = 0; i < n; i++) {
            if (!act[i]) continue;
            orc_kick(par, ip, tl, s, i, s->ti_begstep[i] + s->dt_step[i] / 2, s->ti_begstep[i] + s->dt_step[i], 1);
        }
    }
    free(act);
    if (!rc && tl->stop_code) rc = tl->stop_code;
    return rc;
}
