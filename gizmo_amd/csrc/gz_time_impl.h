// gz_time_impl.h -- the timestep loop around the hot path on the device (FP64; compiled into
// the strict-FP translation unit).  One lane per particle: new power-of-two steps for the
// active particles, the wake-up pass, the two half-step kicks and the drift to the next sync
// point, each restating the reference function named on it for a gas-only, gravity-free,
// non-comoving MFM run.  The host keeps the integer timeline and the bin occupation (60 counts,
// read back once per step) and passes the active bins as a bit mask.
#pragma once
#include "gz_kernels_impl.h"

namespace gz {

constexpr long long kTimebase = 1LL << kTimebins;  // TIMEBASE, src/declarations/typedefs.h:8-9

// GET_INTEGERTIME_FROM_TIMEBIN, src/declarations/macros.h:7
__device__ __forceinline__ long long bin_len(int bin) { return bin ? (1LL << bin) : 0LL; }
__device__ __forceinline__ bool bin_on(unsigned long long mask, int bin) { return (mask >> bin) & 1ull; }

// get_timestep_bin, src/core/timestep.cc:1525-1546
__device__ __forceinline__ int timestep_bin(long long ti_step)
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

// internal_energy_positivity_catch without a material floor, src/core/proto.h:417-421
__device__ __forceinline__ double positivity(double u_old, double u_new) { return (u_new < 0.5 * u_old) ? 0.5 * u_old : u_new; }

// kernel_radius_drift_log_change, src/core/timestep_functions.h:312,322-328
__device__ __forceinline__ double drift_log_change(double x)
{
    if (x > +0.3) return +0.3;
    if (x < -0.3) return -0.3;
    return x;
}

// the start of a sync point (threads 0-63 of one workgroup): the step's counters and maxima
// cleared (the bin occupation and the running pair count kept)
__device__ __forceinline__ void step_prep(int *__restrict__ counters, unsigned long long *__restrict__ u64c)
{
    const int t = threadIdx.x;
    if (t < 16) counters[t] = 0;
    if (t == 0) { u64c[64] = 0; u64c[66] = 0; u64c[67] = 0; u64c[68] = 0; u64c[72] = 0; u64c[73] = 0; u64c[75] = 0; u64c[76] = 0; }
}

// the largest kicked velocity component of the calling lanes' particles (valid lanes only) into
// the bound *vkmax: one conditional atomic per wavefront (all 64 lanes call it together)
__device__ __forceinline__ void raise_vk_bound(const ParticleView &S, long i, bool valid)
{
    double m = valid ? fmax(fabs(S.VKX[i]), fmax(fabs(S.VKY[i]), fabs(S.VKZ[i]))) : 0.0;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0) atomic_max_bits(S.vkmax, m);
}

// get_timestep, src/core/timestep.cc:357-1423 for a gas particle here: acceleration criterion
// with the kernel radius as softening (:417-440, src/gravity/forcetree.cc:85-91), Courant on
// MaxSignalVel (:711-728), velocity divergence (:1092-1098), caps and integer step (:1360-1417).
// zero_step: the step failed the integer-timeline check that ends the run with stop 818
// (:1398-1411); the step handed back is then 2, as there.  After the floor of 2 and the cap at
// half the timeline the check cannot fail for a finite or NaN dt (NaN takes the minimum step),
// so the flag only guards the timeline's invariants.
__device__ long long get_timestep(const TimeParams &T, const ParticleView &S, long i, bool &zero_step)
{
    const long n = S.n;
    double ax = 0.0, ay = 0.0, az = 0.0;
    ax += S.ACC[i];
    ay += S.ACC[n + i];
    az += S.ACC[2 * n + i];
    double ac = sqrt(ax * ax + ay * ay + az * az);
    if (ac == 0) ac = 1.0e-30;
    double dt;
    {
        double h_for_accel_dt = T.core_size * S.SOFT[i];
        if (h_for_accel_dt <= 0) h_for_accel_dt = 1.61199 * S.H[i] / S.NN[i] * 1.0;
        dt = sqrt(2 * T.err_tol * 1.0 * h_for_accel_dt / ac);
    }
    if (S.M[i] > 0) {
        const double csnd = 0.5 * S.MSV[i];
        const double L_particle = 1.61199 * S.H[i] / S.NN[i];
        const double dt_courant = T.courant_fac * (L_particle * 1.0) / csnd;
        if (dt_courant < dt) dt = dt_courant;
        const double divVel = S.DIVV[i];
        if (divVel != 0) {
            const double dt_divv = 1.5 / fabs(1.0 * divVel);
            if (dt_divv < dt) dt = dt_divv;
        }
    }
    dt *= 1.0;
    if (dt >= T.max_dt) dt = T.max_dt;
    if (dt >= T.max_dt) dt = T.max_dt;
    if ((dt < T.min_dt) || !(dt / T.tbi >= 2.0)) dt = T.min_dt;
    if (dt > 0.5 * kTimebase * T.tbi) dt = 0.5 * kTimebase * T.tbi;
    long long ti_step = (long long)(dt / T.tbi);
    if (ti_step <= 1) ti_step = 2;
    zero_step = !(ti_step > 0 && ti_step < kTimebase);
    if (zero_step) ti_step = 2;
    return ti_step;
}

// do_the_kick, src/core/kicks.cc:284-557, MFM gas without gravity (the energy/entropy switch
// cannot fire without gravitational work, :356-376)
__device__ void kick(const TimeParams &T, const ParticleView &S, long i, long long tstart, long long tend, int mode)
{
    if (!bin_on(T.active_mask, S.TB[i])) return;
    const long n = S.n;
    const double mass = S.M[i];
    const double dt_hydrokick = (tend - tstart) * T.tbi;
    const double du_tot = S.DTU[i] * dt_hydrokick + 0.0;
    const double dEnt = S.UK[i] + du_tot;
    double u = positivity(S.UK[i], dEnt);
    if (T.min_egy) {
        if (u < T.min_egy) { u = T.min_egy; S.DTU[i] = 0; }
    }
    S.UK[i] = u;
    const double fac = mass * 1.0 * dt_hydrokick;
    const double inv = 1.0 / mass;
    double *vk[3] = {S.VKX, S.VKY, S.VKZ};
    for (int k = 0; k < 3; k++) {
        double dp = 0.0;
        dp += S.ACC[(size_t)k * n + i] * fac;
        vk[k][i] += dp * inv;
    }
    if (mode == 1) {
        S.VX[i] = S.VKX[i]; S.VY[i] = S.VKY[i]; S.VZ[i] = S.VKZ[i];
        S.U[i] = S.UK[i];
    }
}

// find_timesteps, src/core/timestep.cc:39-355: the new power-of-two step of each particle of
// the active list; a larger bin only where it is synchronised now
__global__ void __launch_bounds__(256) k_find_timesteps(TimeParams T, ParticleView S, int nrows, const int *__restrict__ act,
                                                         int *__restrict__ err, unsigned long long *__restrict__ count, BoxTable bt,
                                                         int kick_first)
{
    // count (one box, may be null): TimeBinCount kept up to date with the moves of the rows,
    // added up per block in LDS
    __shared__ int moved[64];
    if (count && threadIdx.x < 64) moved[threadIdx.x] = 0;
    if (count) __syncthreads();
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    double vkmax_row = 0;  // the row's largest kicked velocity component after its kick
    if (r < nrows) {
        const int i = act[r];
        resolve_time(bt, i, T);
        // the row's inputs loaded ahead of its stores (for the compiler the stores may alias
        // them, so loads after them would wait for each other in turn)
        const long n = S.n;
        const int binold = S.TB[i];
        const long long tib0 = S.TIB[i], dts0 = S.DTS[i];
        const double m = S.M[i];
        double dtu = 0, uk = 0, a[3] = {0, 0, 0}, vk[3] = {0, 0, 0};
        if (kick_first) {
            vk[0] = S.VKX[i]; vk[1] = S.VKY[i]; vk[2] = S.VKZ[i];
            dtu = S.DTU[i]; uk = S.UK[i];
            a[0] = S.ACC[i]; a[1] = S.ACC[n + i]; a[2] = S.ACC[2 * n + i];
        }
        bool zero_step;
        long long ti_step = get_timestep(T, S, i, zero_step);
        if (zero_step) {
            if (bt.b) atomicMax(&bt.slot[(size_t)bt.BOX[i] * kBoxSlots + kSlotStop], 818ull);
            else atomicMax(err, 818);
        }
        ti_step = (long long)(((double)ti_step) / 1.0);
        long long ti_min = kTimebase;
        while (ti_min > ti_step) ti_min >>= 1;
        ti_step = ti_min;
        int bin = timestep_bin(ti_step);
        if (bin > binold) {
            while (!bin_on(T.active_mask, bin) && bin > binold) bin--;
            ti_step = bin_len(bin);
        }
        if (T.ti_current >= kTimebase) { ti_step = 0; bin = 0; }
        if ((kTimebase - T.ti_current) < ti_step) {
            if (bt.b) atomicMax(&bt.slot[(size_t)bt.BOX[i] * kBoxSlots + kSlotStop], 90001005ull);
            else atomicMax(err, 90001005);
            ti_step = kTimebase - T.ti_current;
            ti_min = kTimebase;
            while (ti_min > ti_step) ti_min >>= 1;
            ti_step = ti_min;
        }
        S.TB[i] = bin;
        const long long tib = tib0 + dts0;
        S.TIB[i] = tib;
        S.DTS[i] = ti_step;
        // do_first_halfstep_kick (src/core/kicks.cc:34-97) of the row, as k_halfstep_kick takes
        // it (kick(), mode 0, on the values loaded above): it reads only the row's own fields,
        // which process_wake_ups (inactive particles only) does not touch
        if (kick_first && m > 0 && bin_on(T.active_mask, bin)) {
            const double dt_hydrokick = ((tib + ti_step / 2) - tib) * T.tbi;
            const double du_tot = dtu * dt_hydrokick + 0.0;
            const double dEnt = uk + du_tot;
            double u = positivity(uk, dEnt);
            if (T.min_egy) {
                if (u < T.min_egy) { u = T.min_egy; S.DTU[i] = 0; }
            }
            S.UK[i] = u;
            const double fac = m * 1.0 * dt_hydrokick;
            const double inv = 1.0 / m;
            for (int k = 0; k < 3; k++) {
                double dp = 0.0;
                dp += a[k] * fac;
                vk[k] += dp * inv;
            }
            S.VKX[i] = vk[0]; S.VKY[i] = vk[1]; S.VKZ[i] = vk[2];
        }
        if (kick_first) vkmax_row = fmax(fabs(vk[0]), fmax(fabs(vk[1]), fabs(vk[2])));
        if (count && bin != binold) {
            atomicAdd(&moved[binold & 63], -1);
            atomicAdd(&moved[bin & 63], 1);
        }
    }
    if (S.vkmax) {
        // the bound on the kicked velocities (raise_vk_bound on the values in hand)
        double mv = vkmax_row;
        for (int o = 32; o > 0; o >>= 1) mv = fmax(mv, __shfl_xor(mv, o, 64));
        if ((threadIdx.x & 63) == 0 && mv > 0) atomic_max_bits(S.vkmax, mv);
    }
    if (!count) return;
    __syncthreads();
    if (threadIdx.x < 64 && moved[threadIdx.x]) atomicAdd(&count[threadIdx.x], (unsigned long long)(long long)moved[threadIdx.x]);
}

// TimeBinCount of every box (count[64*box + bin]).  Each block counts a contiguous range of
// particles into an LDS histogram of the box it is in (particles of one box are contiguous, so
// the box changes rarely inside a range; the histogram is added to the global counts whenever
// it does, and at the end); lanes of the same bin are added together (most particles of a wave
// share one bin).  Particles of another box than the block's current one (a range that starts
// inside one box and reaches into the next) are added to the global counts directly.
__global__ void __launch_bounds__(256) k_bin_hist(long n, const int *__restrict__ TB, unsigned long long *__restrict__ count, BoxTable bt)
{
    __shared__ unsigned int h[64];
    __shared__ int cur;
    const int lane = threadIdx.x & 63;
    const long span = (n + gridDim.x - 1) / gridDim.x;
    const long lo = blockIdx.x * span;
    const long hi = lo + span < n ? lo + span : n;
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    if (threadIdx.x == 0) cur = (bt.b && lo < hi) ? bt.BOX[lo] : 0;
    __syncthreads();
    for (long base = lo; base < hi; base += blockDim.x) {
        const int cb = bt.b ? bt.BOX[base] : 0;  // the same for the whole block
        if (cb != cur) {
            __syncthreads();
            if (threadIdx.x < 64) {
                if (h[threadIdx.x]) atomicAdd(&count[(size_t)64 * cur + threadIdx.x], (unsigned long long)h[threadIdx.x]);
                h[threadIdx.x] = 0;
            }
            __syncthreads();
            if (threadIdx.x == 0) cur = cb;
            __syncthreads();
        }
        const long i = base + threadIdx.x;
        int b = -1;
        if (i < hi) {
            const int box = bt.b ? bt.BOX[i] : 0;
            b = TB[i] & 63;
            if (box != cb) {
                atomicAdd(&count[(size_t)64 * box + b], 1ull);
                b = -1;
            }
        }
        while (true) {
            const unsigned long long live = __ballot(b >= 0);
            if (!live) break;
            const int leader = __ffsll((long long)live) - 1;
            const int lb = __shfl(b, leader, 64);
            const unsigned long long same = __ballot(b == lb);
            if (lane == leader) atomicAdd(&h[lb], (unsigned int)__popcll(same));
            if (b == lb) b = -1;
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&count[(size_t)64 * cur + threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// The earliest next kick over the occupied bins of one box (src/core/run.cc:870-890,
// src/core/timestep.cc:1577-1591), one wavefront per box with lane k taking bin k; every lane
// returns the minimum.
__device__ __forceinline__ long long wave_next_kick(const unsigned long long *__restrict__ c, long long tic)
{
    const int k = threadIdx.x & 63;
    long long t = kTimebase;
    if (k < kTimebins && c[k]) {
        if (k > 0) {
            const long long dt_bin = bin_len(k);
            t = (tic / dt_bin) * dt_bin + dt_bin;
        } else {
            t = tic;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const long long u = __shfl_xor(t, o, 64);
        t = u < t ? u : t;
    }
    return t;
}

// find_next_sync_point_and_drift's timeline part for one box, by one wavefront (lane k = bin k):
// the next kick, the bins active there, the particles in them (NumForceUpdate), and whether
// every particle is drifted at it (src/core/run.cc:857-955)
struct NextSync {
    long long ti;
    unsigned long long mask, nf;
    bool full;
};
__device__ __forceinline__ NextSync next_sync_wave(const TimeParams &T, const unsigned long long *__restrict__ count, long long tic)
{
    const int k = threadIdx.x & 63;
    NextSync s;
    s.ti = wave_next_kick(count, tic);
    s.mask = __ballot(k == 0 || (k < kTimebins && (s.ti % bin_len(k)) == 0));
    unsigned long long nf = (k < kTimebins && ((s.mask >> k) & 1ull)) ? count[k] : 0ull;
    for (int o = 32; o > 0; o >>= 1) nf += __shfl_xor(nf, o, 64);
    s.nf = nf;
    // every particle is drifted at a step whose active fraction reaches the domain cadence
    // (src/core/run.cc:360,380-390: the full drift before the decomposition), or where the lazy
    // drift is not in use; otherwise only the active bins (find_next_sync_point_and_drift,
    // src/core/run.cc:979-994) and the others when read
    s.full = !T.lazy_ok || (double)nf > T.full_frac * (double)T.n_total || (long long)nf > T.lazy_rows;
    return s;
}

// the bin limits of process_wake_ups (src/core/timestep.cc:1577-1625) from the occupation
// after find_timesteps, one wavefront per box (block b = box b, lane k = bin k): lim[2b] the
// lowest occupied active bin, lim[2b+1] the highest bin active at the next kick
__global__ void __launch_bounds__(64) k_wakeup_limits(TimeParams T, const unsigned long long *__restrict__ count, int *__restrict__ lim,
                                                      BoxTable bt)
{
    const int b = blockIdx.x;
    const int k = threadIdx.x;
    long long tic = T.ti_current;
    unsigned long long mask = T.active_mask;
    if (bt.b) { tic = bt.b[b].ti_current; mask = bt.b[b].mask; }
    const unsigned long long *c = count + (size_t)64 * b;
    const long long ti_next_kick = wave_next_kick(c, tic);
    const unsigned long long divides = __ballot(k < kTimebins && (ti_next_kick % (1LL << k)) == 0);
    const unsigned long long occupied_active = __ballot(k < kTimebins && bin_on(mask, k) && c[k] > 0);
    if (k == 0) {
        lim[2 * b] = occupied_active ? __ffsll((long long)occupied_active) - 1 : kTimebins;
        lim[2 * b + 1] = 63 - __clzll((long long)divides);  // bit 0 is always set
    }
}

// process_wake_ups, src/core/timestep.cc:1565-1774; the occupation count follows each move.
// A particle drifted to the sync point (P.Ti_current == All.Ti_Current: every particle when the
// drift is eager, and every woken one when it is lazy, since the force pass that flagged it
// read it) completes its old step from the midpoint and starts the new one with its first
// half-kick when the new bin is active now; one still behind keeps a step truncated at the
// sync point (:1727-1731) and takes no kick.
__device__ __forceinline__ bool wakeup_one(const TimeParams &T0, const ParticleView &S, const int *__restrict__ lim,
                                           int wakeup_bin_offset, unsigned long long *__restrict__ count, int *__restrict__ nwoken,
                                           const BoxTable &bt, long i)
{
    const int w = S.WAKE[i];
    if (!w) return false;
    if (S.M[i] <= 0) return false;
    const int box = bt.b ? bt.BOX[i] : 0;
    if (bt.b && !(bt.b[box].stepping && bt.b[box].need_wakeup)) return false;
    TimeParams T = T0;
    resolve_time(bt, (int)i, T);
    const int lowest_occupied_active_bin = lim[2 * box], max_time_bin_active = lim[2 * box + 1];
    const int binold = S.TB[i];
    if (bin_on(T.active_mask, binold)) return false;
    int bin;
    if (w > 0) {
        bin = (w - 1) - wakeup_bin_offset;
        if (bin < 0) bin = 0;
        if (bin < lowest_occupied_active_bin) bin = lowest_occupied_active_bin;
    } else {
        bin = max_time_bin_active;
    }
    if (bin > max_time_bin_active) bin = max_time_bin_active;
    if (bin >= binold) bin = binold;
    if (bin == binold) return false;
    const long long tic = S.TIC ? S.TIC[i] : T.ti_current;
    const long long old_midpoint = S.TIB[i] + S.DTS[i] / 2;
    const bool starts_now = (tic == T.ti_current) && bin_on(T.active_mask, bin) && (S.DTS[i] == bin_len(binold));
    S.TB[i] = bin;
    atomicAdd(nwoken, 1);
    if (bt.b) atomicAdd(&bt.slot[(size_t)box * kBoxSlots + kSlotWoken], 1ull);
    atomicAdd(&count[(size_t)64 * box + binold], ~0ull);  // -1
    atomicAdd(&count[(size_t)64 * box + bin], 1ull);
    if (starts_now) kick(T, S, i, old_midpoint, T.ti_current, 1);
    const long long dt_truncated = T.ti_current - tic;
    if (dt_truncated > 0) {
        S.TIB[i] = tic;
        S.DTS[i] = dt_truncated;
    } else {
        S.TIB[i] = T.ti_current;
        S.DTS[i] = bin_len(bin);
    }
    if (S.TIC && tic < T.ti_current) S.TIC[i] = T.ti_current;
    if (starts_now) kick(T, S, i, S.TIB[i], S.TIB[i] + S.DTS[i] / 2, 0);
    return starts_now;
}

__global__ void __launch_bounds__(256) k_wakeups(TimeParams T, ParticleView S, const int *__restrict__ lim, int wakeup_bin_offset,
                                                  unsigned long long *__restrict__ count, int *__restrict__ nwoken, BoxTable bt)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool kicked = i < S.n && wakeup_one(T, S, lim, wakeup_bin_offset, count, nwoken, bt, i);
    if (S.vkmax) raise_vk_bound(S, kicked ? i : 0, kicked);
}

// do_first_halfstep_kick / do_second_halfstep_kick, src/core/kicks.cc:34-97,99-185
__global__ void __launch_bounds__(256) k_halfstep_kick(TimeParams T, ParticleView S, int nrows, const int *__restrict__ act, int second,
                                                        BoxTable bt, int *__restrict__ prep_counters,
                                                        unsigned long long *__restrict__ prep_u64c)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (prep_counters && blockIdx.x == 0 && threadIdx.x < 64) step_prep(prep_counters, prep_u64c);
    const int i = r < nrows ? act[r] : 0;
    const bool row = r < nrows && S.M[i] > 0;
    if (row) {
        resolve_time(bt, i, T);
        const long long half = S.DTS[i] / 2;
        if (second) kick(T, S, i, S.TIB[i] + half, S.TIB[i] + S.DTS[i], 1);
        else kick(T, S, i, S.TIB[i], S.TIB[i] + half, 0);
    }
    if (S.vkmax) raise_vk_bound(S, i, row);
}

// drift_particle_impl, src/core/drift_particle_functions.h:86-250 (MFM gas; the kernel radius
// moves with the flow, src/declarations/precompiler_logic.h:78) of particle i from time0 to
// time1, then one periodic wrap per axis.  Returns the particle's periodic distance (largest
// axis) from where it was when the particles were last ordered (X0, Y0, Z0), or without those
// positions this drift's length along its longest axis; h_after (if given) receives the radius
// after the drift.
__device__ __forceinline__ double drift_one(const TimeParams &T, const ParticleView &S, long i, long long time0, long long time1,
                                            const double *__restrict__ X0, const double *__restrict__ Y0,
                                            const double *__restrict__ Z0, double *h_after = nullptr)
{
    const double dt_drift = (time1 - time0) * T.tbi * 1.0;
    const double vkx = S.VKX[i], vky = S.VKY[i], vkz = S.VKZ[i];
    double x = S.X[i] + vkx * dt_drift;
    double y = S.Y[i] + vky * dt_drift;
    double z = S.Z[i] + vkz * dt_drift;
    const double divv_fac = drift_log_change(S.DIVV[i] * dt_drift);
    const double h_drift_fac = exp(drift_log_change(S.DIVV[i] * dt_drift) / ((double)3));
    if (S.M[i] > 0) {
        const long n = S.n;
        const double dt_hydrokick = (time1 - time0) * T.tbi;
        S.VX[i] += 0.0 + S.ACC[i] * (dt_hydrokick * 1.0);
        S.VY[i] += 0.0 + S.ACC[n + i] * (dt_hydrokick * 1.0);
        S.VZ[i] += 0.0 + S.ACC[2 * n + i] * (dt_hydrokick * 1.0);
        const double rho = S.RHO[i] * exp(-divv_fac);
        const double etmp = S.U[i] + S.DTU[i] * dt_hydrokick;
        double u = positivity(S.U[i], etmp);
        if (u < T.min_egy) u = T.min_egy;
        double h = S.H[i] * h_drift_fac;
        if (h < T.min_h) h = T.min_h;
        if (h > T.max_h) h = T.max_h;
        S.RHO[i] = rho;
        S.U[i] = u;
        S.H[i] = h;
        S.PR[i] = (T.gamma - 1) * u * rho;
        if (h_after) *h_after = h;
    } else if (h_after) {
        *h_after = S.H[i];
    }
    if (x >= T.box) x -= T.box; else if (x < 0) x += T.box;
    if (y >= T.box) y -= T.box; else if (y < 0) y += T.box;
    if (z >= T.box) z -= T.box; else if (z < 0) z += T.box;
    S.X[i] = x; S.Y[i] = y; S.Z[i] = z;
    if (!X0) {
        // no ordering positions: this drift's length along the longest axis, padded for the
        // rounding of the position update (the caller sums these into a bound)
        return fmax(fabs(vkx), fmax(fabs(vky), fabs(vkz))) * dt_drift * (1.0 + 1e-12) + 1e-13 * T.box;
    }
    const double half = 0.5 * T.box;
    const double ex = fabs(wrap(x - X0[i], T.box, half)), ey = fabs(wrap(y - Y0[i], T.box, half)),
                 ez = fabs(wrap(z - Z0[i], T.box, half));
    return fmax(ex, fmax(ey, ez));
}

// every particle drifted to its box's next sync point (grid-stride), and the per-box largest
// displacement since the last ordering and largest kernel radius after the drift
__global__ void __launch_bounds__(256) k_drift(TimeParams T, ParticleView S, long long time0, const unsigned long long *__restrict__ tl,
                                                const double *__restrict__ X0, const double *__restrict__ Y0,
                                                const double *__restrict__ Z0, unsigned long long *__restrict__ disp,
                                                unsigned long long *__restrict__ hmax, BoxTable bt)
{
    const long long time1_single = bt.b ? 0 : (long long)tl[0];  // the next sync point, found on the device
    const long stride = (long)gridDim.x * blockDim.x;
    const long base0 = blockIdx.x * (long)blockDim.x + (threadIdx.x & ~63);
    const long nround = (S.n + stride - 1) / stride * stride;
    double d1 = 0, h1 = 0;
    for (long base = base0; base < nround; base += stride) {
        const long i = base + (threadIdx.x & 63);
        const bool valid = i < S.n;
        double d = 0, hh = 0;
        int box = 0;
        if (valid) {
            TimeParams Ti = T;
            long long t0 = time0, t1 = time1_single;
            if (bt.b) {
                box = bt.BOX[i];
                const BoxDev &B = bt.b[box];
                resolve_time(bt, (int)i, Ti);
                t0 = B.ti_current;
                t1 = B.ti_next;
            }
            if (t1 != t0) d = drift_one(Ti, S, i, t0, t1, X0, Y0, Z0);
            hh = S.H[i];
        }
        if (bt.b) {
            wave_max_to(d, valid, box, disp, bt, kSlotDisp);
            wave_max_to(hh, valid, box, hmax, bt, kSlotHmax);
        } else {
            d1 = fmax(d1, d);
            h1 = fmax(h1, hh);
        }
    }
    if (!bt.b) {
        for (int o = 32; o > 0; o >>= 1) {
            d1 = fmax(d1, __shfl_xor(d1, o, 64));
            h1 = fmax(h1, __shfl_xor(h1, o, 64));
        }
        __shared__ double sm[4], sh[4];
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) { sm[w] = d1; sh[w] = h1; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int k = 1; k < (int)(blockDim.x >> 6); k++) { d1 = fmax(d1, sm[k]); h1 = fmax(h1, sh[k]); }
            if (d1 > 0) atomic_max_bits(disp, d1);
            if (h1 > 0) atomic_max_bits(hmax, h1);
        }
    }
}

__global__ void k_step_prep(int *__restrict__ counters, unsigned long long *__restrict__ u64c)
{
    step_prep(counters, u64c);
}

// find_next_sync_point_and_drift's timeline part (src/core/run.cc:857-930) from the occupation,
// one wavefront per box (block b = box b, lane k = bin k): the next kick (All.Ti_Current after
// the step) and the bins active there (single box: tl[0], tl[1]; a box not stepping keeps its
// time and has no active bin)
__global__ void __launch_bounds__(64) k_next_sync(TimeParams T, const unsigned long long *__restrict__ count,
                                                  unsigned long long *__restrict__ tl, BoxTable bt)
{
    const int b = blockIdx.x;
    const int k = threadIdx.x;
    long long tic = T.ti_current;
    if (bt.b) {
        tic = bt.b[b].ti_current;
        if (!bt.b[b].stepping) {
            if (k == 0) { bt.b[b].need_wakeup = 0; bt.b[b].ti_next = tic; bt.b[b].mask_next = 0; }
            return;
        }
    }
    const NextSync ns = next_sync_wave(T, count + (size_t)64 * b, tic);
    if (k != 0) return;
    if (bt.b) {
        bt.b[b].need_wakeup = 0;  // consumed by the wake-up pass of this step
        bt.b[b].ti_next = ns.ti;
        bt.b[b].mask_next = ns.mask;
    } else {
        tl[0] = (unsigned long long)ns.ti;
        tl[1] = ns.mask;
        tl[2] = ns.full ? 1ull : 0ull;
        tl[3] = ns.nf;
    }
}

// make_list_of_active_particles (src/core/run.cc:997-1009) as flags for the scan: a particle of
// an active bin with mass; its TimeBinActive byte (the force pass tests neighbours with it); the
// softening cache of the active particles (src/gravity/gravtree.cc:267); and the largest
// radius of the active particles after density()'s clamp (src/hydro/density_loop.cc:818-838)
__global__ void __launch_bounds__(256) k_active_flags(long n, const unsigned long long *__restrict__ tl, const int *__restrict__ TB,
                                                       const double *__restrict__ M, const double *__restrict__ H, double hclamp,
                                                       uint8_t *__restrict__ ACT, double *__restrict__ SOFT, int *__restrict__ flag,
                                                       unsigned long long *__restrict__ hmax_act, BoxTable bt)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const bool valid = i < n;
    double h = 0;
    int box = 0;
    if (valid) {
        unsigned long long mask;
        if (bt.b) {
            box = bt.BOX[i];
            mask = bt.b[box].mask_next;
            hclamp = 0.99 * bt.b[box].max_h;
        } else {
            mask = tl[1];
        }
        const bool on = bin_on(mask, TB[i]);
        ACT[i] = on ? 1 : 0;
        const bool row = on && (M[i] > 0);
        flag[i] = row ? 1 : 0;
        if (row) {
            h = H[i];
            SOFT[i] = h;
            if ((h < 0) || !isfinite(h) || (h > hclamp)) h = hclamp;
        }
    }
    if (i == n) flag[n] = 0;
    wave_max_to(h, valid, box, hmax_act, bt, kSlotHrows);
}

// The drift to the next sync point (drift_one, as k_drift does for boxes) fused with the first
// half of make_list_of_active_particles (src/core/run.cc:997-1009): block b takes the kActChunk
// particles of chunk b (thread t takes particles t, t+256, ...) and keeps each wavefront's 64
// row flags as a ballot and the block's row count; k_active_compact then writes each block's rows
// at its offset (the counts of the blocks before it) in ascending order.  No scan over every
// particle, and the flags read the drifted radius from registers instead of a second sweep.
// Which particles move: tl[2] set (the full drift of the step) every particle from its own
// P.Ti_current; else the particles of the active bins only (src/core/run.cc:979-994), and a
// particle of another bin reads only its bin.  disp receives the step's bound on how far any
// particle can move, drifted now or later (the bound on the kicked velocities times the step,
// padded for rounding): the caller sums these into its bound on the displacement since the last
// ordering, so no ordering positions are read.  hmax receives the largest radius drifted.
__global__ void __launch_bounds__(256) k_drift_active(TimeParams T, ParticleView S, long long time0, unsigned long long *__restrict__ tl,
                                                      unsigned long long *__restrict__ disp,
                                                      unsigned long long *__restrict__ hmax, double hclamp, uint8_t *__restrict__ ACT,
                                                      unsigned long long *__restrict__ bits, int *__restrict__ bcount,
                                                      unsigned long long *__restrict__ hmax_act,
                                                      const unsigned long long *__restrict__ occupation)
{
    __shared__ unsigned long long s_tl[3];
    const unsigned long long *count_bins = occupation;
    const long base = (long)blockIdx.x * kActChunk;
    const int tid = threadIdx.x, w = tid >> 6;
    // every time bin of the thread's particles loaded first, before any store (the stores may
    // alias them for the compiler, and a load issued after a store waits for it) and while the
    // next sync point is worked out: the sweep takes one memory round trip, not one per
    // particle; tbs[0] is the current particle's, shifted on
    constexpr int kPer = kActChunk / 256;
    int tbs[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const long i = base + k * 256 + tid;
        tbs[k] = i < S.n ? S.TB[i] : 0;
    }
    if (count_bins) {
        // the next sync point from the occupation, in every block (k_next_sync's work; the
        // first block also writes it for the host and the later launches)
        if (threadIdx.x < 64) {
            const NextSync ns = next_sync_wave(T, count_bins, T.ti_current);
            if (threadIdx.x == 0) {
                s_tl[0] = (unsigned long long)ns.ti;
                s_tl[1] = ns.mask;
                s_tl[2] = ns.full ? 1ull : 0ull;
                if (blockIdx.x == 0) { tl[0] = s_tl[0]; tl[1] = s_tl[1]; tl[2] = s_tl[2]; tl[3] = ns.nf; }
            }
        }
        __syncthreads();
    }
    const long long time1 = (long long)(count_bins ? s_tl[0] : tl[0]);
    const unsigned long long mask = count_bins ? s_tl[1] : tl[1];
    const bool full = (count_bins ? s_tl[2] : tl[2]) != 0;
    double d1 = 0, h1 = 0, hm = 0;
    int count = 0;
    for (int k = 0; k < kPer; k++) {
        const long i = base + k * 256 + tid;
        const int tb = tbs[0];
#pragma unroll
        for (int q = 0; q + 1 < kPer; q++) tbs[q] = tbs[q + 1];
        bool row = false;
        if (i < S.n) {
            // the drifted radius taken from the drift
            const bool on = bin_on(mask, tb);
            if (full || on) {
                const double m = S.M[i];
                const long long t0 = S.TIC ? S.TIC[i] : time0;
                double d = 0, hh;
                if (time1 != t0) {
                    d = drift_one(T, S, i, t0, time1, nullptr, nullptr, nullptr, &hh);
                    if (S.TIC) S.TIC[i] = time1;
                } else {
                    hh = S.H[i];
                }
                d1 = fmax(d1, d);
                h1 = fmax(h1, hh);
                row = on && (m > 0);
                if (row) {
                    S.SOFT[i] = hh;
                    double h = hh;
                    if ((h < 0) || !isfinite(h) || (h > hclamp)) h = hclamp;
                    hm = fmax(hm, h);
                }
            }
            ACT[i] = on ? 1 : 0;
        }
        const unsigned long long b = __ballot(row);
        if ((tid & 63) == 0) bits[(long)blockIdx.x * (kActChunk / 64) + k * 4 + w] = b;
        count += __popcll(b);
    }
    __shared__ int wcount[4];
    __shared__ double wd[4], wh[4], whm[4];
    for (int o = 32; o > 0; o >>= 1) {
        d1 = fmax(d1, __shfl_xor(d1, o, 64));
        h1 = fmax(h1, __shfl_xor(h1, o, 64));
        hm = fmax(hm, __shfl_xor(hm, o, 64));
    }
    if ((tid & 63) == 0) { wcount[w] = count; wd[w] = d1; wh[w] = h1; whm[w] = hm; }
    __syncthreads();
    if (tid == 0) {
        bcount[blockIdx.x] = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        d1 = fmax(fmax(wd[0], wd[1]), fmax(wd[2], wd[3]));
        h1 = fmax(fmax(wh[0], wh[1]), fmax(wh[2], wh[3]));
        hm = fmax(fmax(whm[0], whm[1]), fmax(whm[2], whm[3]));
        if (blockIdx.x == 0 && S.vkmax) {
            // every particle moves at most the bound on the kicked velocities per unit time
            const double vk = __longlong_as_double((long long)*S.vkmax);
            d1 = fmax(d1, vk * ((time1 - time0) * T.tbi) * (1.0 + 1e-12) + 1e-13 * T.box);
        }
        if (d1 > 0) atomic_max_bits(disp, d1);
        if (h1 > 0) atomic_max_bits(hmax, h1);
        if (hm > 0) atomic_max_bits(hmax_act, hm);
    }
}

__global__ void __launch_bounds__(256) k_active_compact(long n, int nblocks, const unsigned long long *__restrict__ bits,
                                                        const int *__restrict__ bcount, int *__restrict__ act, int *__restrict__ count_out,
                                                        DensityStart ds)
{
    __shared__ int spart[4];
    __shared__ int sprefix[kActChunk / 64 + 1];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    // a block without rows has nothing to place (the last block still writes the count)
    if (bcount[blockIdx.x] == 0 && (int)blockIdx.x != nblocks - 1) return;
    // the rows of the blocks before this one
    int before = 0;
    const int nprev = (int)blockIdx.x;
    for (int b = tid; b < nprev; b += 1024) {  // four loads in flight per thread
        const int v0 = bcount[b];
        const int v1 = (b + 256 < nprev) ? bcount[b + 256] : 0;
        const int v2 = (b + 512 < nprev) ? bcount[b + 512] : 0;
        const int v3 = (b + 768 < nprev) ? bcount[b + 768] : 0;
        before += (v0 + v1) + (v2 + v3);
    }
    for (int o = 32; o > 0; o >>= 1) before += __shfl_xor(before, o, 64);
    if (lane == 0) spart[w] = before;
    const unsigned long long *mb = bits + (long)blockIdx.x * (kActChunk / 64);
    if (tid == 0) {
        int acc = 0;
        for (int q = 0; q < kActChunk / 64; q++) { sprefix[q] = acc; acc += __popcll(mb[q]); }
        sprefix[kActChunk / 64] = acc;
    }
    __syncthreads();
    const int off = spart[0] + spart[1] + spart[2] + spart[3];
    const long base = (long)blockIdx.x * kActChunk;
    for (int k = 0; k < kActChunk / 256; k++) {
        const int q = k * 4 + w;
        const unsigned long long m = mb[q];
        if ((m >> lane) & 1ull) {
            const int below = __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
            const int r = off + sprefix[q] + below;
            const long i = base + k * 256 + tid;
            act[r] = (int)i;
            if (ds.H) {
                // density()'s pre-iteration radius clamp (src/hydro/density_loop.cc:818-838) and the
                // row's iteration state, as k_density_init sets them
                double h = ds.H[i];
                if ((h < 0) || !isfinite(h) || (h > ds.hclamp)) { h = ds.hclamp; ds.H[i] = h; }
                ds.hcur[r] = h;
                ds.left[r] = 0; ds.right[r] = 0; ds.cn_cur[r] = 0; ds.cn_init[r] = 0; ds.iter[r] = 0;
                ds.need[r] = 1;
            }
        }
    }
    if (blockIdx.x == nblocks - 1 && tid == 0) *count_out = off + sprefix[kActChunk / 64];
}

// the scanned flags to the active list; the count to count_out
__global__ void __launch_bounds__(256) k_compact_count(long n, const int *__restrict__ flag, const int *__restrict__ scan,
                                                        int *__restrict__ act, int *__restrict__ count_out)
{
    const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s < n && flag[s]) act[scan[s]] = (int)s;
    if (s == n) *count_out = scan[n];
}

// the boxes that stepped move to their next sync point (All.Ti_Current, All.Time, TimeBinActive)
__global__ void k_advance_boxes(BoxTable bt)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (!bt.b || b >= bt.nbox) return;
    BoxDev &B = bt.b[b];
    if (!B.stepping) return;
    B.ti_current = B.ti_next;
    B.mask = B.mask_next;
    B.time = B.time_begin + B.ti_current * B.tbi;
}

// TimeBinActive of each particle (the byte the force pass tests a neighbour with), and the
// softening cache of the active list (compute_all_force_softening(0),
// src/gravity/gravtree.cc:267)
__global__ void __launch_bounds__(256) k_active_from_bins(long n, unsigned long long mask, const int *__restrict__ TB,
                                                           const double *__restrict__ M, const double *__restrict__ H,
                                                           uint8_t *__restrict__ ACT, double *__restrict__ SOFT)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool on = bin_on(mask, TB[i]);
    ACT[i] = on ? 1 : 0;
    if (on && M[i] > 0) SOFT[i] = H[i];
}

// The particles a pass over the active rows may read, drifted to time1 before it reads them
// (the lazy drift's hook in the neighbour search, src/mesh/gpu_neighbor_list.cc:2443-2510: the
// members the walk found behind the search time are drifted to it).  The walks of this library
// visit the 27 cells around each row's home cell of the grid g, so those cells are the set here
// (a superset of the pairs accepted, as the reference's walk's is).  Rows in one region share
// most of their cells, so the cells are listed once each first (k_stencil_cells: a thread per
// row and stencil cell; the first to mark a cell appends it, one atomic per wavefront for the
// list), and one workgroup per listed cell drifts its particles that stand behind time1
// (k_drift_cells; cells are disjoint index ranges, so no particle has two writers) and clears
// the cell's mark.  hmax receives the largest radius drifted.
__global__ void __launch_bounds__(256) k_stencil_cells(GridView g, int nrows, const int *__restrict__ act,
                                                       const uint32_t *__restrict__ key, int *__restrict__ cmark,
                                                       int *__restrict__ cells, int *__restrict__ ncells)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    bool fresh = false;
    int c = 0;
    if (t < 27L * nrows) {
        const int r = (int)(t / 27), q = (int)(t % 27);
        const uint32_t home = key[act[r]] >> g.shift;
        const int nc = 1 << g.level;
        const int cx = compact10(home >> 2), cy = compact10(home >> 1), cz = compact10(home);
        const int dx = q / 9 - 1, dy = (q / 3) % 3 - 1, dz = q % 3 - 1;
        const int a = (cx + dx + nc) % nc, b = (cy + dy + nc) % nc, d = (cz + dz + nc) % nc;
        c = (int)((spread10(a) << 2) | (spread10(b) << 1) | spread10(d));
        fresh = atomicExch(&cmark[c], 1) == 0;
    }
    const unsigned long long won = __ballot(fresh);
    if (!won) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)won) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(ncells, __popcll(won));
    base = __shfl(base, leader, 64);
    if (fresh) cells[base + __popcll(won & ((1ull << lane) - 1))] = c;
}

__global__ void __launch_bounds__(256) k_drift_cells(TimeParams T, ParticleView S, GridView g, const int *__restrict__ cells,
                                                     const int *__restrict__ ncells, int *__restrict__ cmark, long long time1,
                                                     unsigned long long *__restrict__ hmax)
{
    __shared__ double wmax[4];
    const int count = *ncells;
    double h1 = 0;
    for (int k = blockIdx.x; k < count; k += gridDim.x) {
        const int c = cells[k];
        const int st = g.cstart[c], en = g.cend[c];
        for (int j = st + threadIdx.x; j < en; j += blockDim.x) {
            const long long t0 = S.TIC[j];
            if (t0 == time1) continue;
            double hh;
            drift_one(T, S, j, t0, time1, nullptr, nullptr, nullptr, &hh);
            S.TIC[j] = time1;
            h1 = fmax(h1, hh);
        }
        if (threadIdx.x == 0) cmark[c] = 0;
    }
    for (int o = 32; o > 0; o >>= 1) h1 = fmax(h1, __shfl_xor(h1, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = h1;
    __syncthreads();
    if (threadIdx.x == 0) {
        h1 = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
        if (h1 > 0) atomic_max_bits(hmax, h1);
    }
}

// every particle behind time1 drifted to it (gizmo_full_drift_to, src/core/predict.cc:148-171:
// before a re-ordering, which reads every position, and for a snapshot of the whole state)
__global__ void __launch_bounds__(256) k_drift_behind(TimeParams T, ParticleView S, long long time1, unsigned long long *__restrict__ hmax)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double h1 = 0;
    if (i < S.n) {
        const long long t0 = S.TIC[i];
        if (t0 != time1) {
            drift_one(T, S, i, t0, time1, nullptr, nullptr, nullptr, &h1);
            S.TIC[i] = time1;
        }
    }
    for (int o = 32; o > 0; o >>= 1) h1 = fmax(h1, __shfl_xor(h1, o, 64));
    if ((threadIdx.x & 63) == 0 && h1 > 0 && hmax) atomic_max_bits(hmax, h1);
}

// the bound on the kicked velocities over every particle (the start of a run)
__global__ void __launch_bounds__(256) k_vk_max(ParticleView S)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    raise_vk_bound(S, i < S.n ? i : 0, i < S.n);
}

}  // namespace gz