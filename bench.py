"""Benchmark of the hydro hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 50] [--warmup 2] [--n-side 128] [--no-cpu] [--no-full-pass]
                    [--no-serve] [--no-tune] [--no-legs] [--workload sedov|ensemble|lognormal]

Workload (the 2M-particle Sedov job): a 128^3 jittered lattice in a periodic unit box, cold
medium rho = 1, P = 1e-6, gamma 5/3, E = 1 deposited kernel-weighted into the 64 central
particles; Wendland C2 kernel, 64 neighbours; the timestep loop with adaptive power-of-two
bins (CourantFac 0.05, ErrTolIntAccuracy 0.002, MaxSizeTimestep 0.01 as in the reference's
own Sedov test, TimeMax 1).  A step is one sync point of the run loop: new steps and
wake-ups, the first half-kick, the drift of every particle to the next sync point, the hot
path (neighbour builds, density() with all its radius iterations, hydro_gradient_calc(),
hydro_force(), final ops) over the particles active there, and the second half-kick.
`value` is the active-particle updates per second of the timed steps (sum of the active
counts over the wall time), the job starting from gz_run_begin; the state is resident in HBM
for the whole timed region.  Warmup: W steps of an identical untimed job.

Both arithmetic forms run the same job: the plain FP64 form (the reported `value`, as the
job is the FP64-at-scale configuration) and the FP32-mixed form beside it.  `roofline` covers
the dominant kernel of the timed job: of the three pair kernels (density rounds, gradients,
force; the units of work of the pass) the one with the most launch time, its launches timed
with HIP events on the stream the kernels run on.  Its algorithmic bytes follow a pair-gather
model: rows x (S_i + S_out) + pairs x (S_j + 4), the rows' own fields in and results out, and
per pair the neighbour's fields and a 4 B list index, pairs counted on the device (PAIR_MODEL,
FORCE_MODEL below).  With a few hundred active rows per step these launches are short and
latency-bound.  `roofline_force` is the force kernel over the same job whatever dominates.
`full_pass` adds one all-active pass of the hot path over the 2M particles (the kernels at
full occupancy) and `roofline_force_full_pass` the force kernel there.
Before each form's timed job its launch options (few_rows, row_schedule, speculate: shapes
only, never results) are autotuned on the job's first 10 sync points (`autotune`,
gizmo_amd/tune.py; --no-tune keeps the library's defaults).
The serving loop (`serving`, p50/p99 job latency), compact legs of the other two workloads and
the CPU baseline follow: `ensemble` (48 boxes of the 256-box mix, 30 steps: batched FP32-mixed,
batched FP64 and looped FP64, so that batching and arithmetic form are reported apart) and
`lognormal` (1M particles of the lognormal field, one pass per form and the FP32-mixed errors
against FP64); --no-legs skips them, --workload ensemble|lognormal runs them at full size.
The CPU baseline is the FP64 oracle (C + OpenMP) running the first steps of the same job on
this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "hydro active-particle updates/s at 2M (mixed vs FP64); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
# pair-gather model of the force pass, bytes per particle value-set
FORCE_MODEL = {"fp64": dict(S_i=288, S_out=40, S_j=288), "mixed": dict(S_i=144, S_out=20, S_j=144)}
INTEG = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
             time_begin=0.0, time_max=1.0, min_egy_spec=0.0, domain_build_active_fraction=0.5)
# launch families the library times with HIP events on its stream (gz_timing)
FAMILIES = ("timeline", "drift", "sort", "ngb_build", "density", "density_after", "pack_grad", "gradients", "pack_force",
            "force")
# algorithmic bytes of the three pair kernels (per row: the row's own fields in and its results
# out; per pair: the neighbour's fields and its list index), FP64 / FP32-mixed:
#   density (k_density_rows): row 56 B in (position, velocity, radius) + 104 B out (13 sums);
#     pair 60 B (position, velocity, mass, a 4 B index), the same in both forms (FP64 state);
#   gradients (k_gradients_rows): row 136 B in + 208 B out (15 gradients, 10 extrema, distance,
#     written as FP64); pair 84 B (position 24 B + 7 fields 56 B + index);
#   force: the FORCE_MODEL above.
PAIR_MODEL = {"density": dict(S_i=56, S_out=104, S_j=56), "gradients": dict(S_i=136, S_out=208, S_j=80)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def physical_cores():
    try:
        cores = set()
        phys = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                cores.add((phys, line.split(":")[1].strip()))
        return len(cores) or os.cpu_count()
    except OSError:
        return os.cpu_count()


def new_hydro(ic, precision, options=None):
    from gizmo_amd.hydro import Hydro
    g = Hydro(dict(ic["params"]), len(ic["mass"]), precision)
    for k, v in (options or {}).items():
        g.set_option(k, v)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    return g


def new_job(ic, precision, options=None):
    g = new_hydro(ic, precision, options)
    g.run_begin(INTEG)
    return g


def tune_job(ic, precision, steps=10):
    """The job's launch options (few_rows, row_schedule, speculate) autotuned on its first
    `steps` sync points (gizmo_amd/tune.py), each candidate a fresh job."""
    from gizmo_amd.tune import SINGLE_BOX_SPACE, job_sample, tune
    best, table = tune(job_sample(lambda: new_hydro(ic, precision), INTEG, steps), SINGLE_BOX_SPACE, reps=2, log=log)
    return {"options": best, "sample": f"sync points 1..{steps} of the same job, best of 2",
            "table_ms": [[o, round(1e3 * t, 3)] for o, t in table]}


def run_job(ic, precision, steps, warmup, options=None):
    """The timed job (no events on the stream), then the same job again with HIP events around
    every launch family for the per-kernel times and the roofline."""
    if warmup > 0:
        g = new_job(ic, precision, options)
        g.run_steps(warmup)
        g.close()
    g = new_job(ic, precision, options)
    g.synchronize()
    g.timing_reset()
    t0 = time.perf_counter()
    active, woken = g.run_steps(steps)  # run()'s loop over the sync points, in the library
    g.synchronize()
    dt = time.perf_counter() - t0
    res = dict(seconds=dt, active=active, woken=woken, resorts=g.stat("resorts"), updates=g.stat("active_updates"),
               tl=g.timeline())
    g.close()
    g = new_job(ic, precision, options)
    g.synchronize()
    g.timing_enable(True)
    g.timing_reset()
    active2, _ = g.run_steps(steps)
    g.synchronize()
    assert active2 == active, "the profiled job took other steps than the timed one"
    res["kernels"] = {k: g.timing(k) for k in FAMILIES}
    res["pairs"] = g.stat("force_pairs_total")
    res["pairs_density"] = g.stat("density_pairs_total")
    g.timing_enable(False)
    g.close()
    return res


def full_pass(ic, precision, reps=3):
    """All particles active through one hot-path pass (after the radii converged)."""
    from gizmo_amd.hydro import Hydro
    g = Hydro(dict(ic["params"]), len(ic["mass"]), precision)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    for _ in range(2):
        g.hydro_pass()
    g.synchronize()
    g.timing_enable(True)
    g.timing_reset()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.hydro_pass()
    g.synchronize()
    dt = (time.perf_counter() - t0) / reps
    n_act = g.stat("n_active")
    kern = {k: round(g.timing(k)[1] / reps, 3) for k in FAMILIES if g.timing(k)[0]}
    pairs = g.stat("force_pairs_total") / reps
    g.close()
    return dict(ms=round(1e3 * dt, 3), updates_per_s=round(n_act / dt, 1), kernel_ms=kern, n_active=n_act, pairs=pairs)


def roofline(res, form):
    """The force kernel over the timed job: pair-gather bytes of all its launches over their time."""
    m = FORCE_MODEL[form]
    launches, total_ms = res["kernels"]["force"]
    n_act = sum(res["active"])
    bytes_total = n_act * (m["S_i"] + m["S_out"]) + res["pairs"] * (m["S_j"] + 4)
    achieved = bytes_total / (total_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_force_rows<6,{'double' if form == 'fp64' else 'float'}>",
            "launches": launches, "bytes_per_launch": int(bytes_total / max(launches, 1)),
            "avg_launch_ms": round(total_ms / max(launches, 1), 4)}


def roofline_pair_kernel(res, form, family):
    """A pair kernel of the timed job (the unit of work: density, gradients or force) over all
    its launches: algorithmic bytes (rows x (S_i + S_out) + pairs x (S_j + 4), pairs counted on
    the device) over the launches' HIP-event time.  The job's rows per launch are a few hundred,
    so these launches are latency-bound."""
    if family == "force":
        return roofline(res, form)
    m = PAIR_MODEL[family]
    launches, total_ms = res["kernels"][family]
    if family == "density":
        # rows: every active row of each step times the step's rounds (an upper bound: rows that
        # converged earlier skip later rounds), pairs: counted by the row kernels
        rows = sum(res["active"]) * launches / max(len(res["active"]), 1)
        pairs = res["pairs_density"]
        name = f"k_density_rows<6,{'double' if form == 'fp64' else 'float'},256>"
    else:
        rows = sum(res["active"])
        pairs = res["pairs"]
        name = f"k_gradients_rows<6,{'double' if form == 'fp64' else 'float'},128>"
    bytes_total = rows * (m["S_i"] + m["S_out"]) + pairs * (m["S_j"] + 4)
    achieved = bytes_total / (total_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name, "launches": launches,
            "bytes_per_launch": int(bytes_total / max(launches, 1)), "avg_launch_ms": round(total_ms / max(launches, 1), 4),
            "bytes_model": f"rows x ({m['S_i']} + {m['S_out']}) B + pairs x ({m['S_j']} + 4) B"}


def roofline_dominant(res, form):
    """The pair kernel that takes the most time in the timed job."""
    fam = max(("density", "gradients", "force"), key=lambda k: res["kernels"][k][1])
    return roofline_pair_kernel(res, form, fam)


def roofline_full_pass(fp, form):
    """The force kernel at full occupancy (every particle active, lane-per-row form)."""
    m = FORCE_MODEL[form]
    ms = fp["kernel_ms"]["force"]
    bytes_total = fp["n_active"] * (m["S_i"] + m["S_out"]) + fp["pairs"] * (m["S_j"] + 4)
    achieved = bytes_total / (ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": f"k_force<6,{'double' if form == 'fp64' else 'float'}>",
            "launches": 1, "bytes_per_launch": int(bytes_total), "avg_launch_ms": ms}


def serve_bench(precision, n_jobs=48, seed=1005, capacity=4 * 1024 * 1024, max_boxes=64, load=0.8, calib_jobs=8):
    """The serving loop: a Poisson stream of n_jobs jobs (N log-uniform in [8192, 1048576], rms
    Mach uniform in [0.1, 5], steps uniform in [10, 200]) into a batch with a budget of
    `capacity` resident particles.  The arrival rate is set to `load` times the device's
    capacity for this job mix, measured first on calib_jobs other jobs of the same mix run
    together (particle-steps per second).  Every request's particles are made before the clock
    starts (the request carries them); admission, upload, start, steps, read-back of the
    result and removal are inside the latencies."""
    from gizmo_amd.serve import Server, draw_jobs, latency_summary, poisson_arrivals
    jobs = draw_jobs(n_jobs, seed)
    cal = draw_jobs(calib_jobs, seed + 1)
    for j in jobs + cal:
        j.fields = j.ic()
    for j in cal:
        j.arrival = 0.0
    s = Server(capacity, max_boxes, precision, keep=["u"])
    wall_cal = s.run(cal)
    ps_rate = sum(j.n * j.steps for j in cal) / wall_cal
    mean_work = float(np.mean([j.n * j.steps for j in jobs]))
    rate = load * ps_rate / mean_work
    poisson_arrivals(jobs, rate, seed)
    log(f"serve: calibration {wall_cal:.2f} s ({ps_rate / 1e6:.1f} M particle-steps/s); arrival rate {rate:.2f} jobs/s")
    wall = s.run(jobs)
    dev_bytes = s.batch.stat("device_bytes")
    s.close()
    out = latency_summary(jobs, wall)
    out.update(arrival_rate_per_s=round(rate, 4), target_load=load, capacity_particles=capacity,
               device_bytes=int(dev_bytes), bytes_per_particle=round(dev_bytes / capacity, 1),
               calibration_particle_steps_per_s=round(ps_rate, 1),
               max_queue_wait_s=round(max(j.admitted - j.arrival for j in jobs), 4),
               jobs_n=[j.n for j in jobs], jobs_steps=[j.steps for j in jobs])
    for k in ("p50_s", "p99_s", "mean_s", "max_s", "wall_s"):
        out[k] = round(out[k], 4)
    out["updates_per_s"] = round(out["updates_per_s"], 1)
    return out


def ensemble_boxes(n_boxes=256, seed=1004):
    """256 independent boxes: N log-uniform in [4096, 262144] (the nearest lattice cube), rms
    Mach uniform in [0.1, 5], gamma 5/3, Wendland C2 with 64 neighbours."""
    from gizmo_amd.serve import draw_jobs
    return draw_jobs(n_boxes, seed, n_min=4096, n_max=262144, steps=(100, 100))


def ensemble_bench(steps=100, n_boxes=256, seed=1004, tune_sample=(12, 5)):
    """The ensemble: every box started together in one batch and advanced `steps` sync points,
    in the FP32-mixed and the plain FP64 form, against looping the plain FP64 form box by box
    (one context, begin and `steps` steps per box): batched FP64 over looped FP64 is the gain
    of batching alone, batched mixed over batched FP64 that of the arithmetic form.  The
    batch's launch shape (few_rows) is autotuned first on a sample (the first boxes, a few
    steps, gizmo_amd/tune.py).  Every leg times the steps only; the start of each box is
    untimed."""
    from gizmo_amd.batch import Batch
    from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64, Hydro
    from gizmo_amd.serve import JOB_INTEG, job_params
    from gizmo_amd.tune import BATCH_SPACE, batch_sample, tune
    boxes = ensemble_boxes(n_boxes, seed)
    fields = [j.ic() for j in boxes]
    total = sum(j.n for j in boxes)
    log(f"ensemble: {n_boxes} boxes, {total} particles")

    def make_batch(prec, flist):
        def make():
            b = Batch(job_params(), sum(len(f["mass"]) for f in flist), len(flist), prec)
            for f in flist:
                b.add(f, job_params(), JOB_INTEG)
            return b
        return make

    tuned = {}
    if tune_sample:
        nb, ns = tune_sample
        best, table = tune(batch_sample(make_batch(GZ_FP32_MIXED, fields[:nb]), ns), BATCH_SPACE, reps=1, log=log)
        tuned = {"options": best, "sample": f"first {nb} boxes, {ns} steps, FP32-mixed",
                 "table_ms": [[o, round(1e3 * t, 3)] for o, t in table]}

    def batched(prec, name):
        b = make_batch(prec, fields)()
        for k, v in tuned.get("options", {}).items():
            b.set_option(k, v)
        t0 = time.perf_counter()
        b.begin()
        b.synchronize()
        t_begin = time.perf_counter() - t0
        u0 = b.stat("active_updates")
        t0 = time.perf_counter()
        for _ in range(steps):
            b.step()
        b.synchronize()
        t_b = time.perf_counter() - t0
        upd = b.stat("active_updates") - u0
        dev_bytes = b.stat("device_bytes")
        b.close()
        log(f"ensemble batched {name}: {t_b:.2f} s, {upd / t_b / 1e6:.1f} M updates/s (start {t_begin:.2f} s)")
        return {"seconds": round(t_b, 3), "updates": int(upd), "updates_per_s": round(upd / t_b, 1),
                "start_seconds": round(t_begin, 3), "device_bytes": int(dev_bytes)}

    bm = batched(GZ_FP32_MIXED, "mixed")
    bf = batched(GZ_FP64, "fp64")
    t_l, upd_l = 0.0, 0.0
    for k, f in enumerate(fields):
        g = Hydro(job_params(), len(f["mass"]), GZ_FP64)
        g.load(f)
        g.run_begin(JOB_INTEG)
        g.synchronize()
        u0 = g.stat("active_updates")
        t0 = time.perf_counter()
        g.run_steps(steps)
        g.synchronize()
        t_l += time.perf_counter() - t0
        upd_l += g.stat("active_updates") - u0
        g.close()
    log(f"ensemble looped fp64: {t_l:.2f} s, {upd_l / t_l / 1e6:.1f} M updates/s")
    lf = {"seconds": round(t_l, 3), "updates": int(upd_l), "updates_per_s": round(upd_l / t_l, 1)}
    return {"boxes": n_boxes, "particles": total, "steps": steps, "autotune": tuned,
            "batched_mixed": bm, "batched_fp64": bf, "looped_fp64": lf,
            "batched_fp64_over_looped_fp64": round(bf["updates_per_s"] / lf["updates_per_s"], 3),
            "batched_mixed_over_batched_fp64": round(bm["updates_per_s"] / bf["updates_per_s"], 3),
            "batched_mixed_over_looped_fp64": round(bm["updates_per_s"] / lf["updates_per_s"], 3)}


LOGNORMAL_QUANTITIES = {"h": "h", "rho": "rho", "grad_rho": "grad_rho", "grad_p": "grad_p", "grad_v": "grad_v",
                        "dv_dt": "accel", "du_dt": "dtu"}


def lognormal_bench(n_particles=4 * 1024 * 1024, ng=256, reps=2):
    """The FP32-mixed form against the plain FP64 form on identical inputs: particles sampled
    from a lognormal density field (sigma(ln rho) = 2, P(k) ~ k^-3 on an ng^3 grid), isothermal
    (gamma 1.001, P = rho), solenoidal velocities at rms Mach 2; one hot-path pass per form
    after the radii converged, with the search over many grid levels.  Reports per-quantity
    errors of the mixed form against FP64 (max and median over particles, each relative to the
    quantity's largest magnitude, and the median per-particle relative error) and both rates."""
    from gizmo_amd import ics
    from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64, Hydro
    t0 = time.perf_counter()
    ic = ics.lognormal_box(n_particles, 1003, ng=ng)
    log(f"lognormal: {n_particles} particles built in {time.perf_counter() - t0:.1f} s")
    par = dict(ic["params"])
    # the radii converged once (FP64), then both forms start from them
    g = Hydro(par, n_particles, GZ_FP64)
    g.set_option("wide_search", 1)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    t0 = time.perf_counter()
    g.density()
    g.synchronize()
    log(f"lognormal: first density {time.perf_counter() - t0:.2f} s, {g.stat('density_iterations'):.0f} rounds")
    h0 = g.store(["h"])["h"]
    g.close()
    fields = dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=h0)
    out, rate = {}, {}
    names = sorted(set(LOGNORMAL_QUANTITIES.values()))
    for form, prec in (("fp64", GZ_FP64), ("mixed", GZ_FP32_MIXED)):
        g = Hydro(par, n_particles, prec)
        g.set_option("wide_search", 1)
        g.load(fields)
        g.hydro_pass()  # untimed: allocations and tables
        g.synchronize()
        # timed passes from the same inputs (their upload stays outside)
        dts = []
        for _ in range(reps):
            g.load(fields)
            g.synchronize()
            t1 = time.perf_counter()
            g.hydro_pass()
            g.synchronize()
            dts.append(time.perf_counter() - t1)
        dt = min(dts)
        rate[form] = dict(ms=round(1e3 * dt, 2), updates_per_s=round(n_particles / dt, 1),
                          density_iterations=g.stat("density_iterations"),
                          force_rare_rows_all_passes=g.stat("force_rare_rows"))
        out[form] = g.store(names)
        g.close()
        log(f"lognormal {form}: {1e3 * dt:.1f} ms per pass, {n_particles / dt / 1e6:.1f} M updates/s")
    errs = {}
    for q, k in LOGNORMAL_QUANTITIES.items():
        a = np.asarray(out["mixed"][k], dtype=np.float64).reshape(n_particles, -1)
        b = np.asarray(out["fp64"][k], dtype=np.float64).reshape(n_particles, -1)
        scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
        e_scale = (np.abs(a - b) / scale).max(axis=1)
        mag = np.maximum(np.sqrt((b * b).sum(axis=1)), 1e-300)
        e_part = np.sqrt(((a - b) ** 2).sum(axis=1)) / mag
        errs[q] = {"max_rel_to_scale": float(e_scale.max()), "median_rel_to_scale": float(np.median(e_scale)),
                   "median_rel_per_particle": float(np.median(e_part)),
                   "p99_rel_per_particle": float(np.percentile(e_part, 99))}
    hq = out["fp64"]["h"]
    return {"particles": n_particles, "grid": ng, "h_min": float(hq.min()), "h_max": float(hq.max()),
            "h_span": float(hq.max() / hq.min()), "fp64": rate["fp64"], "mixed": rate["mixed"],
            "mixed_over_fp64": round(rate["fp64"]["ms"] / rate["mixed"]["ms"], 3), "errors_mixed_vs_fp64": errs}


def cpu_baseline(ic, steps, target_s=20.0):
    """The oracle (C + OpenMP, FP64) on the same job: run_begin, then sync points until the
    step budget or the time budget is spent; the rate counts the timed steps only."""
    from oracle import oracle as O
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    r = O.Run(par, INTEG, st)
    t0 = time.perf_counter()
    r.begin()
    log(f"cpu oracle: run_begin {time.perf_counter() - t0:.1f} s")
    n_upd, k = 0, 0
    t0 = time.perf_counter()
    while k < steps and (time.perf_counter() - t0) < target_s:
        n_upd += r.step()
        k += 1
    dt = time.perf_counter() - t0
    log(f"cpu oracle: {k} steps, {n_upd} updates in {dt:.1f} s")
    thr = O.num_threads()
    return {"value": round(n_upd / dt, 1), "unit": "active-particle updates/s", "cores": thr, "kind": "port",
            "sample": f"steps 1..{k} of the same {len(st.mass)}-particle Sedov job ({n_upd} active-particle updates in "
                      f"{dt:.1f} s after the untimed start; {thr} OpenMP threads on {physical_cores()} physical cores)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-side", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full-pass", action="store_true")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop latency runs")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--no-tune", action="store_true", help="the library's default launch options, no autotuning")
    ap.add_argument("--no-legs", action="store_true", help="skip the compact ensemble and lognormal legs")
    ap.add_argument("--workload", choices=["sedov", "ensemble", "lognormal"], default="sedov",
                    help="sedov: the 2M Sedov job (+ serving); ensemble: 256 boxes batched vs looped; "
                         "lognormal: 4M lognormal field, FP32-mixed against FP64")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {a.gpus}); "
              "a single box shards by spatial domain and independent boxes shard by job, neither is built here",
              file=sys.stderr)
        sys.exit(2)

    from gizmo_amd import ics
    from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64

    if a.workload == "ensemble":
        e = ensemble_bench(steps=a.steps if a.steps != 50 else 100)
        print(json.dumps({
            "metric": "hydro active-particle updates/s, 256-box ensemble (batched FP32-mixed vs looped FP64)",
            "value": e["batched_mixed"]["updates_per_s"], "unit": "active-particle updates/s", "n_gpus": 1,
            "steps": e["steps"], "warmup": 0, "ms_per_step": round(1e3 * e["batched_mixed"]["seconds"] / e["steps"], 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32-mixed",
            "data": "synthetic: seeded jittered lattices with Gaussian velocities (numpy Philox)",
            "config": {"workload": "ensemble_256_boxes_N_loguniform_4096_262144_mach_0.1_5_100_steps"},
            "ensemble": e}), flush=True)
        return

    if a.workload == "lognormal":
        e = lognormal_bench()
        print(json.dumps({
            "metric": "hydro active-particle updates/s, 4M lognormal field (FP32-mixed vs FP64, per-quantity error)",
            "value": e["mixed"]["updates_per_s"], "unit": "active-particle updates/s", "n_gpus": 1, "steps": 1,
            "warmup": 1, "ms_per_step": e["mixed"]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32-mixed",
            "data": "synthetic: seeded lognormal density field, solenoidal Mach-2 velocities (numpy Philox)",
            "config": {"workload": "lognormal_4M_sigma2_k-3_256grid_isothermal_mach2_one_pass"},
            "lognormal": e}), flush=True)
        return

    t0 = time.perf_counter()
    ic = ics.sedov_box(a.n_side, 1002, kernel=6, des_num_ngb=64.0)
    n = len(ic["mass"])
    log(f"workload: {a.n_side}^3 = {n} particles, built in {time.perf_counter() - t0:.1f} s")

    forms = {"fp64": GZ_FP64, "mixed": GZ_FP32_MIXED}
    todo = ["fp64", "mixed"] if a.only == "both" else [a.only]
    res, fp, tuned = {}, {}, {}
    for f in todo:
        if not a.no_tune:
            tuned[f] = tune_job(ic, forms[f])
            log(f"{f}: tuned options {tuned[f]['options']}")
        log(f"{f}: {a.steps} steps ...")
        res[f] = run_job(ic, forms[f], a.steps, a.warmup, tuned.get(f, {}).get("options"))
        r = res[f]
        log(f"{f}: {1e3 * r['seconds']:.1f} ms, {sum(r['active'])} updates, resorts {r['resorts']:.0f}, "
            f"kernels {{{', '.join(f'{k}: {v[1]:.1f}' for k, v in r['kernels'].items())}}} ms")
        if not a.no_full_pass:
            fp[f] = full_pass(ic, forms[f])
            log(f"{f}: full pass {fp[f]['ms']} ms, {fp[f]['updates_per_s'] / 1e6:.1f} M updates/s")

    def summary(r):
        ups = sum(r["active"]) / r["seconds"]
        return {"value": round(ups, 1), "ms_per_step": round(1e3 * r["seconds"] / a.steps, 3),
                "updates": int(sum(r["active"])),
                "kernel_ms": {k: round(v[1], 3) for k, v in r["kernels"].items() if v[0]}}

    head = "fp64" if "fp64" in res else "mixed"
    r = res[head]
    out = {
        "metric": METRIC,
        "value": summary(r)["value"],
        "unit": "active-particle updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": summary(r)["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "fp32-mixed (fp64 positions, 3x3 inverse, h test)",
        "data": "synthetic: seeded Sedov blast in a jittered lattice (numpy Philox)",
        "config": {"workload": f"sedov_{a.n_side}^3_wendland_c2_64ngb_adaptive_timebins_{a.steps}_steps",
                   "n_particles": n, "kernel": "wendland_c2", "des_num_ngb": 64, "integ": INTEG,
                   "step": "one sync point: timesteps + wake-ups, kick, drift, neighbour builds, density (all radius "
                           "iterations), gradients, force, final ops, kick"},
        "active_per_step": r["active"],
        "woken_per_step": r["woken"],
        "resorts": r["resorts"],
        "kernel_ms": summary(r)["kernel_ms"],
        "roofline": roofline_dominant(r, head),
        "roofline_force": roofline(r, head),
        "p50_ms": None, "p99_ms": None,
    }
    for f in res:
        if f != head:
            out[f] = dict(summary(res[f]), roofline=roofline_dominant(res[f], f), roofline_force=roofline(res[f], f))
    if tuned:
        out["autotune"] = tuned
    if "fp64" in res and "mixed" in res:
        out["mixed_over_fp64"] = round(summary(res["mixed"])["value"] / summary(res["fp64"])["value"], 3)
    if fp:
        out["full_pass"] = fp
        out["roofline_force_full_pass"] = {f: roofline_full_pass(fp[f], f) for f in fp}
        if "fp64" in fp and "mixed" in fp:
            out["full_pass_mixed_over_fp64"] = round(fp["mixed"]["updates_per_s"] / fp["fp64"]["updates_per_s"], 3)
    if not a.no_serve:
        serving = {}
        for f in todo:
            log(f"serve {f}: {a.serve_jobs} jobs ...")
            serving[f] = serve_bench(forms[f], n_jobs=a.serve_jobs)
            sv = serving[f]
            log(f"serve {f}: p50 {sv['p50_s']:.3f} s, p99 {sv['p99_s']:.3f} s, {sv['updates_per_s'] / 1e6:.1f} M updates/s")
        out["serving"] = serving
        sv = serving["mixed" if "mixed" in serving else head]
        out["p50_ms"] = round(1e3 * sv["p50_s"], 1)
        out["p99_ms"] = round(1e3 * sv["p99_s"], 1)
        out["latency_form"] = "mixed" if "mixed" in serving else head
    if not a.no_legs:
        # compact legs of the other two workloads, timed by the same run
        out["ensemble"] = ensemble_bench(steps=30, n_boxes=48, tune_sample=None if a.no_tune else (8, 5))
        out["lognormal"] = lognormal_bench(n_particles=1 << 20, ng=128)
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(ic, a.steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
