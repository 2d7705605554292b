"""Benchmark of the hydro hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n-side 128] [--no-cpu]

Workload: the 2M-particle Sedov blast (128^3 jittered lattice in a periodic unit box, cold
medium rho = 1, P = 1e-6, gamma 5/3, E = 1 deposited kernel-weighted into the 64 central
particles; Wendland C2 kernel, 64 neighbours).  A step is one pass of the hot path over every
particle: the neighbour builds, density() with all its kernel-radius iterations,
hydro_gradient_calc(), hydro_force() and the final operations.  The warmup passes converge the
radii from the initial guess; the timed steps then start each density() from the converged
radius of the previous pass, as consecutive timesteps do.  Inputs are resident in HBM for
the whole timed region.

Both arithmetic forms are timed (FP32-mixed is the reported `value`, plain FP64 beside it).
The roofline entry covers the force kernel, the dominant one: algorithmic bytes per launch
(pair-gather model: n_active (S_i + S_out) + n_pairs (S_j + 4), with S_i = S_j = 288 B and
S_out = 40 B per value-set in FP64, 144/20 B in FP32) over its average launch time, measured
with HIP events on the stream the kernels run on.  The CPU baseline is the FP64 oracle (C +
OpenMP) timed on this host on a random sample of the same state, as active particles.
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
# pair-gather model of the force pass, bytes per value-set
FORCE_MODEL = {"fp64": dict(S_i=288, S_out=40, S_j=288), "mixed": dict(S_i=144, S_out=20, S_j=144)}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def physical_cores():
    try:
        cores = set()
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                core = line.split(":")[1].strip()
                cores.add((phys, core))
        return len(cores) or os.cpu_count()
    except OSError:
        return os.cpu_count()


def run_gpu(ic, precision, steps, warmup):
    from gizmo_amd.hydro import Hydro
    par = dict(ic["params"])
    n = len(ic["mass"])
    g = Hydro(par, n, precision)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    for w in range(warmup):
        t0 = time.perf_counter()
        g.hydro_pass()
        g.synchronize()
        log(f"warmup {w + 1}/{warmup}: {1e3 * (time.perf_counter() - t0):.1f} ms, "
            f"density iterations {g.stat('density_iterations'):.0f}")
    g.timing_enable(True)
    g.timing_reset()
    g.synchronize()
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        g.hydro_pass()
        iters.append(g.stat("density_iterations"))
    g.synchronize()
    dt = time.perf_counter() - t0
    n_act = g.stat("n_active")
    pairs = g.stat("pairs_symmetric")
    kern = {k: g.timing(k) for k in ("ngb_build", "density", "gradients", "force")}
    g.timing_enable(False)
    state = g.store() if precision == 0 else None
    g.close()
    return dict(seconds=dt, n_active=n_act, pairs=pairs, kernels=kern, density_iterations=iters, state=state)


def roofline(res, form):
    m = FORCE_MODEL[form]
    launches, total_ms = res["kernels"]["force"]
    avg_s = total_ms / max(launches, 1) / 1e3
    bytes_per_launch = res["n_active"] * (m["S_i"] + m["S_out"]) + res["pairs"] * (m["S_j"] + 4)
    achieved = bytes_per_launch / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_force<6,{'double' if form == 'fp64' else 'float'}>",
            "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_s * 1e3, 3)}


def cpu_baseline(ic, state, n_side, target_s=12.0):
    from oracle import oracle as O
    par = dict(ic["params"])
    st = O.State(len(ic["mass"]))
    for k in O.FIELDS:
        if k in state:
            getattr(st, k)[...] = state[k]
    rng = np.random.default_rng(12345)
    frac = 0.002
    while True:
        active = (rng.uniform(size=st.n) < frac).astype(np.uint8)
        s = st.copy()
        t0 = time.perf_counter()
        O.hydro_pass(par, s, active=active)
        dt = time.perf_counter() - t0
        log(f"cpu oracle: {int(active.sum())} active particles in {dt:.2f} s")
        if dt > 0.5 * target_s or frac >= 1.0:
            break
        frac = min(1.0, frac * max(2.0, 0.8 * target_s / max(dt, 1e-3)))
    n_act = int(active.sum())
    return {"value": round(n_act / dt, 1), "unit": "active-particle updates/s", "cores": O.num_threads(),
            "kind": "port", "sample": f"{n_act} randomly chosen active particles of the converged {n_side}^3 Sedov state "
                                      f"(one density+gradients+force pass, {dt:.1f} s, "
                                      f"{O.num_threads()} OpenMP threads, {physical_cores()} physical cores)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-side", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["both", "fp64", "mixed"], default="both")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for {a.gpus}); "
              "a single box shards by spatial domain and independent boxes shard by job, neither is built here",
              file=sys.stderr)
        sys.exit(2)

    from gizmo_amd import ics
    from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64

    t0 = time.perf_counter()
    ic = ics.sedov_box(a.n_side, 1002, kernel=6, des_num_ngb=64.0)
    n = len(ic["mass"])
    log(f"workload: {a.n_side}^3 = {n} particles, built in {time.perf_counter() - t0:.1f} s")

    res = {}
    if a.only in ("both", "fp64"):
        log("fp64 ...")
        res["fp64"] = run_gpu(ic, GZ_FP64, a.steps, a.warmup)
    if a.only in ("both", "mixed"):
        log("fp32-mixed ...")
        res["mixed"] = run_gpu(ic, GZ_FP32_MIXED, a.steps, a.warmup)
    head = "mixed" if "mixed" in res else "fp64"
    r = res[head]
    ups = r["n_active"] * a.steps / r["seconds"]
    out = {
        "metric": METRIC,
        "value": round(ups, 1),
        "unit": "active-particle updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * r["seconds"] / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 positions, 3x3 inverse, h test)" if head == "mixed" else "f64",
        "data": "synthetic: seeded Sedov blast in a jittered lattice (numpy Philox)",
        "config": {"workload": f"sedov_{a.n_side}^3_wendland_c2_64ngb_all_active_hot_path_pass", "n_particles": n,
                   "kernel": "wendland_c2", "des_num_ngb": 64,
                   "step": "neighbour builds + density (all radius iterations) + gradients + force + final ops"},
        "density_iterations_per_step": r["density_iterations"],
        "kernel_ms_per_step": {k: round(v[1] / a.steps, 3) for k, v in r["kernels"].items()},
        "pairs_symmetric": r["pairs"],
        "roofline": roofline(r, head),
        "p50_ms": None, "p99_ms": None,
    }
    if "fp64" in res and "mixed" in res:
        f = res["fp64"]
        ups64 = f["n_active"] * a.steps / f["seconds"]
        out["fp64"] = {"value": round(ups64, 1), "ms_per_step": round(1e3 * f["seconds"] / a.steps, 3),
                       "kernel_ms_per_step": {k: round(v[1] / a.steps, 3) for k, v in f["kernels"].items()},
                       "roofline": roofline(f, "fp64")}
        out["mixed_over_fp64"] = round(ups / ups64, 3)
    if not a.no_cpu and "fp64" in res:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(ic, res["fp64"]["state"], a.n_side)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
