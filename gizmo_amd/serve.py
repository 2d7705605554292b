"""A request loop over a batch of boxes: continuous batching of simulation jobs on one GPU.

A job is a seeded periodic box (jittered lattice of N particles with a Gaussian velocity field
of a given rms Mach number) to be advanced by a given number of sync points of its own
timestep loop.  Jobs arrive over time; the server admits them in arrival order as long as the
device memory budget (particles) and the box slots allow, starts them together
(``Batch.begin``), advances every resident job with one shared ``Batch.step`` per iteration, and
retires a job as soon as it has taken its steps, storing its result and freeing its room for
the next arrivals.  A job's result does not depend on what else was on the device.

Latency of a job = completion - arrival (wall clock); throughput = active-particle updates
of all jobs / wall time of the serving run.
"""
import math
import time

import numpy as np

from . import ics
from .batch import Batch
from .hydro import GZ_FP64

# the shared hot-path parameters of every job (Wendland C2, 64 neighbours, gamma 5/3)
KERNEL = 6
DES_NUM_NGB = 64.0
GAMMA = 5.0 / 3.0


def job_params():
    return dict(box=1.0, kernel=KERNEL, gamma=GAMMA, des_num_ngb=DES_NUM_NGB, max_num_ngb_dev=0.1,
                min_kernel_radius=0.0, max_kernel_radius=0.25, time=0.0, time_begin=0.0)


# timestep-loop parameters of every job (the end of the timeline lies beyond any job's steps)
JOB_INTEG = dict(courant_fac=0.2, err_tol_int_accuracy=0.02, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=100.0, min_egy_spec=0.0)


class Job:
    __slots__ = ("jid", "seed", "n_side", "mach", "steps", "arrival", "admitted", "done", "box", "result",
                 "updates", "fields")

    def __init__(self, jid, seed, n_side, mach, steps, arrival=0.0):
        self.jid, self.seed, self.n_side, self.mach, self.steps, self.arrival = jid, seed, n_side, mach, steps, arrival
        self.admitted = self.done = None
        self.box = None
        self.result = None
        self.updates = 0
        self.fields = None  # the request's particles, when made before it arrives

    @property
    def n(self):
        return self.n_side ** 3

    def ic(self):
        """The job's particles: a jittered lattice (n_side^3 particles, the cube nearest the
        drawn N), velocities Gaussian with rms Mach `mach` (per component mach/sqrt(3) c_s)."""
        ic = ics.lattice_box(self.n_side, self.seed, gamma=GAMMA, mach_sigma=self.mach / math.sqrt(3.0),
                             kernel=KERNEL, des_num_ngb=DES_NUM_NGB)
        return dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"])


def draw_jobs(count, seed, n_min=8192, n_max=1048576, mach=(0.1, 5.0), steps=(10, 200)):
    """Seeded job parameters: N log-uniform in [n_min, n_max] (rounded to the nearest lattice
    cube within the range), rms Mach uniform, steps uniform over the integers in the range."""
    rng = np.random.Generator(np.random.Philox(int(seed)))
    lo, hi = int(math.ceil(n_min ** (1 / 3) - 1e-9)), int(math.floor(n_max ** (1 / 3) + 1e-9))
    jobs = []
    for k in range(count):
        n = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        side = min(max(int(round(n ** (1 / 3))), lo), hi)
        m = rng.uniform(*mach)
        s = int(rng.integers(steps[0], steps[1] + 1))
        jobs.append(Job(k, int(seed) * 100003 + k, side, float(m), s))
    return jobs


def poisson_arrivals(jobs, rate, seed):
    """Arrival times of a Poisson stream of the given rate (jobs per second), in job order."""
    rng = np.random.Generator(np.random.Philox(int(seed) + 7))
    t = 0.0
    for j in jobs:
        t += rng.exponential(1.0 / rate)
        j.arrival = t


class Server:
    """Continuous batching of jobs on one device within a particle budget."""

    def __init__(self, capacity, max_boxes, precision=GZ_FP64, keep=None, device=0):
        self.batch = Batch(job_params(), capacity, max_boxes, precision, device)
        self.capacity = capacity
        self.max_boxes = max_boxes
        self.keep = keep  # field names stored as each job's result (None: none)

    def close(self):
        self.batch.close()

    def run(self, jobs, clock=time.perf_counter, sleep=time.sleep, log=None):
        """Serve the jobs (sorted by arrival; arrival times in seconds from the start).  Returns
        the wall time of the run; each job gets admitted/done times and its active updates."""
        b = self.batch
        waiting = sorted(jobs, key=lambda j: j.arrival)
        for j in waiting:
            if j.n > self.capacity:
                raise ValueError(f"job {j.jid}: {j.n} particles exceed the capacity {self.capacity}")
        resident = {}
        resident_n = 0
        t0 = clock()
        qi = 0
        while qi < len(waiting) or resident:
            now = clock() - t0
            admitted = []
            while qi < len(waiting) and waiting[qi].arrival <= now:
                j = waiting[qi]
                if resident_n + j.n > self.capacity or len(resident) >= self.max_boxes:
                    break
                j.box = b.add(j.fields if j.fields is not None else j.ic(), job_params(), JOB_INTEG)
                j.admitted = clock() - t0
                resident[j.box] = j
                resident_n += j.n
                admitted.append(j)
                qi += 1
            if admitted:
                b.begin()
                for j in admitted:
                    j.updates += j.n
            if not resident:
                # idle until the next arrival
                wait = waiting[qi].arrival - (clock() - t0)
                if wait > 0:
                    sleep(wait)
                continue
            b.step()
            finished = []
            for bid, j in resident.items():
                tl = b.timeline(bid)
                j.updates += int(tl["n_active"])
                if tl["num_steps"] >= j.steps or tl["stop_code"]:
                    finished.append(bid)
            for bid in finished:
                j = resident.pop(bid)
                if self.keep:
                    j.result = b.store(bid, self.keep)
                b.remove(bid)
                resident_n -= j.n
                j.done = clock() - t0
                if log:
                    log(f"job {j.jid} done: n={j.n} steps={j.steps} latency={j.done - j.arrival:.3f}s")
        return clock() - t0


def latency_summary(jobs, wall):
    lat = np.array([j.done - j.arrival for j in jobs])
    updates = sum(j.updates for j in jobs)
    return dict(jobs=len(jobs), p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                mean_s=float(lat.mean()), max_s=float(lat.max()), wall_s=wall, updates=int(updates),
                updates_per_s=updates / wall if wall > 0 else 0.0)
