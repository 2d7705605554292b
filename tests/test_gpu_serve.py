"""The serving loop (gizmo_amd/serve.py): jobs that arrive while others run are admitted as the
particle budget and the box slots free up, share every step with the resident jobs, and each
job's result equals the same job run alone, bit for bit, in both the FP64 and the FP32-mixed
form (a box's sums run over its own particles in its own order whatever else is resident).
"""
import numpy as np
import pytest

from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64
from gizmo_amd.serve import Server, draw_jobs, latency_summary

pytestmark = pytest.mark.gpu

KEEP = ["pos", "vel", "u", "h", "rho", "press", "accel", "dtu", "grad_v", "timebin", "dt_step"]


def small_jobs():
    jobs = draw_jobs(6, 4242, n_min=1728, n_max=4096, steps=(4, 16))
    # all but the first arrive while the first runs; the budget holds about two at once
    for k, j in enumerate(jobs):
        j.arrival = 0.0 if k == 0 else 0.01 * k
    return jobs


@pytest.mark.parametrize("precision", [GZ_FP64, GZ_FP32_MIXED], ids=["fp64", "fp32_mixed"])
def test_each_job_equals_alone(precision):
    jobs = small_jobs()
    cap = max(j.n for j in jobs) * 2
    print(f"serve mix: precision={precision} sizes={[j.n for j in jobs]} steps={[j.steps for j in jobs]}", flush=True)
    s = Server(cap, 3, precision, keep=KEEP)
    wall = s.run(jobs)
    s.close()
    summ = latency_summary(jobs, wall)
    assert summ["jobs"] == len(jobs)
    # the budget forced overlap and queueing: some job waited, some job shared its steps
    assert any(j.admitted > j.arrival + 1e-6 for j in jobs)
    assert max(j.admitted for j in jobs) < max(j.done for j in jobs)
    for j in jobs:
        alone = draw_jobs(6, 4242, n_min=1728, n_max=4096, steps=(4, 16))[j.jid]
        a = Server(alone.n, 1, precision, keep=KEEP)
        a.run([alone])
        a.close()
        assert alone.updates == j.updates, j.jid
        for k in KEEP:
            assert np.array_equal(j.result[k], alone.result[k]), (j.jid, k)


KEEP_ORACLE = ["pos", "vel", "vel_kick", "u", "u_kick", "h", "rho", "press", "accel", "dtu", "maxsig", "timebin",
               "ti_begstep", "dt_step"]


def test_jobs_of_the_serving_sizes(oracle_lib):
    """Jobs of N in [8192, 32768] (the lower end of the serving mix, lattice sides 21 to 32)
    through the server with a budget that holds about two: each job equals itself alone bit
    for bit, and the first job equals the oracle's run() of the same box (its integer timeline
    exactly, its fields within 1e-9 per particle, tests/parity.py)."""
    import parity
    from gizmo_amd.serve import JOB_INTEG, job_params
    O = oracle_lib

    def jobs_():
        js = draw_jobs(4, 5151, n_min=8192, n_max=32768, steps=(6, 12))
        for k, j in enumerate(js):
            j.arrival = 0.0 if k == 0 else 0.005 * k
        return js

    jobs = jobs_()
    cap = max(j.n for j in jobs) + sorted(j.n for j in jobs)[-2]
    print(f"serve sizes: {[j.n for j in jobs]} steps={[j.steps for j in jobs]} capacity={cap}", flush=True)
    s = Server(cap, 3, GZ_FP64, keep=KEEP_ORACLE)
    s.run(jobs)
    s.close()
    assert all(8192 <= j.n <= 32768 for j in jobs)
    for j in jobs:
        alone = jobs_()[j.jid]
        a = Server(alone.n, 1, GZ_FP64, keep=KEEP_ORACLE)
        a.run([alone])
        a.close()
        assert alone.updates == j.updates, j.jid
        for k in KEEP_ORACLE:
            assert np.array_equal(j.result[k], alone.result[k]), (j.jid, k)
    j = jobs[0]
    par = job_params()
    st = O.State.from_ic(j.ic())
    r = O.Run(par, JOB_INTEG, st)
    r.begin()
    updates = st.n
    for _ in range(j.steps):
        updates += r.step()
    print(f"job 0 vs oracle: n={j.n} steps={j.steps} updates={updates}", flush=True)
    assert updates == j.updates
    out = j.result
    assert np.array_equal(out["timebin"], st.timebin)
    assert np.array_equal(out["ti_begstep"], st.ti_begstep)
    assert np.array_equal(out["dt_step"], st.dt_step)
    assert parity.periodic_err(out["pos"], st.pos, par["box"]) <= 1e-9 * par["box"]
    fields = ("vel", "vel_kick", "u", "u_kick", "h", "rho", "press", "accel", "dtu", "maxsig")
    errs = parity.errs(fields, out, st, parity.scales(st, par["gamma"]))
    print("per-particle:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= 1e-9, (k, e)


def test_overlapping_jobs_share_steps():
    # two jobs resident together: a shared step advances both
    jobs = draw_jobs(2, 77, n_min=1728, n_max=2197, steps=(6, 6))
    for j in jobs:
        j.arrival = 0.0
    s = Server(sum(j.n for j in jobs), 2, GZ_FP64)
    wall = s.run(jobs)
    assert s.batch.stat("steps") == 6
    s.close()
    assert wall > 0 and all(j.done is not None for j in jobs)
