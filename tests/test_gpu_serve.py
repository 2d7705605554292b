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
