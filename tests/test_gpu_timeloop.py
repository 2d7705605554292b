"""The device timestep loop (gz_run_begin / gz_run_step through the C ABI) against the oracle's
restatement of run() (oracle/gz_oracle_time.c).

Bars, FP64 plain form:
  * the integer timeline is identical at every sync point: All.Ti_Current, the number of
    active particles, the occupation of every time bin, and at the end every particle's
    TimeBin, Ti_begstep and dt_step;
  * every floating-point field of every particle within 1e-9 of the oracle, relative to the
    particle's own value with a floor at its physical scale of the field (tests/parity.py),
    after the whole run (the device sums in its Morton order and the oracle in the caller's, so
    rounding differences enter at 1e-16 and grow through the steps);
  * positions compared as periodic differences.
Workloads: global steps on a 32^3 cubic-spline lattice (every particle in one bin, the forced
single-bin recipe of the reference's own tests), adaptive bins on a 32^3 Sedov blast whose
hot centre runs up to 2^8 times finer steps than the cold medium and wakes its neighbours, and
the first 10 sync points of the benchmark's 128^3 (2M-particle) Sedov job.
"""
import numpy as np
import pytest

import parity
from gizmo_amd import ics
from gizmo_amd.hydro import GZ_FP64, Hydro

pytestmark = pytest.mark.gpu

TOL = 1e-9
FLOAT_FIELDS = ("vel", "vel_kick", "u", "u_kick", "h", "rho", "press", "numngb", "divvel", "accel", "dtu", "maxsig",
                "grad_rho", "grad_p", "nvt", "cn")


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return float((np.abs(a - b) / scale).max())


def periodic_diff(a, b, box):
    d = np.asarray(a) - np.asarray(b)
    return np.abs(d - box * np.round(d / box)).max()


def run_both(oracle_lib, ic, integ, steps, resort_fraction=None, few_rows=None):
    O = oracle_lib
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    r = O.Run(par, integ, st)
    r.begin()
    g = Hydro(par, st.n, GZ_FP64)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    if resort_fraction is not None:
        g.set_option("resort_fraction", resort_fraction)
    if few_rows is not None:
        g.set_option("few_rows", few_rows)
    g.run_begin(integ)
    print(f"run_both: n={st.n} steps={steps} resort_fraction={resort_fraction} few_rows={few_rows} "
          f"domain_build_active_fraction={integ.get('domain_build_active_fraction', 0.0)}", flush=True)
    hist = []
    for k in range(steps):
        na_o = r.step()
        na_g = g.run_step()
        tl = g.timeline()
        assert tl["ti_current"] == r.ti_current, (k, tl["ti_current"], r.ti_current)
        assert na_g == na_o, (k, na_g, na_o)
        assert np.array_equal(tl["bin_count"], np.array(r.tl.bin_count[:], dtype=np.int64)), k
        assert tl["n_woken"] == r.tl.n_woken, (k, tl["n_woken"], r.tl.n_woken)
        # the device drifts every particle the oracle drifts (its cells around the active rows
        # hold every partner the oracle's pair test accepts), and at least at the same steps
        tic = g.store(["ti_current"])["ti_current"]
        assert (tic >= st.ti_current).all(), k
        if r.tl.full_drift:
            assert tl["full_drift"] == 1 and (tic == r.ti_current).all(), k
        hist.append((na_o, int(r.tl.n_woken), int(r.tl.full_drift), int((tic != r.ti_current).sum())))
    # every particle to the last sync point on both sides (gizmo_full_drift_to), then compare
    r.drift_all()
    g.drift_all()
    out = g.store(["pos", "timebin", "ti_begstep", "dt_step", "wakeup", "ti_current"] + list(FLOAT_FIELDS))
    assert (out["ti_current"] == r.ti_current).all() and (st.ti_current == r.ti_current).all()
    out["resorts"] = g.stat("resorts")
    g.close()
    return st, out, hist, par


def check_fields(st, out, par):
    assert np.array_equal(out["timebin"], st.timebin)
    assert np.array_equal(out["ti_begstep"], st.ti_begstep)
    assert np.array_equal(out["dt_step"], st.dt_step)
    assert periodic_diff(out["pos"], st.pos, par["box"]) <= TOL * par["box"]
    print("field-max:", {k: f"{relerr(out[k], getattr(st, k)):.1e}" for k in FLOAT_FIELDS})
    errs = parity.errs(FLOAT_FIELDS, out, st, parity.scales(st, par["gamma"]))
    print("per-particle:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k, e in errs.items():
        assert e <= TOL, (k, e)


def test_global_steps_lattice_c1(oracle_lib):
    """The CPU-runnable configuration at its size: 32^3 jittered lattice, cubic spline, 32
    neighbours, 10 global steps (one forced bin: the maximum step below every Courant step)."""
    ic = ics.lattice_box(32, 1001, kernel=3, des_num_ngb=32.0)
    integ = dict(courant_fac=0.2, err_tol_int_accuracy=0.02, max_size_timestep=0.002, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
    st, out, hist, par = run_both(oracle_lib, ic, integ, 10)
    assert all(h[0] == st.n for h in hist)
    check_fields(st, out, par)


@pytest.mark.parametrize("resort_fraction,few_rows,frac", [(None, None, 0.0), (0.01, None, 0.0), (None, None, 0.5),
                                                           (0.01, None, 0.5), (None, 0, 0.5), (0.01, 0, 0.5)])
def test_adaptive_bins_sedov(oracle_lib, resort_fraction, few_rows, frac):
    """Adaptive bins with the eager drift (DomainBuild_ActiveFraction 0: every particle at every
    sync point) and the reference's lazy drift (0.5: the blast's steps drift their active bins
    and the particles their passes read), through the one-wavefront-per-row kernels (few active
    rows, the default) and the one-lane-per-row kernels (few_rows 0), with and without the
    particles being ordered again.  32^3 particles: the search grid's cells then span an eighth of
    the box, so the cells around the blast's rows hold 12-36% of the particles (at 20^3 their
    27-cell stencils cover the whole box and nothing stays behind)."""
    ic = ics.sedov_box(32, 1002, kernel=6, des_num_ngb=64.0)
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0, domain_build_active_fraction=frac)
    st, out, hist, par = run_both(oracle_lib, ic, integ, 40, resort_fraction, few_rows)
    if resort_fraction is not None:
        assert out["resorts"] > 0, "the particles were never ordered again"
    del out["resorts"]
    active = [h[0] for h in hist]
    assert min(active) < st.n // 10, "bins never split"
    assert sum(h[1] for h in hist) > 0, "no wake-up happened"
    if frac > 0:
        lazy = [h for h in hist if not h[2]]
        assert len(lazy) >= 20 and max(h[3] for h in lazy) > st.n // 2, "the cold medium was drifted every step"
    check_fields(st, out, par)


def test_sedov128_job_against_oracle(oracle_lib):
    """The 2M-particle job of the benchmark at its size (128^3 Sedov blast, Wendland C2, 64
    neighbours, CourantFac 0.05, ErrTolIntAccuracy 0.002, MaxSizeTimestep 0.01,
    DomainBuild_ActiveFraction 0.5): its first 10 sync points against the oracle's run(), the
    timeline identical at each and every field within 1e-9 per particle at the end."""
    ic = ics.sedov_box(128, 1002, kernel=6, des_num_ngb=64.0)
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0, domain_build_active_fraction=0.5)
    st, out, hist, par = run_both(oracle_lib, ic, integ, 10)
    del out["resorts"]
    print("sedov128 steps (active, woken, full drift, behind):", hist, flush=True)
    assert max(h[0] for h in hist) < st.n // 10, "the job's steps are partial steps"
    check_fields(st, out, par)


@pytest.mark.parametrize("few_rows,frac", [(None, 0.0), (None, 0.5), (0, 0.5)])
def test_run_steps_same_as_single_steps(few_rows, frac):
    """gz_run_steps(k) takes the same k sync points as k calls of gz_run_step: identical active
    and woken counts per step, timeline and state, bit for bit (the same launches in the same
    order)."""
    ic = ics.sedov_box(20, 1002, kernel=6, des_num_ngb=64.0)
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0, domain_build_active_fraction=frac)
    par = dict(ic["params"])
    outs, counts = [], []
    for multi in (False, True):
        print(f"run_steps: multi={multi}", flush=True)
        g = Hydro(par, len(ic["mass"]), GZ_FP64)
        if few_rows is not None:
            g.set_option("few_rows", few_rows)
        g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
        g.run_begin(integ)
        if multi:
            act, woken = g.run_steps(25)
        else:
            act, woken = [], []
            for _ in range(25):
                act.append(g.run_step())
                woken.append(g.timeline()["n_woken"])
        tl = g.timeline()
        counts.append((act, woken, tl["ti_current"], tl["num_steps"], tl["bin_count"].tolist()))
        outs.append(g.store(["pos", "timebin", "ti_begstep", "dt_step", "ti_current"] + list(FLOAT_FIELDS)))
        g.close()
    assert counts[0] == counts[1]
    assert sum(counts[0][1]) > 0, "no wake-up happened"
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
