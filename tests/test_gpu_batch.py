"""Batches of independent boxes (include/gizmo_batch.h) through the C ABI.

Bars, FP64 plain form:
  * a batch holding one box reproduces the single-box run loop (gz_run_begin / gz_run_step,
    itself held to the oracle's restatement of run() in test_gpu_timeloop.py) bit for bit: the
    integer timeline at every sync point and every field at the end;
  * a box computes the same bits alone or in a mix of boxes of other sizes, geometries and
    timestep parameters, also when boxes leave the batch and others join it between steps
    (continuous batching);
  * one box of a mix against the oracle directly, at the 1e-9 bar of the single-box loop.
"""
import numpy as np
import pytest

from gizmo_amd import ics
from gizmo_amd.batch import Batch
from gizmo_amd.hydro import GZ_FP64, Hydro

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "vel_kick", "u", "u_kick", "h", "rho", "press", "numngb", "divvel", "accel", "dtu", "maxsig",
          "grad_rho", "grad_p", "grad_v", "nvt", "cn", "fce", "soft", "timebin", "ti_begstep", "dt_step", "wakeup")
TIMELINE = ("ti_current", "n_active", "n_woken", "num_steps", "bin_active", "stop_code")

INTEG_A = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
               time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
INTEG_B = dict(courant_fac=0.1, err_tol_int_accuracy=0.01, max_size_timestep=0.02, min_size_timestep=0.0,
               time_begin=0.5, time_max=2.5, min_egy_spec=0.0)


def box_ic(kind):
    if kind == "sedov20":
        return ics.sedov_box(20, 1002, kernel=6, des_num_ngb=64.0), INTEG_A
    if kind == "sedov12_b2":
        return ics.sedov_box(12, 2001, kernel=6, des_num_ngb=64.0, box=2.0, energy=3.0), INTEG_B
    if kind == "lattice14":
        return ics.lattice_box(14, 2002, kernel=6, des_num_ngb=64.0, mach_sigma=0.3), INTEG_A
    if kind == "sedov16":
        return ics.sedov_box(16, 2003, kernel=6, des_num_ngb=64.0), INTEG_B
    raise ValueError(kind)


def particles(ic):
    return dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"])


def same_timeline(a, b, where):
    for k in TIMELINE:
        assert a[k] == b[k], (where, k, a[k], b[k])
    assert np.array_equal(a["bin_count"], b["bin_count"]), where


def same_fields(a, b, where):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), (where, k, float(np.abs(np.asarray(a[k], float) - np.asarray(b[k], float)).max()))


def test_one_box_equals_run_loop():
    ic, integ = box_ic("sedov20")
    par = ic["params"]
    n = len(ic["mass"])
    print(f"one box: n={n} steps=40", flush=True)
    g = Hydro(par, n, GZ_FP64)
    g.load(particles(ic))
    g.run_begin(integ)
    b = Batch(par, n, 4, GZ_FP64)
    bid = b.add(particles(ic), par, integ)
    b.begin()
    for k in range(40):
        g.run_step()
        b.step()
        same_timeline(b.timeline(bid), g.timeline(), k)
    same_fields(b.store(bid, FIELDS), g.store(FIELDS), "end")
    assert b.stat("active_updates") == g.stat("active_updates")
    g.close()
    b.close()


def run_alone(kind, steps):
    ic, integ = box_ic(kind)
    par = ic["params"]
    n = len(ic["mass"])
    b = Batch(par, n, 1, GZ_FP64)
    bid = b.add(particles(ic), par, integ)
    b.begin()
    tls = []
    for _ in range(steps):
        b.step()
        tls.append(b.timeline(bid))
    out = b.store(bid, FIELDS)
    b.close()
    return tls, out


# few_rows None: the default split (passes over more than 8192 rows walk row groups, one
# wavefront per cell of each box's grid); 0: every pass in row groups.  The boxes alone take one
# wavefront per row (each holds at most 8000 particles).
@pytest.mark.parametrize("few_rows", [None, 0])
def test_mix_equals_alone_with_joins_and_leaves(few_rows):
    # sedov20 and lattice14 from the start; sedov12_b2 joins after 5 steps; lattice14 leaves
    # after 12; sedov16 joins after 15 into the room it left
    steps = {"sedov20": 30, "lattice14": 12, "sedov12_b2": 25, "sedov16": 15}
    alone = {k: run_alone(k, s) for k, s in steps.items()}
    par = box_ic("sedov20")[0]["params"]
    size = {k: len(box_ic(k)[0]["mass"]) for k in steps}
    cap = max(size["sedov20"] + size["lattice14"] + size["sedov12_b2"], size["sedov20"] + size["sedov12_b2"] + size["sedov16"])
    b = Batch(par, cap, 3, GZ_FP64)
    if few_rows is not None:
        b.set_option("few_rows", few_rows)
    ids, done, tls = {}, {}, {k: [] for k in steps}

    def join(kind):
        ic, integ = box_ic(kind)
        ids[kind] = b.add(particles(ic), ic["params"], integ)

    join("sedov20")
    join("lattice14")
    b.begin()
    for t in range(30):
        if t == 5:
            join("sedov12_b2")
            b.begin()
        if t == 15:
            join("sedov16")
            b.begin()
        print(f"mix step {t}: few_rows={few_rows} boxes={b.boxes()} particles={b.particles()}", flush=True)
        b.step()
        for kind, bid in list(ids.items()):
            tls[kind].append(b.timeline(bid))
            if len(tls[kind]) == steps[kind]:
                done[kind] = b.store(bid, FIELDS)
                if kind == "lattice14":
                    b.remove(bid)
                del ids[kind]
    b.close()
    assert set(done) == set(steps)
    for kind in steps:
        for k, (x, y) in enumerate(zip(tls[kind], alone[kind][0])):
            same_timeline(x, y, (kind, k))
        same_fields(done[kind], alone[kind][1], kind)


def test_box_in_mix_against_oracle(oracle_lib):
    O = oracle_lib
    ic, integ = box_ic("sedov12_b2")
    par = ic["params"]
    other, integ_o = box_ic("lattice14")
    st = O.State.from_ic(ic)
    r = O.Run(par, integ, st)
    r.begin()
    n = len(ic["mass"]) + len(other["mass"])
    b = Batch(par, n, 2, GZ_FP64)
    b.add(particles(other), other["params"], integ_o)
    bid = b.add(particles(ic), par, integ)
    b.begin()
    print(f"box vs oracle: n={st.n} steps=20", flush=True)
    for k in range(20):
        na_o = r.step()
        b.step()
        tl = b.timeline(bid)
        assert tl["ti_current"] == r.ti_current, k
        assert tl["n_active"] == na_o, k
        assert np.array_equal(tl["bin_count"], np.array(r.tl.bin_count[:], dtype=np.int64)), k
    out = b.store(bid, ["pos", "timebin", "ti_begstep", "dt_step", "u", "rho", "h", "accel", "vel"])
    b.close()
    assert np.array_equal(out["timebin"], st.timebin)
    assert np.array_equal(out["ti_begstep"], st.ti_begstep)
    assert np.array_equal(out["dt_step"], st.dt_step)
    d = out["pos"] - st.pos
    box = par["box"]
    assert np.abs(d - box * np.round(d / box)).max() <= 1e-9 * box
    import parity
    sc = parity.scales(st, par["gamma"])
    for k in ("u", "rho", "h", "accel", "vel"):
        err = parity.err(k, out[k], getattr(st, k), sc)
        assert err <= 1e-9, (k, err)


PASS_FIELDS = ("h", "numngb", "drkern", "divvel", "rho", "press", "nvt", "nvt_fw", "cn", "fce", "grad_rho", "grad_p",
               "grad_v", "accel", "dtu", "maxsig")


def test_box_in_mix_fp32_against_oracle(oracle_lib):
    """The FP32-mixed batch (the ensemble's form): a box of a mix against the FP64 oracle after
    the batch's first pass (every particle active), at the oracle's converged radii with a band
    so wide that the first evaluation stands (identical inputs and neighbour sets): every
    quantity of every particle within 1e-5 per particle (tests/parity.py), du/dt within 2e-5
    (the blast's cold particles next to its hot centre; see test_gpu_parity.py)."""
    import parity
    O = oracle_lib
    ic, integ = box_ic("sedov12_b2")
    other, integ_o = box_ic("lattice14")
    par = dict(ic["params"])
    s1 = O.State.from_ic(ic)
    O.density(par, s1)
    pfix = dict(par, max_num_ngb_dev=1e30)
    ref = s1.copy()
    O.density(pfix, ref)
    O.gradients(pfix, ref)
    O.force(pfix, ref)
    n = s1.n + len(other["mass"])
    from gizmo_amd.hydro import GZ_FP32_MIXED
    b = Batch(pfix, n, 2, GZ_FP32_MIXED)
    b.add(particles(other), other["params"], integ_o)
    bid = b.add(dict(pos=s1.pos, vel=s1.vel, mass=s1.mass, u=s1.u, h=s1.h), par, integ)
    print(f"fp32 box vs oracle: n={s1.n} in a mix of {n}", flush=True)
    b.begin()
    out = b.store(bid, PASS_FIELDS)
    b.close()
    pe = parity.errs(PASS_FIELDS, out, ref, parity.scales(ref, par["gamma"]))
    print("per-particle:", " ".join(f"{k}={v:.1e}" for k, v in pe.items()), flush=True)
    for k, e in pe.items():
        assert e <= (2e-5 if k == "dtu" else 1e-5), (k, e)


def test_batch_rejects_what_does_not_fit():
    ic, integ = box_ic("lattice14")
    par = ic["params"]
    n = len(ic["mass"])
    b = Batch(par, n + 10, 2, GZ_FP64)
    b.add(particles(ic), par, integ)
    with pytest.raises(Exception, match="capacity"):
        b.add(particles(ic), par, integ)
    bad = dict(integ, time_max=integ["time_begin"])
    small = ics.lattice_box(2, 1, kernel=6, des_num_ngb=64.0)
    with pytest.raises(Exception, match="time_max"):
        b.add(particles(small), par, bad)
    b.close()
