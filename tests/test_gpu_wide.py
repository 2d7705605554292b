"""The search over many grid levels (option "wide_search") on particles whose kernel radii span
orders of magnitude: a lognormal density field (sigma(ln rho) = 2, P(k) ~ k^-3).

Bars, FP64 plain form:
  * symmetric neighbour sets identical to the oracle's, row by row;
  * a full hot-path pass (density with its radius iteration, gradients, force) identical, bit
    for bit, to the pass with one grid sized by the largest radius (the walk visits the same
    candidates in the same ascending order, only fewer of them), and within 1e-12 of the oracle.
"""
import numpy as np
import pytest

from gizmo_amd import ics
from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64, Hydro

pytestmark = pytest.mark.gpu

FIELDS = ("h", "numngb", "rho", "press", "nvt", "cn", "fce", "drkern", "divvel", "grad_rho", "grad_p", "grad_v",
          "accel", "dtu", "maxsig")


@pytest.fixture(scope="module")
def wide_case(oracle_lib):
    ic = ics.lognormal_box(24000, 3003, ng=32)
    par = dict(ic["params"])
    O = oracle_lib
    s0 = O.State.from_ic(ic)
    s1 = s0.copy()
    O.density(par, s1)
    s3 = s1.copy()
    O.gradients(par, s3)
    O.force(par, s3)
    return dict(ic=ic, par=par, s1=s1, s3=s3)


def device(ic, par, wide, precision=GZ_FP64, **options):
    g = Hydro(par, len(ic["mass"]), precision)
    g.set_option("wide_search", wide)
    for k, v in options.items():
        g.set_option(k, v)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    return g


def test_radii_span_levels(wide_case):
    h = wide_case["s1"].h
    assert h.max() / h.min() > 10.0


# few_rows 0: every row in the cell walk of row groups (one wavefront per cell of each row's own
# level); 10**9: one wavefront per row
@pytest.mark.parametrize("few_rows", [0, 10**9])
def test_symmetric_sets_exact(wide_case, oracle_lib, few_rows):
    s1, par = wide_case["s1"], wide_case["par"]
    g = Hydro(par, s1.n, GZ_FP64)
    g.set_option("wide_search", 1)
    g.set_option("few_rows", few_rows)
    g.load(dict(pos=s1.pos, vel=s1.vel, mass=s1.mass, u=s1.u, h=s1.h))
    print(f"wide sets: n={s1.n} few_rows={few_rows} h in [{s1.h.min():.2e}, {s1.h.max():.2e}]", flush=True)
    off_g, nb_g = g.neighbors(1)
    g.close()
    off_o, nb_o = oracle_lib.neighbors(par, s1.pos, s1.h, 1)
    assert np.array_equal(off_g, off_o)
    for i in range(s1.n):
        row = nb_g[off_g[i]:off_g[i + 1]]
        assert np.array_equal(np.sort(row), nb_o[off_o[i]:off_o[i + 1]]), i


def test_full_pass_same_bits_as_one_grid(wide_case):
    ic, par = wide_case["ic"], wide_case["par"]
    outs = {}
    # one grid; many levels with the default split; every list in row groups; every list one
    # wavefront per row; the force pass's lanes in plain row order
    cases = {"one grid": (0, {}), "wide": (1, {}), "wide groups": (1, dict(few_rows=0)), "wide rows": (1, dict(few_rows=10**9)),
             "wide rows in order": (1, dict(row_schedule=0)), "one grid rows in order": (0, dict(row_schedule=0)),
             "wide, force at two waves": (1, dict(force_waves=2)), "one grid, force at three waves": (0, dict(force_waves=3))}
    for name, (wide, opts) in cases.items():
        print(f"full pass: {name}", flush=True)
        g = device(ic, par, wide, **opts)
        g.hydro_pass()
        outs[name] = g.store(FIELDS)
        g.close()
    for name in cases:
        for k in FIELDS:
            assert np.array_equal(outs["one grid"][k], outs[name][k]), (name, k)


def test_full_pass_against_oracle(wide_case):
    """Also the path of a symmetric list that overflows its first width: the lognormal field's
    longest lists exceed the initial width, the gradient and force passes run on the cut lists
    (their lengths stored clamped to the width), the overflow read back with the force pass
    calls for a wider list and both passes run again."""
    ic, par, s3 = wide_case["ic"], wide_case["par"], wide_case["s3"]
    g = device(ic, par, 1)
    g.hydro_pass()
    widened = g.stat("ell_widen")
    out = g.store(FIELDS)
    g.close()
    assert widened >= 1
    for k in FIELDS:
        a = np.asarray(out[k], dtype=np.float64).reshape(s3.n, -1)
        b = np.asarray(getattr(s3, k), dtype=np.float64).reshape(s3.n, -1)
        err = float((np.abs(a - b) / np.maximum(np.abs(b).max(axis=0), 1e-300)).max())
        assert err <= 1e-12, (k, err)


@pytest.mark.parametrize("few_rows", [0, 10**9])
def test_fp32_mixed_wide_against_oracle(wide_case, oracle_lib, few_rows):
    """The FP32-mixed form on the lognormal field (gamma = 1.001, radii spanning 130x) against
    the FP64 oracle at the oracle's converged radii, with the band so wide that the first
    evaluation stands for every particle that has its neighbours: 99% of the particles within
    1e-5 per particle in every quantity (measured on the MI355X: p99 <= 1.7e-6).  The largest
    errors sit at the few particles whose radius the iteration still moves (a handful of
    neighbours, where the radius search never settles): there the FP32 sums move the radius
    itself (up to 2e-4), and with it the neighbour sets."""
    import parity
    O = oracle_lib
    par = dict(wide_case["par"], max_num_ngb_dev=1e30)
    s1 = wide_case["s1"]
    ref = s1.copy()
    O.density(par, ref)
    O.gradients(par, ref)
    O.force(par, ref)
    g = Hydro(par, s1.n, GZ_FP32_MIXED)
    g.set_option("wide_search", 1)
    g.set_option("few_rows", few_rows)
    g.load(dict(pos=s1.pos, vel=s1.vel, mass=s1.mass, u=s1.u, h=s1.h))
    print(f"wide fp32: n={s1.n} few_rows={few_rows}", flush=True)
    g.hydro_pass()
    out = g.store(FIELDS)
    g.close()
    sc = parity.scales(ref, par["gamma"])
    for k in FIELDS:
        a = np.asarray(out[k], dtype=np.float64).reshape(s1.n, -1)
        b = np.asarray(getattr(ref, k), dtype=np.float64).reshape(s1.n, -1)
        d = np.sqrt(((a - b) ** 2).sum(1))
        den = np.maximum(np.sqrt((b ** 2).sum(1)), sc.get(k, 0.0))
        e = d / np.where(den > 0, den, 1e-300)
        print(f"  {k}: p99 {np.quantile(e, 0.99):.2e} max {e.max():.2e}", flush=True)
        assert np.isfinite(a).all(), k
        assert np.quantile(e, 0.99) <= 1e-5, (k, np.quantile(e, 0.99))


def test_fp32_mixed_loop_near_isothermal(wide_case):
    """The FP32-mixed form's Riemann ladder at gamma = 1.001 (the lognormal field, Mach 2) over 30
    sync points of the timestep loop with the search over many levels: no face fails the ladder
    (a failed face would stop the loop with code 1234) and every field stays finite, as in the
    FP64 run of the same loop (the rungs each form took are printed)."""
    ic, par = wide_case["ic"], wide_case["par"]
    integ = dict(courant_fac=0.1, err_tol_int_accuracy=0.01, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
    counts = {}
    for name, precision in (("fp64", GZ_FP64), ("fp32_mixed", GZ_FP32_MIXED)):
        g = device(ic, par, 1, precision)
        g.set_option("riemann_stats", 1)
        g.run_begin(integ)
        print(f"near-isothermal loop {name}: n={len(ic['mass'])} gamma={par['gamma']}", flush=True)
        acts = [g.run_step() for _ in range(30)]
        tl = g.timeline()
        out = g.store(["vel", "u", "rho", "h", "accel", "dtu"])
        rungs = g.rung_counts()
        g.close()
        assert tl["stop_code"] == 0
        for k, v in out.items():
            assert np.isfinite(v).all(), (name, k)
        counts[name] = (sum(acts), tl["ti_current"], {k: v for k, v in rungs.items() if v})
        print(f"near-isothermal loop {name}: {sum(acts)} active updates, ti {tl['ti_current']}, rungs {counts[name][2]}",
              flush=True)
    r32 = counts["fp32_mixed"][2]
    assert r32.get("riemann_faces", 0) > 0 and r32.get("riemann_fail", 0) == 0, r32
