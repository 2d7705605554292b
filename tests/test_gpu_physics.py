"""End-to-end physics bars of the reference's own tests, re-created with this project's ICs in
3D and run through the device's timestep loop (gz_run_*) in both arithmetic forms:

  * Sedov-Taylor blast (test/sedov/test_sedov.py:103): the radially binned median density against
    the self-similar solution, L1 < 0.3 (tests/analytic.py supplies the solution, pinned in
    tests/test_analytic.py);
  * Sod shock tube (test/shocktube/test_shocktube.py:71-84): L1 of density < 0.05, of velocity and
    internal energy < 0.1 against the exact Riemann solution, and conservation from the first
    sync point to the last: energy < 1e-5 (relative), momentum < 1e-10 (in units of
    sqrt(2 M E)), mass < 1e-12.  The reference runs the tube in 1D; in 3D the bars need the
    waves resolved by ~128 particles per box side on the dense side (measured on the MI355X,
    FP64 at t = 0.13: 48 per side gives L1 rho/v/u 0.078/0.28/0.041, 96 gives 0.047/0.146/0.026,
    128 with the Wendland C2 kernel and 64 neighbours 0.031/0.096/0.018);
  * sound wave (test/soundwave/test_soundwave.py:22): after one period the density, velocities
    and internal energy of every particle equal their initial values within rtol 1e-5,
    atol 1e-8 (FP64; the FP32-mixed form's pair arithmetic cannot resolve a 1e-6 wave against
    pressure terms of order 1, so it is held at 1e-3 amplitude to the same relative bar).  The
    velocity bar (1% of the wave's amplitude) is met from 64^3 particles on: the returned wave
    lags by a phase that falls as the square of the resolution (measured: 0.034, 0.016 and
    0.0088 rad at 32^3, 48^3 and 64^3 with the cubic spline and 32 neighbours, amplitude ratio
    0.997, 0.9991, 0.9996).
The blast and the tube exercise the Riemann ladder's fallbacks and the entropic switch (counted
on the device for the blast).
"""
import numpy as np
import pytest

import analytic as A
from gizmo_amd import ics
from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64, Hydro

pytestmark = pytest.mark.gpu

FORMS = {"fp64": GZ_FP64, "fp32_mixed": GZ_FP32_MIXED}
HALF_TIMELINE = 1 << 59  # time_max = 2 t_end: t_end is the sync point 2^59 of every bin


def run_until_half(g, max_steps=20000):
    """Steps the loop to the middle of the timeline (t_end; no sync point skips it, since every
    bin's step divides 2^59), returning the active counts."""
    acts = []
    while g.timeline()["ti_current"] < HALF_TIMELINE:
        acts.append(g.run_step())
        assert len(acts) < max_steps, "too many steps"
    assert g.timeline()["ti_current"] == HALF_TIMELINE
    return acts


def make_ctx(ic, precision):
    g = Hydro(dict(ic["params"]), len(ic["mass"]), precision)
    g.load(dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]))
    return g


@pytest.mark.parametrize("form", list(FORMS))
def test_sedov_blast_l1(form):
    n_side, E, rho0 = 32, 1.0, 1.0
    ic = ics.sedov_box(n_side, 1002, kernel=6, des_num_ngb=64.0, energy=E)
    st = A.SedovTaylor(5.0 / 3.0)
    t_end = (0.3 / st.xi0) ** 2.5  # shock radius 0.3 of the unit box
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=2 * t_end, min_egy_spec=0.0, domain_build_active_fraction=0.5)
    print(f"sedov {form}: n={n_side}^3 t_end={t_end:.4f}", flush=True)
    g = make_ctx(ic, FORMS[form])
    g.set_option("riemann_stats", 1)
    g.run_begin(integ)
    acts = run_until_half(g)
    tl = g.timeline()
    g.drift_all()
    out = g.store(["pos", "vel", "rho"])
    rungs = g.rung_counts()
    g.close()
    assert abs(tl["time"] - t_end) < 1e-12 * t_end
    d = out["pos"] - 0.5
    r = np.sqrt((d * d).sum(1))
    vr = (out["vel"] * d).sum(1) / np.maximum(r, 1e-30)
    R = st.shock_radius(E, rho0, t_end)
    edges = np.linspace(0, 1.5 * R, 31)
    cen = 0.5 * (edges[1:] + edges[:-1])
    idx = np.digitize(r, edges) - 1
    med = np.array([np.median(out["rho"][idx == k]) if (idx == k).any() else np.nan for k in range(30)])
    vmed = np.array([np.median(vr[idx == k]) if (idx == k).any() else np.nan for k in range(30)])
    ex_rho, ex_v = st.profile(cen, E, rho0, t_end)
    good = ex_rho > 0.01
    l1 = np.nanmean(np.abs(med[good] - ex_rho[good])) / np.nanmean(np.abs(ex_rho[good]))
    l1v = np.nanmean(np.abs(vmed - ex_v)) / np.nanmean(np.abs(ex_v))
    print(f"sedov {form}: {len(acts)} steps, {sum(acts)} active updates; density L1 {l1:.3f}, radial velocity L1 {l1v:.3f}",
          flush=True)
    print(f"sedov {form}: rungs", {k: v for k, v in rungs.items() if v}, flush=True)
    assert l1 < 0.3
    assert rungs["riemann_faces"] > 0 and rungs["entropic_applied"] > 0


def _sod_exact(x, t, g=1.4):
    """The two Sod problems of the slab at x = 3/4 (dense on the left) and its mirror at 1/4."""
    L, R = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    right = x >= 0.5
    rho = np.empty_like(x)
    u = np.empty_like(x)
    p = np.empty_like(x)
    r1, u1, p1 = A.riemann_sample(g, L, R, (x[right] - 0.75) / t)
    r2, u2, p2 = A.riemann_sample(g, L, R, (0.25 - x[~right]) / t)
    rho[right], u[right], p[right] = r1, u1, p1
    rho[~right], u[~right], p[~right] = r2, -u2, p2
    return rho, u, p


def _totals(out):
    m = out["mass"]
    e = (m * (out["u_kick"] + 0.5 * (out["vel_kick"] ** 2).sum(1))).sum()
    return m.sum(), (m[:, None] * out["vel_kick"]).sum(0), e


@pytest.mark.parametrize("form", list(FORMS))
def test_shocktube_sod(form):
    ic = ics.shocktube_box(128, 606, kernel=6, des_num_ngb=64.0)
    # before the shocks of the two faces meet across the light slab (t = 0.143)
    t_end = 0.13
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.0025, max_size_timestep=2e-4, min_size_timestep=0.0,
                 time_begin=0.0, time_max=2 * t_end, min_egy_spec=0.0)
    print(f"shocktube {form}: n={len(ic['mass'])} t_end={t_end}", flush=True)
    g = make_ctx(ic, FORMS[form])
    g.run_begin(integ)
    keys = ["mass", "vel_kick", "u_kick"]
    m0, p0, e0 = _totals(g.store(keys))
    acts = run_until_half(g)
    g.drift_all()
    out = g.store(["pos", "vel", "rho", "u"] + keys)
    g.close()
    m1, p1, e1 = _totals(out)
    x = out["pos"][:, 0]
    rho_x, u_x, p_x = _sod_exact(x, t_end)
    u_ex = p_x / (rho_x * 0.4)
    l1_rho = np.mean(np.abs(out["rho"] - rho_x)) / np.mean(np.abs(rho_x))
    l1_v = np.mean(np.abs(out["vel"][:, 0] - u_x)) / (np.mean(np.abs(u_x)) + 1e-10)
    l1_u = np.mean(np.abs(out["u"] - u_ex)) / np.mean(np.abs(u_ex))
    e_err = abs(e1 - e0) / e0
    p_err = np.abs(p1 - p0).max() / np.sqrt(2 * m0 * e0)
    m_err = abs(m1 - m0) / m0
    glob = all(a == len(ic["mass"]) for a in acts)
    print(f"shocktube {form}: {len(acts)} steps (global: {glob}); L1 rho {l1_rho:.4f} v {l1_v:.4f} u {l1_u:.4f}; "
          f"energy {e_err:.2e} momentum {p_err:.2e} mass {m_err:.2e}", flush=True)
    assert l1_rho < 0.05 and l1_v < 0.1 and l1_u < 0.1
    assert e_err < 1e-5 and m_err < 1e-12
    assert glob, "the tube's steps are global (one forced bin), as momentum conservation to roundoff needs"
    # FP64: the reference's bar.  FP32-mixed: each side of a pair rounds its own flux in FP32, so
    # the pair forces cancel to FP32 rounding (measured 1.3e-10 on the MI355X)
    assert p_err < (1e-10 if form == "fp64" else 1e-9)


@pytest.mark.parametrize("form,amp,n_side", [("fp64", 1e-6, 64), ("fp32_mixed", 1e-3, 64)])
def test_soundwave_returns(form, amp, n_side):
    ic = ics.soundwave_box(n_side, amplitude=amp, kernel=3, des_num_ngb=32.0)
    T = ic["period"]
    integ = dict(courant_fac=0.2, err_tol_int_accuracy=0.01, max_size_timestep=0.05, min_size_timestep=0.0,
                 time_begin=0.0, time_max=2 * T, min_egy_spec=0.0)
    print(f"soundwave {form}: n={n_side}^3 amplitude {amp} period {T}", flush=True)
    g = make_ctx(ic, FORMS[form])
    g.run_begin(integ)
    keys = ["rho", "vel_kick", "u_kick"]
    first = g.store(keys)
    acts = run_until_half(g)
    last = g.store(keys)
    g.close()
    scale = amp / 1e-6
    rtol, atol = 1e-5 * scale, 1e-8 * scale
    worst = {}
    for k in keys:
        a, b = last[k], first[k]
        worst[k] = float((np.abs(a - b) - rtol * np.abs(b)).max())
    vamp = np.abs(first["vel_kick"][:, 0]).max()
    print(f"soundwave {form}: {len(acts)} steps; velocity amplitude {vamp:.3e}; largest |diff| - rtol|ref|: "
          + " ".join(f"{k}={v:.2e}" for k, v in worst.items()), flush=True)
    for k in keys:
        np.testing.assert_allclose(last[k], first[k], rtol=rtol, atol=atol, err_msg=k)
