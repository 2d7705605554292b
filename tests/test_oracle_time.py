"""Pins of the oracle's timestep loop (oracle/gz_oracle_time.c) by what the mathematics fixes:

  * kick-drift-kick is exact for a constant acceleration: x(T) = x0 + v0 T + a T^2 / 2 and
    v(T) = v0 + a T after any number of steps of any length;
  * the drift's predicted density, radius and energy follow their closed forms
    (rho0 exp(-divv dt), h0 exp(divv dt / 3), u0 + du/dt dt), compose over split intervals,
    cap the log change at 0.3 and keep P = (gamma - 1) u rho;
  * the kick's positivity catch halves the energy instead of letting it fall below half;
  * get_timestep: the maximum step binds exactly; the acceleration criterion scales as
    |a|^(-1/2) and the Courant criterion as the kernel radius (bins shift by one);
  * the run loop with one forced bin advances every particle every step by the same power of
    two, and conserves total momentum to roundoff (antisymmetric pair forces, equal kicks).
"""
import math

import numpy as np
import pytest

from gizmo_amd import ics

TBI = 1.0 / (1 << 40)


def one_particle(O, pos=(10.0, 20.0, 30.0), vel=(0.3, -0.2, 0.1), acc=(0.5, 0.25, -1.0), mass=2.0):
    s = O.State(1)
    s.pos[0] = pos
    s.vel[0] = vel
    s.vel_kick[0] = vel
    s.accel[0] = acc
    s.mass[0] = mass
    s.u[0] = s.u_kick[0] = 1.0
    s.rho[0] = 1.0
    s.h[0] = 0.5
    s.numngb[0] = 4.0
    s.maxsig[0] = 1.0
    s.soft[0] = 0.5
    return s


def runner(O, s, box=1e6, **integ):
    par = dict(box=box, kernel=6, gamma=5.0 / 3.0, des_num_ngb=64.0, max_num_ngb_dev=0.1, min_kernel_radius=0.0,
               max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    ip = dict(courant_fac=0.2, err_tol_int_accuracy=0.02, max_size_timestep=0.01, min_size_timestep=0.0,
              time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
    ip.update(integ)
    r = O.Run(par, ip, s)
    r.tl.timebase_interval = TBI
    for k in range(60):
        r.tl.bin_active[k] = 1
    return r


def test_kdk_exact_for_constant_acceleration(oracle_lib):
    O = oracle_lib
    s = one_particle(O)
    r = runner(O, s)
    x0, v0, a = s.pos[0].copy(), s.vel_kick[0].copy(), s.accel[0].copy()
    t = 0
    for dt in (1 << 20, 1 << 22, 1 << 18, 1 << 22, 1 << 21):
        half = dt // 2
        r.kick(0, t, t + half, 0)
        r.drift(0, t, t + dt)
        r.kick(0, t + half, t + dt, 1)
        t += dt
    T = t * TBI
    assert np.allclose(s.pos[0], x0 + v0 * T + 0.5 * a * T * T, rtol=0, atol=1e-12 * np.abs(x0).max())
    assert np.allclose(s.vel_kick[0], v0 + a * T, rtol=1e-14, atol=1e-16)
    assert np.array_equal(s.vel[0], s.vel_kick[0])  # mode 1: predicted = kicked


@pytest.mark.parametrize("divv", [0.7, -1.3])
def test_drift_closed_forms_and_composition(oracle_lib, divv):
    O = oracle_lib
    out = []
    for split in (False, True):
        s = one_particle(O)
        s.divvel[0] = divv
        s.dtu[0] = -0.25
        r = runner(O, s)
        T = 1 << 36  # dt = 1/16: |divv dt| < 0.3
        if split:
            r.drift(0, 0, T // 2)
            r.drift(0, T // 2, T)
        else:
            r.drift(0, 0, T)
        out.append(s)
    a, b = out
    dt = (1 << 36) * TBI
    assert a.rho[0] == pytest.approx(math.exp(-divv * dt), rel=1e-14)
    assert a.h[0] == pytest.approx(0.5 * math.exp(divv * dt / 3), rel=1e-14)
    assert a.u[0] == pytest.approx(1.0 - 0.25 * dt, rel=1e-14)
    assert a.press[0] == pytest.approx((5.0 / 3.0 - 1) * a.u[0] * a.rho[0], rel=1e-15)
    for k in ("rho", "h", "u", "press"):
        assert getattr(a, k)[0] == pytest.approx(getattr(b, k)[0], rel=1e-14), k
    assert np.allclose(a.pos, b.pos, rtol=1e-15)


def test_drift_log_change_capped(oracle_lib):
    O = oracle_lib
    s = one_particle(O)
    s.divvel[0] = 100.0
    r = runner(O, s)
    r.drift(0, 0, 1 << 36)
    assert s.rho[0] == math.exp(-0.3)
    assert s.h[0] == 0.5 * math.exp(0.3 / 3)


def test_kick_energy_positivity(oracle_lib):
    O = oracle_lib
    s = one_particle(O)
    s.dtu[0] = -1e6
    r = runner(O, s)
    r.kick(0, 0, 1 << 30, 1)
    assert s.u_kick[0] == 0.5 and s.u[0] == 0.5


def test_get_timestep_limits_and_scalings(oracle_lib):
    O = oracle_lib
    # the maximum step binds: a tiny acceleration, a slow signal, no divergence
    s = one_particle(O, acc=(1e-12, 0, 0))
    s.maxsig[0] = 1e-9
    r = runner(O, s, max_size_timestep=0.01)
    assert r.get_timestep(0) == int(0.01 / TBI)
    # acceleration criterion alone: sqrt(2 eta core h / |a|) scales as |a|^(-1/2)
    s = one_particle(O, acc=(1.0, 0, 0))
    s.maxsig[0] = 1e-9
    r = runner(O, s, max_size_timestep=1e3, err_tol_int_accuracy=0.02)
    t1 = r.get_timestep(0)
    s.accel[0] = (4.0, 0, 0)
    t4 = r.get_timestep(0)
    assert abs(t1 - 2 * t4) <= 2
    assert t1 * TBI == pytest.approx(math.sqrt(2 * 0.02 * 0.4 * 0.5 / 1.0), rel=1e-9)
    # Courant criterion alone scales as the particle size (kernel radius)
    s = one_particle(O, acc=(1e-12, 0, 0))
    r = runner(O, s, max_size_timestep=1e3, err_tol_int_accuracy=1e30)
    c1 = r.get_timestep(0)
    s.h[0] *= 2
    c2 = r.get_timestep(0)
    assert abs(c2 - 2 * c1) <= 2


def test_run_loop_single_bin_conserves_momentum(oracle_lib):
    O = oracle_lib
    ic = ics.lattice_box(8, 4001, kernel=3, des_num_ngb=32.0, mach_sigma=0.3)
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    # the forced single bin: MaxSizeTimestep = MinSizeTimestep below every Courant step
    integ = dict(courant_fac=0.2, err_tol_int_accuracy=0.02, max_size_timestep=1e-4, min_size_timestep=1e-4,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
    r = O.Run(par, integ, st)
    r.begin()
    p0 = (st.mass[:, None] * st.vel_kick).sum(axis=0)
    scale = (st.mass[:, None] * np.abs(st.vel_kick)).sum()
    tis = []
    for _ in range(6):
        assert r.step() == st.n
        tis.append(r.ti_current)
    steps = np.diff([0] + tis)
    assert len(set(steps)) == 1 and steps[0] & (steps[0] - 1) == 0
    p1 = (st.mass[:, None] * st.vel_kick).sum(axis=0)
    assert np.abs(p1 - p0).max() <= 1e-13 * scale


# ------------------------------------------------------------------------------------------
# the lazy drift (src/core/run.cc:979-994, src/core/predict.cc:175-210,
# src/mesh/gpu_neighbor_list.cc:2443-2510) and the wake-up bin rule
# ------------------------------------------------------------------------------------------
SEDOV_INTEG = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                   time_begin=0.0, time_max=1.0, min_egy_spec=0.0)


def _sedov_run(O, frac, steps, n_side=14):
    ic = ics.sedov_box(n_side, 1002, kernel=6, des_num_ngb=64.0)
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    r = O.Run(par, dict(SEDOV_INTEG, domain_build_active_fraction=frac), st)
    r.begin()
    hist = []
    for _ in range(steps):
        na = r.step()
        hist.append((r.ti_current, na, int(r.tl.n_woken), int(r.tl.full_drift), tuple(r.tl.bin_count[:])))
    return r, st, par, hist


def test_lazy_drift_reads_only_current_particles(oracle_lib):
    """After every lazy step, no particle still behind the sync point is a partner of an active
    particle under the pair test at the current time (its predicted position and radius): every
    particle a pass read was drifted first, and the active ones always are."""
    O = oracle_lib
    ic = ics.sedov_box(14, 1002, kernel=6, des_num_ngb=64.0)
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    r = O.Run(par, dict(SEDOV_INTEG, domain_build_active_fraction=1.0), st)
    r.begin()
    lazy_steps = 0
    box = par["box"]
    for _ in range(25):
        r.step()
        now = r.ti_current
        act = np.array([r.tl.bin_active[b] for b in st.timebin], dtype=bool) & (st.mass > 0)
        behind = st.ti_current != now
        assert not (act & behind).any()
        if r.tl.full_drift:
            assert not behind.any()
            continue
        lazy_steps += 1
        assert behind.any(), "a lazy step of a Sedov blast leaves the cold medium behind"
        # predicted positions/radii of the particles behind (the oracle's own drift, applied to a copy)
        cp = st.copy()
        rc = O.Run(par, dict(SEDOV_INTEG, domain_build_active_fraction=1.0), cp)
        rc.tl = r.tl
        rc.drift_all()
        ai = np.flatnonzero(act)
        bj = np.flatnonzero(behind)
        d = cp.pos[ai][:, None, :] - cp.pos[bj][None, :, :]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(-1)
        hm = np.maximum(cp.h[ai][:, None], cp.h[bj][None, :])
        assert not (r2 < hm * hm).any(), "a partner of an active particle was left behind"
    assert lazy_steps >= 10


def test_lazy_drift_agrees_with_eager_to_rounding(oracle_lib):
    """Lazy (a particle drifted in one piece when read) and eager (every particle at every sync
    point) runs take the same integer timeline and agree to rounding once every particle is
    drifted to the end: the cold medium the lazy run leaves behind has zero velocity divergence
    and energy rate, so neither the 0.3 log cap nor the positivity catch can bind on it."""
    O = oracle_lib
    ra, sa, par, ha = _sedov_run(O, 0.0, 30)
    rb, sb, _, hb = _sedov_run(O, 0.5, 30)
    assert [h[:3] + h[4:] for h in ha] == [h[:3] + h[4:] for h in hb]
    assert all(h[3] == 1 for h in ha) and sum(1 - h[3] for h in hb) >= 20
    ra.drift_all()
    rb.drift_all()
    box = par["box"]
    d = sa.pos - sb.pos
    assert np.abs(d - box * np.round(d / box)).max() < 1e-13
    for k in ("vel", "rho", "u", "h", "press", "accel", "dtu"):
        a, b = getattr(sa, k), getattr(sb, k)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), k


def test_lazy_full_drift_at_the_domain_cadence(oracle_lib):
    """A step whose active bins hold more than DomainBuild_ActiveFraction of the particles drifts
    every particle (src/core/run.cc:360,380-390), any other step only the active bins."""
    O = oracle_lib
    ic = ics.sedov_box(14, 1002, kernel=6, des_num_ngb=64.0)
    par = dict(ic["params"])
    st = O.State.from_ic(ic)
    r = O.Run(par, dict(SEDOV_INTEG, domain_build_active_fraction=0.5), st)
    r.begin()
    seen = set()
    for _ in range(25):
        bins_before = np.array(r.tl.bin_count[:])
        r.step()
        active = np.array([r.tl.bin_active[b] for b in range(60)], dtype=bool)
        nf = int(np.array(r.tl.bin_count[:])[active].sum())
        full = int(r.tl.full_drift)
        assert full == int(nf > 0.5 * st.n), (nf, full)
        seen.add(full)
        if not full:
            inactive = ~np.array([r.tl.bin_active[b] for b in st.timebin], dtype=bool)
            assert (st.ti_current[inactive] < r.ti_current).any()
    assert 0 in seen
    # DomainBuild_ActiveFraction 0: every step drifts every particle
    st = O.State.from_ic(ic)
    r = O.Run(par, dict(SEDOV_INTEG, domain_build_active_fraction=0.0), st)
    r.begin()
    for _ in range(10):
        r.step()
        assert r.tl.full_drift == 1 and (st.ti_current == r.ti_current).all()


def test_wake_up_bins_steps_and_kicks(oracle_lib):
    """process_wake_ups (src/core/timestep.cc:1565-1774) on a hand-made state at Ti_Current =
    3 * 2^10 (bins 0-10 active; the next kick is bin 8's at 3328 = 13 * 2^8, so bin 8 is the
    highest bin active there).  The bin rule, step truncation and kicks are checked case by case,
    and the woken particle's velocity and energy against the closed form of a constant
    acceleration and energy rate: kicked from its old step's start to the midpoint of its new
    step, whether the wake-up falls before the old midpoint or not."""
    O = oracle_lib
    T = 3 << 10
    n = 7
    s = O.State(n)
    s.mass[:] = 1.0
    s.rho[:] = 1.0
    s.h[:] = 0.1
    s.u[:] = s.u_kick[:] = 1.0
    s.accel[:] = (0.5, -0.25, 1.0)
    s.dtu[:] = 0.01
    v_start = np.array([0.1, 0.2, -0.3])
    bins = [8, 14, 12, 14, 14, 13, 10]
    begstep = [3072 - 256, 0, 0, 0, 0, 0, 2048]
    dt = [256, 1 << 14, 1 << 12, 1 << 14, 1 << 14, 1 << 13, 1024]
    tic = [T, T, T, T, T, 2048, T]
    # waker bin + 1 (hydro wake-up), -1 (generic), 0 (none)
    wake = [0, 9, 16, -1, 0, 9, 9]
    for i in range(n):
        s.timebin[i] = bins[i]
        s.ti_begstep[i] = begstep[i]
        s.dt_step[i] = dt[i]
        s.ti_current[i] = tic[i]
        s.wakeup[i] = wake[i]
        # the first half-kick of the old step applied: kicked to its midpoint
        mid = begstep[i] + dt[i] // 2
        s.vel_kick[i] = v_start + s.accel[i] * mid * TBI
        s.u_kick[i] = 1.0 + s.dtu[i] * mid * TBI
        s.vel[i] = s.vel_kick[i]
    before = s.copy()
    r = runner(O, s)
    r.tl.ti_current = T
    for b in range(60):
        r.tl.bin_active[b] = 1 if (b == 0 or T % (1 << b) == 0) else 0
        r.tl.bin_count[b] = bins.count(b)
    r.tl.need_wakeup = 1
    r.process_wake_ups()
    # 1: waker bin 8 - 3 = 5, raised to the lowest occupied active bin (8); 2: waker bin 15 - 3 =
    # 12, lowered to the highest bin active at the next kick (8); 3: generic, that same bin; 5:
    # as 1
    assert list(s.timebin) == [8, 8, 8, 8, 14, 8, 10]
    assert r.tl.n_woken == 4
    assert [r.tl.bin_count[b] for b in (8, 10, 12, 13, 14)] == [5, 1, 0, 0, 1]
    # drifted to now and on a full bin step: the new step starts now, with the old step completed
    # (1, 3: the wake-up falls before the old midpoint; 2: after it) and the new one's first
    # half-kick; the predicted velocity and energy are those of now
    for i in (1, 2, 3):
        assert (s.ti_begstep[i], s.dt_step[i], s.ti_current[i]) == (T, 256, T)
        np.testing.assert_allclose(s.vel_kick[i], v_start + s.accel[i] * (T + 128) * TBI, rtol=0, atol=1e-15)
        assert abs(s.u_kick[i] - (1.0 + s.dtu[i] * (T + 128) * TBI)) < 1e-15
        np.testing.assert_allclose(s.vel[i], v_start + s.accel[i] * T * TBI, rtol=0, atol=1e-15)
        assert abs(s.u[i] - (1.0 + s.dtu[i] * T * TBI)) < 1e-15
    # not drifted to now: the step ends now (truncated from where the particle stands), no kick
    assert (s.ti_begstep[5], s.dt_step[5], s.ti_current[5]) == (2048, T - 2048, T)
    np.testing.assert_array_equal(s.vel_kick[5], before.vel_kick[5])
    # untouched: no flag (4), active (6), the waker (0)
    for i in (0, 4, 6):
        assert (s.timebin[i], s.ti_begstep[i], s.dt_step[i], s.ti_current[i]) == (bins[i], begstep[i], dt[i], tic[i])
        np.testing.assert_array_equal(s.vel_kick[i], before.vel_kick[i])
