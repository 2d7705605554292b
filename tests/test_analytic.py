"""Pins of the analytic solutions the physics tests use (tests/analytic.py).

  * Sedov-Taylor: the energy constant xi0 takes its published values (1.1517 for gamma 5/3,
    1.0328 for 7/5), the profile starts at the strong-shock jump (rho = rho0 (g+1)/(g-1),
    v = 2/(g+1) dR/dt) and its density falls to zero at the centre; the energy the profile holds
    equals E.
  * exact Riemann sampler: Toro's Sod problem (test 1: p* = 0.30313, u* = 0.92745) and the star
    pressures of his tests 2-5 (Table 4.2), and the oracle's exact solver's star state on random
    problems.
"""
import numpy as np
import pytest

import analytic as A


@pytest.mark.parametrize("gamma,xi0", [(5.0 / 3.0, 1.1517), (1.4, 1.0328)])
def test_sedov_taylor_constants_and_energy(gamma, xi0):
    s = A.SedovTaylor(gamma)
    assert abs(s.xi0 - xi0) < 5e-4, s.xi0
    E, rho0, t = 2.0, 3.0, 0.7
    R = s.shock_radius(E, rho0, t)
    rho, v = s.profile(np.array([R * (1 - 1e-9)]), E, rho0, t)
    assert abs(rho[0] / rho0 - (gamma + 1) / (gamma - 1)) < 1e-6
    assert abs(v[0] - 2.0 / (gamma + 1) * 0.4 * R / t) < 1e-6 * v[0]
    rc, _ = s.profile(np.array([1e-3 * R]), E, rho0, t)
    assert rc[0] < 1e-3 * rho0
    # the energy inside the shock: kinetic plus thermal, from the similarity profile of P too
    lam = s.lam
    Rdot = 0.4 * R / t
    e = (0.5 * rho0 * s.D * (Rdot * s.U) ** 2 + rho0 * Rdot ** 2 * s.P / (gamma - 1)) * 4 * np.pi * (lam * R) ** 2
    assert abs(np.trapezoid(e, lam * R) / E - 1.0) < 2e-3


TORO = [  # (left rho,u,p), (right rho,u,p), p*, u*  (Toro Table 4.2, gamma 1.4)
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.00000),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975),
]


@pytest.mark.parametrize("case", range(len(TORO)))
def test_riemann_star_toro(case):
    L, R, ps, us = TORO[case]
    p, u = A.riemann_star(1.4, L, R)
    assert abs(p - ps) <= 5e-5 * ps + 1e-5
    assert abs(u - us) <= 5e-5 * abs(us) + 1e-5


def test_riemann_sample_sod_states():
    rho, u, p = A.riemann_sample(1.4, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), np.array([-2.0, 0.5, 1.2, 1.70, 2.0]))
    assert np.allclose([rho[0], u[0], p[0]], [1.0, 0.0, 1.0])
    assert np.allclose([rho[4], u[4], p[4]], [0.125, 0.0, 0.1])
    # between the contact (0.927) and the shock (1.752): the shocked right state
    assert abs(rho[2] - 0.26557) < 1e-4 and abs(u[2] - 0.92745) < 1e-4 and abs(p[2] - 0.30313) < 1e-4
    # left star state between the rarefaction's tail (-0.0703) and the contact
    assert abs(rho[1] - 0.42632) < 1e-4


def test_riemann_star_agrees_with_oracle_exact_solver(oracle_lib):
    rng = np.random.default_rng(2)
    for _ in range(50):
        L = (np.exp(rng.uniform(-2, 2)), rng.normal(), np.exp(rng.uniform(-2, 2)))
        R = (np.exp(rng.uniform(-2, 2)), rng.normal(), np.exp(rng.uniform(-2, 2)))
        cl, cr = np.sqrt(1.4 * L[2] / L[0]), np.sqrt(1.4 * R[2] / R[0])
        if R[1] - L[1] >= 2 / 0.4 * (cl + cr):
            continue
        p, u = A.riemann_star(1.4, L, R)
        po, uo = oracle_lib.riemann_exact(1.4, [L[0], L[2], L[1]], [R[0], R[2], R[1]])
        assert abs(p - po) <= 1e-5 * p and abs(u - uo) <= 1e-5 * (abs(u) + cl + cr)
