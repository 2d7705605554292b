"""Analytic solutions the end-to-end physics tests compare the device's runs against.

  * sedov_taylor(gamma): the self-similar point explosion in a uniform cold medium (Sedov 1959;
    Landau & Lifshitz, Fluid Mechanics, sec. 106), from the similarity ODEs integrated inward
    from the strong-shock jump conditions, with the energy integral fixing xi0 in
    R(t) = xi0 (E t^2 / rho0)^(1/5).
  * riemann_sample(gamma, left, right, xi): the exact solution of the 1D Riemann problem at
    x/t = xi (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, ch. 4: the pressure
    function's Newton root, then the wave pattern sampled).
"""
import numpy as np
from scipy.integrate import quad, solve_ivp
from scipy.interpolate import interp1d


class SedovTaylor:
    """rho(r, t), v_r(r, t) of the blast of energy E in density rho0 (3D, spherical)."""

    def __init__(self, gamma, lam_min=1e-4):
        g = gamma
        self.gamma = g
        D1, U1, P1 = (g + 1) / (g - 1), 2 / (g + 1), 2 / (g + 1)

        def rhs(lam, y):
            D, U, P = y
            w = U - lam
            # continuity, momentum and entropy in the similarity variables
            A = np.array([[w, D, 0.0], [0.0, D * w, 1.0], [-g * w * P / D, 0.0, w]])
            b = np.array([-2.0 * D * U / lam, 1.5 * D * U, 3.0 * P])
            return np.linalg.solve(A, b)

        sol = solve_ivp(rhs, (1.0, lam_min), [D1, U1, P1], method="LSODA", rtol=1e-10, atol=1e-14, dense_output=True)
        lam = np.concatenate([np.geomspace(lam_min, 1e-2, 200), np.linspace(1e-2, 1.0, 4000)[1:]])
        D, U, P = sol.sol(lam)
        self.lam, self.D, self.U, self.P = lam, D, U, P
        integrand = interp1d(lam, (0.5 * D * U ** 2 + P / (g - 1)) * lam ** 2, kind="cubic")
        I = quad(integrand, lam[0], 1.0, limit=400)[0]
        self.xi0 = (25.0 / (16.0 * np.pi * I)) ** 0.2
        self._D = interp1d(lam, D, bounds_error=False, fill_value=(D[0], D[-1]))
        self._U = interp1d(lam, U, bounds_error=False, fill_value=(0.0, U[-1]))

    def shock_radius(self, E, rho0, t):
        return self.xi0 * (E * t * t / rho0) ** 0.2

    def profile(self, r, E, rho0, t):
        R = self.shock_radius(E, rho0, t)
        Rdot = 0.4 * R / t
        lam = np.asarray(r) / R
        inside = lam < 1.0
        rho = np.where(inside, rho0 * self._D(np.minimum(lam, 1.0)), rho0)
        v = np.where(inside, Rdot * self._U(np.minimum(lam, 1.0)), 0.0)
        return rho, v


def _pressure_function(p, rho, pk, ck, g):
    """Toro (4.6)-(4.7): f_K and its derivative."""
    if p > pk:  # shock
        A = 2.0 / ((g + 1) * rho)
        B = (g - 1) / (g + 1) * pk
        q = np.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    pr = p / pk  # rarefaction
    return 2 * ck / (g - 1) * (pr ** ((g - 1) / (2 * g)) - 1), 1.0 / (rho * ck) * pr ** (-(g + 1) / (2 * g))


def riemann_star(g, left, right):
    """(p*, u*) of the Riemann problem left = (rho, u, p), right = (rho, u, p)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(200):
        fl, dl = _pressure_function(p, rl, pl, cl, g)
        fr, dr = _pressure_function(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _pressure_function(p, rl, pl, cl, g)
    fr, _ = _pressure_function(p, rr, pr, cr, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def riemann_sample(g, left, right, xi):
    """rho, u, p of the exact solution at x/t = xi (array)."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    ps, us = riemann_star(g, left, right)
    xi = np.asarray(xi, dtype=np.float64)
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    gm = (g - 1) / (g + 1)
    for k, s in enumerate(xi.ravel()):
        if s <= us:  # left of the contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    v = (rl, ul, pl)
                else:
                    v = (rl * (ps / pl + gm) / (gm * ps / pl + 1), us, ps)
            else:  # left rarefaction
                shl = ul - cl
                csl = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - csl
                if s <= shl:
                    v = (rl, ul, pl)
                elif s >= stl:
                    v = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    uu = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    v = (rl * (c / cl) ** (2 / (g - 1)), uu, pl * (c / cl) ** (2 * g / (g - 1)))
        else:  # right of the contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    v = (rr, ur, pr)
                else:
                    v = (rr * (ps / pr + gm) / (gm * ps / pr + 1), us, ps)
            else:  # right rarefaction
                shr = ur + cr
                csr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + csr
                if s >= shr:
                    v = (rr, ur, pr)
                elif s <= str_:
                    v = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    uu = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    v = (rr * (c / cr) ** (2 / (g - 1)), uu, pr * (c / cr) ** (2 * g / (g - 1)))
        rho.flat[k], u.flat[k], p.flat[k] = v
    return rho, u, p
