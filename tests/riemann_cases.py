"""Face states that drive every rung of the Riemann ladder (shared by the oracle pins and the
device parity tests of the Riemann functions).

Cases use the layout of gz_riemann_probe / orc_riemann_probe: left (rho, p, v[3]), right
(rho, p, v[3]), unit normal n[3], pressure limiter, then the first-order states of the ladder's
retries (p_i, p_j, rho_i, rho_j, v_i[3], v_j[3]).  Random cases span many decades of density,
pressure, velocity and limiter; crafted ones reach the rungs random states do not: unphysical
input, the Kurganov-Tadmor flux (reached only through a NaN star pressure: an empty left state
with equal normal velocities), the zero-velocity retry (a first-order retry that cannot solve)
and the failure of all three.
"""
import numpy as np

IN = 24


def unit(rng, m):
    v = rng.normal(size=(m, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_cases(m, seed=11):
    rng = np.random.default_rng(seed)
    c = np.zeros((m, IN))
    rho = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), (m, 2)))
    p = np.exp(rng.uniform(np.log(1e-6), np.log(1e6), (m, 2)))
    cs = np.sqrt(p / rho)
    vs = np.exp(rng.uniform(np.log(1e-3), np.log(30.0), (m, 1))) * cs.max(axis=1, keepdims=True)
    c[:, 0] = rho[:, 0]; c[:, 1] = p[:, 0]; c[:, 2:5] = rng.normal(size=(m, 3)) * vs
    c[:, 5] = rho[:, 1]; c[:, 6] = p[:, 1]; c[:, 7:10] = rng.normal(size=(m, 3)) * vs
    c[:, 10:13] = unit(rng, m)
    c[:, 13] = np.exp(rng.uniform(np.log(1e-2), np.log(10.0), m)) * p.max(axis=1)
    # the first-order states: the face states before reconstruction (perturbed copies)
    c[:, 14] = c[:, 6] * np.exp(rng.normal(0, 0.1, m))
    c[:, 15] = c[:, 1] * np.exp(rng.normal(0, 0.1, m))
    c[:, 16] = c[:, 5] * np.exp(rng.normal(0, 0.1, m))
    c[:, 17] = c[:, 0] * np.exp(rng.normal(0, 0.1, m))
    c[:, 18:21] = c[:, 7:10] + rng.normal(size=(m, 3)) * 0.1 * vs
    c[:, 21:24] = c[:, 2:5] + rng.normal(size=(m, 3)) * 0.1 * vs
    return c


def crafted_cases():
    rows = []

    def case(L, R, n, ptl, fo=None):
        n = np.asarray(n, float)
        n = n / np.linalg.norm(n)
        if fo is None:
            fo = [R[1], L[1], R[0], L[0], *R[2:5], *L[2:5]]
        rows.append(list(L) + list(R) + list(n) + [ptl] + list(fo))

    n = [0.6, 0.0, 0.8]
    # contact: equal pressure and normal velocity, density jump (first HLLC estimate exact)
    case([1.0, 1.0, 0.3 * 0.6, 0.0, 0.3 * 0.8], [0.125, 1.0, 0.3 * 0.6, 0.0, 0.3 * 0.8], n, 10.0)
    # strong rarefaction: vacuum star state
    case([1.0, 1.0, -20 * 0.6, 0.0, -20 * 0.8], [1.0, 1.0, 20 * 0.6, 0.0, 20 * 0.8], n, 10.0)
    # vacuum star state with a limiter below it: the exact solver's vacuum sample
    case([1.0, 1.0, -20 * 0.6, 0.0, -20 * 0.8], [1.0, 1.0, 20 * 0.6, 0.0, 20 * 0.8], n, 0.0)
    # strong collision over a tight limiter: exact solver, then the first-order retry
    case([1.0, 1.0, 5 * 0.6, 0.0, 5 * 0.8], [1.0, 1.0, -5 * 0.6, 0.0, -5 * 0.8], n, 1.0)
    # negative density: unphysical input
    case([-1.0, 1.0, 0, 0, 0], [1.0, 1.0, 0, 0, 0], n, 10.0)
    # both pressures negative: unphysical input
    case([1.0, -1.0, 0, 0, 0], [1.0, -2.0, 0, 0, 0], n, 10.0)
    # empty left state, equal velocities: NaN star pressure -> Kurganov-Tadmor -> exact solver
    case([0.0, 1.0, 0.5, 0.2, 0.1], [1.0, 1.0, 0.5, 0.2, 0.1], n, 10.0)
    # both densities zero
    case([0.0, 1.0, 0.1, 0, 0], [0.0, 1.0, 0.1, 0, 0], n, 10.0)
    # reconstructed states without a solution (NaN velocity) and first-order states with a NaN
    # velocity too: the zero-velocity retry
    nan = float("nan")
    case([1.0, 1.0, nan, 0, 0], [1.0, 1.0, 0, 0, 0], n, 10.0, fo=[1.0, 1.0, 1.0, 1.0, nan, 0, 0, 0, 0, 0])
    # and with a NaN pressure in the first-order states: no solution anywhere (stop 1234)
    case([1.0, 1.0, nan, 0, 0], [1.0, 1.0, 0, 0, 0], n, 10.0, fo=[nan, 1.0, 1.0, 1.0, nan, 0, 0, 0, 0, 0])
    return np.array(rows, dtype=np.float64)


def all_cases(m=4000, seed=11):
    return np.concatenate([crafted_cases(), random_cases(m, seed)])
