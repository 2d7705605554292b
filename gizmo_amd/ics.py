"""Seeded initial conditions for the hydro workloads.

Every generator returns a dict of numpy arrays holding the particle fields that the hot
path reads (``pos``, ``vel``, ``mass``, ``u``, ``h``) plus the run parameters, in the
caller's index order.  Random numbers come from numpy's counter-based Philox generator
keyed by the seed, so every process (tests, bench, the CPU baseline) sees identical bits.

Workloads (the reference's own test geometry is a jittered periodic lattice,
initial_conditions/make_IC.py):
  * ``lattice_box``: 3D periodic unit box, n^3 equal-mass particles on a cubic lattice,
    each coordinate jittered by U(-jitter, +jitter) lattice spacings; uniform density 1 and
    pressure P; Gaussian velocities of dispersion ``mach_sigma`` * c_s per component.
  * ``sedov_box``: the cold lattice with energy E deposited kernel-weighted (Wendland C2
    weight W(r/h_dep), h_dep = distance to the 65th-nearest particle) into the 64 particles
    nearest the box centre.
"""
import numpy as np

VOLNORM = 4.188790204786  # 4 pi / 3 as the reference writes it (src/declarations/constants.h:445)


def _rng(seed):
    return np.random.Generator(np.random.Philox(int(seed)))


def h_guess(n_total, des_num_ngb, box=1.0):
    """Initial kernel radius enclosing des_num_ngb neighbours at the mean number density."""
    return (des_num_ngb * box ** 3 / (VOLNORM * n_total)) ** (1.0 / 3.0)


def lattice_box(n_side, seed, *, gamma=5.0 / 3.0, pressure=1.0, density=1.0, mach_sigma=0.1,
                jitter=0.1, kernel=3, des_num_ngb=32.0, box=1.0):
    rng = _rng(seed)
    n = n_side ** 3
    dx = box / n_side
    g = (np.arange(n_side, dtype=np.float64) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pos += rng.uniform(-jitter, jitter, size=(n, 3)) * dx
    pos %= box
    cs = np.sqrt(gamma * pressure / density)
    vel = rng.normal(0.0, mach_sigma * cs, size=(n, 3)) if mach_sigma > 0 else np.zeros((n, 3))
    mass = np.full(n, density * box ** 3 / n)
    u = np.full(n, pressure / ((gamma - 1.0) * density))
    h = np.full(n, h_guess(n, des_num_ngb, box))
    params = dict(box=box, kernel=kernel, gamma=gamma, des_num_ngb=float(des_num_ngb),
                  max_num_ngb_dev=max(des_num_ngb / 640.0, 0.05),
                  min_kernel_radius=0.0, max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    return dict(pos=np.ascontiguousarray(pos), vel=np.ascontiguousarray(vel), mass=mass, u=u, h=h,
                params=params)


def _wendland_c2_shape(q):
    q = np.clip(q, 0.0, 1.0)
    return (1.0 - q) ** 4 * (1.0 + 4.0 * q)


def sedov_box(n_side, seed, *, gamma=5.0 / 3.0, pressure=1e-6, energy=1.0, n_deposit=64, jitter=0.1,
              kernel=6, des_num_ngb=64.0, box=1.0):
    ic = lattice_box(n_side, seed, gamma=gamma, pressure=pressure, mach_sigma=0.0, jitter=jitter,
                     kernel=kernel, des_num_ngb=des_num_ngb, box=box)
    pos, mass, u = ic["pos"], ic["mass"], ic["u"]
    d = pos - 0.5 * box
    r = np.sqrt((d * d).sum(axis=1))
    order = np.argsort(r, kind="stable")
    near = order[:n_deposit]
    h_dep = r[order[n_deposit]]
    w = _wendland_c2_shape(r[near] / h_dep)
    w /= w.sum()
    u[near] += energy * w / mass[near]
    return ic


def workload(name, seed=None):
    """The named single-box workloads: 'c1' (32^3 cubic-spline lattice) and 'c2' (128^3 Sedov)."""
    if name == "c1":
        return lattice_box(32, 1001 if seed is None else seed, kernel=3, des_num_ngb=32.0)
    if name == "c2":
        return sedov_box(128, 1002 if seed is None else seed, kernel=6, des_num_ngb=64.0)
    raise ValueError(f"unknown workload {name!r}")
