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


def gradient_box(n_side, *, k=2.0, gamma=5.0 / 3.0, pressure=1.0, kernel=6, des_num_ngb=64.0, box=1.0):
    """Particles whose number density grows as exp(k x) along x: a cubic lattice of n_side^3
    points at cell centres in (xi, y, z), mapped along x by x = box ln(1 + xi (e^k - 1)) / k
    (so that d xi / dx is proportional to exp(k x / box)); equal masses (total mass box^3),
    uniform pressure, at rest.  The periodic box joins the densest and the most dilute slabs at
    x = 0; away from that face, a kernel radius that keeps the neighbour count fixed varies as
    n^(-1/3): dh/dx = -k h / (3 box)."""
    n = n_side ** 3
    d = box / n_side
    g = (np.arange(n_side, dtype=np.float64) + 0.5) * d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xi = X.ravel() / box
    x = box * np.log1p(xi * np.expm1(k)) / k
    pos = np.stack([x, Y.ravel(), Z.ravel()], axis=1) % box
    mass = np.full(n, box ** 3 / n)
    nloc = n * k * np.exp(k * x / box) / (np.expm1(k) * box ** 3)  # particles per volume at x
    h = (des_num_ngb / (VOLNORM * nloc)) ** (1.0 / 3.0)
    rho = mass * nloc
    u = pressure / ((gamma - 1.0) * rho)
    params = dict(box=box, kernel=kernel, gamma=gamma, des_num_ngb=float(des_num_ngb),
                  max_num_ngb_dev=max(des_num_ngb / 640.0, 0.05),
                  min_kernel_radius=0.0, max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    return dict(pos=np.ascontiguousarray(pos), vel=np.zeros((n, 3)), mass=mass, u=u, h=h, params=params, k=k)


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


def _gaussian_field(ng, slope, rng, box=1.0):
    """A real Gaussian random field on an ng^3 periodic grid with power spectrum P(k) ~ k^slope
    (the k = 0 mode removed), unit variance; returned with the wave vectors for projections."""
    k1 = np.fft.fftfreq(ng, d=box / ng) * 2.0 * np.pi
    kz1 = np.fft.rfftfreq(ng, d=box / ng) * 2.0 * np.pi
    kx, ky, kz = np.meshgrid(k1, k1, kz1, indexing="ij", sparse=True)
    k2 = kx * kx + ky * ky + kz * kz
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = k2[nz] ** (0.25 * slope)  # sqrt(P(k)) = k^(slope/2)
    shape = (ng, ng, ng // 2 + 1)
    white = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    return white * amp, (kx, ky, kz, k2)


def lognormal_box(n_particles, seed, *, ng=256, sigma_ln=2.0, slope=-3.0, mach=2.0, gamma=1.001, kernel=6,
                  des_num_ngb=64.0, box=1.0):
    """Particles sampled from a lognormal density field: delta is a Gaussian random field with
    P(k) ~ k^slope on an ng^3 grid, normalised to std(delta) = sigma_ln, rho ~ exp(delta); the
    n_particles (equal mass, total mass 1) are spread over the cells by a multinomial draw with
    probabilities rho_c / sum(rho), uniformly inside each cell.  Isothermal: P = rho with
    gamma = 1.001 (u = 1/(gamma-1)).  Velocities: a solenoidal Gaussian field (three fields of
    the same spectrum, longitudinal part projected out) at the particle's cell, scaled to rms
    Mach `mach` (3D, c_s = sqrt(gamma P/rho)).  The first radius guess encloses des_num_ngb
    neighbours at the cell's density."""
    rng = _rng(seed)
    fk, (kx, ky, kz, k2) = _gaussian_field(ng, slope, rng, box)
    delta = np.fft.irfftn(fk, s=(ng, ng, ng), axes=(0, 1, 2))
    delta -= delta.mean()
    delta *= sigma_ln / delta.std()
    rho = np.exp(delta).ravel()
    del delta, fk
    prob = rho / rho.sum()
    counts = rng.multinomial(n_particles, prob)
    cells = np.repeat(np.arange(ng ** 3, dtype=np.int64), counts)
    del prob
    dx = box / ng
    ci = np.stack(np.unravel_index(cells, (ng, ng, ng)), axis=1).astype(np.float64)
    pos = (ci + rng.uniform(0.0, 1.0, size=(n_particles, 3))) * dx
    pos %= box
    del ci
    # solenoidal velocity field
    vk = [rng.normal(size=(ng, ng, ng // 2 + 1)) + 1j * rng.normal(size=(ng, ng, ng // 2 + 1)) for _ in range(3)]
    amp = np.zeros_like(k2)
    nz = k2 > 0
    amp[nz] = k2[nz] ** (0.25 * slope)
    vk = [v * amp for v in vk]
    kd = np.where(nz, (kx * vk[0] + ky * vk[1] + kz * vk[2]) / np.where(nz, k2, 1.0), 0.0)
    vk = [vk[0] - kx * kd, vk[1] - ky * kd, vk[2] - kz * kd]
    del kd, amp
    vel = np.empty((n_particles, 3))
    for c in range(3):
        vg = np.fft.irfftn(vk[c], s=(ng, ng, ng), axes=(0, 1, 2)).ravel()
        vel[:, c] = vg[cells]
        del vg
    del vk
    cs = np.sqrt(gamma)  # P = rho: c_s^2 = gamma P / rho
    vel *= mach * cs / np.sqrt((vel * vel).sum(axis=1).mean())
    mass = np.full(n_particles, 1.0 / n_particles)
    u = np.full(n_particles, 1.0 / (gamma - 1.0))
    nloc = rho[cells] / rho.mean() * n_particles / box ** 3
    h = (des_num_ngb / (VOLNORM * nloc)) ** (1.0 / 3.0)
    h = np.minimum(h, 0.2 * box)
    params = dict(box=box, kernel=kernel, gamma=gamma, des_num_ngb=float(des_num_ngb),
                  max_num_ngb_dev=max(des_num_ngb / 640.0, 0.05),
                  min_kernel_radius=0.0, max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    return dict(pos=np.ascontiguousarray(pos), vel=np.ascontiguousarray(vel), mass=mass, u=u, h=h, params=params)


def shocktube_box(n_dense, seed, *, gamma=1.4, left=(1.0, 1.0), right=(0.125, 0.1), jitter=0.02, kernel=3,
                  des_num_ngb=32.0, box=1.0):
    """Sod's shock tube in a periodic box (the reference's shocktube test, 3D): the dense state
    (rho, P) = left fills the slab box/4 <= x < 3 box/4, the light state (right) the rest, at rest,
    gamma 1.4.  Equal-mass particles on two cubic lattices: n_dense per box side in the slab and
    n_dense * (rho_right/rho_left)^(1/3) (rounded) outside, each coordinate jittered by
    U(-jitter, +jitter) spacings.  The slab's faces at 3 box/4 and box/4 are Sod's problem and its
    mirror image."""
    rng = _rng(seed)
    (rl, pl), (rr, pr) = left, right
    n_light = int(round(n_dense * (rr / rl) ** (1.0 / 3.0)))
    parts = []
    for n_side, x0, x1 in ((n_dense, 0.25 * box, 0.75 * box), (n_light, 0.75 * box, 1.25 * box)):
        dx = box / n_side
        nx = int(round((x1 - x0) / dx))
        gx = x0 + (np.arange(nx) + 0.5) * dx
        g = (np.arange(n_side) + 0.5) * dx
        X, Y, Z = np.meshgrid(gx, g, g, indexing="ij")
        p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        p += rng.uniform(-jitter, jitter, size=p.shape) * dx
        parts.append(p % box)
    pos = np.concatenate(parts)
    n = len(pos)
    dense = np.zeros(n, dtype=bool)
    dense[:len(parts[0])] = True
    m = rl * 0.5 * box ** 3 / len(parts[0])
    mass = np.full(n, m)
    u = np.where(dense, pl / ((gamma - 1.0) * rl), pr / ((gamma - 1.0) * rr))
    nloc = np.where(dense, rl, rr) / m
    h = (des_num_ngb / (VOLNORM * nloc)) ** (1.0 / 3.0)
    params = dict(box=box, kernel=kernel, gamma=gamma, des_num_ngb=float(des_num_ngb),
                  max_num_ngb_dev=max(des_num_ngb / 640.0, 0.05),
                  min_kernel_radius=0.0, max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    return dict(pos=np.ascontiguousarray(pos), vel=np.zeros((n, 3)), mass=mass, u=u, h=h, params=params,
                dense=dense)


def soundwave_box(n_side, *, amplitude=1e-6, gamma=5.0 / 3.0, kernel=3, des_num_ngb=32.0, box=1.0):
    """A linear sound wave (the reference's soundwave test, 3D): a perfect cubic lattice (an
    equilibrium of the pair forces) at density 1 and pressure 3/5 (c_s = 1 for gamma 5/3) with a
    plane wave of wavelength box along x: particle displacement (amplitude/k) cos(k x0), so that
    rho = 1 + amplitude sin(k x); velocity c_s amplitude sin(k x) along x (a wave moving to +x);
    adiabatic energy u0 (rho)^(gamma-1).  One period takes box / c_s."""
    n = n_side ** 3
    dx = box / n_side
    g = (np.arange(n_side, dtype=np.float64) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k = 2.0 * np.pi / box
    x0 = pos[:, 0].copy()
    pos[:, 0] = (x0 + amplitude / k * np.cos(k * x0)) % box
    p0 = 1.0 / gamma
    cs = np.sqrt(gamma * p0)
    s = np.sin(k * x0)
    vel = np.zeros((n, 3))
    vel[:, 0] = cs * amplitude * s
    u = p0 / (gamma - 1.0) * (1.0 + amplitude * s) ** (gamma - 1.0)
    mass = np.full(n, box ** 3 / n)
    h = np.full(n, h_guess(n, des_num_ngb, box))
    params = dict(box=box, kernel=kernel, gamma=gamma, des_num_ngb=float(des_num_ngb),
                  max_num_ngb_dev=max(des_num_ngb / 640.0, 0.05),
                  min_kernel_radius=0.0, max_kernel_radius=0.25 * box, time=0.0, time_begin=0.0)
    return dict(pos=np.ascontiguousarray(pos), vel=vel, mass=mass, u=u, h=h, params=params, period=box / cs)


def workload(name, seed=None):
    """The named single-box workloads: 'c1' (32^3 cubic-spline lattice) and 'c2' (128^3 Sedov)."""
    if name == "c1":
        return lattice_box(32, 1001 if seed is None else seed, kernel=3, des_num_ngb=32.0)
    if name == "c2":
        return sedov_box(128, 1002 if seed is None else seed, kernel=6, des_num_ngb=64.0)
    raise ValueError(f"unknown workload {name!r}")
