"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use this module as the checker or the timed CPU reference.  The product package never
imports it.

``State`` holds the P[]/CellP[] hot-path fields as numpy arrays in the caller's index order;
``density``, ``gradients`` and ``force`` restate density(), hydro_gradient_calc() and
hydro_force() on it in place.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_up = C.POINTER(C.c_ubyte)
_llp = C.POINTER(C.c_longlong)


class Params(C.Structure):
    _fields_ = [("box", C.c_double), ("kernel", C.c_int), ("gamma", C.c_double),
                ("des_num_ngb", C.c_double), ("max_num_ngb_dev", C.c_double),
                ("min_kernel_radius", C.c_double), ("max_kernel_radius", C.c_double),
                ("time", C.c_double), ("time_begin", C.c_double), ("fd_correction", C.c_int)]


class _CState(C.Structure):
    _fields_ = [("n", C.c_long), ("pos", _dp), ("mass", _dp), ("h", _dp), ("numngb", _dp),
                ("drkern", _dp), ("divvel", _dp), ("timebin", _ip), ("wakeup", _ip),
                ("vel", _dp), ("u", _dp), ("rho", _dp), ("press", _dp), ("nvt", _dp),
                ("nvt_fw", _dp), ("cn", _dp), ("fce", _dp), ("grad_rho", _dp), ("grad_p", _dp),
                ("grad_v", _dp), ("accel", _dp), ("dtu", _dp), ("maxsig", _dp),
                ("vel_kick", _dp), ("u_kick", _dp), ("soft", _dp), ("ti_begstep", _llp), ("dt_step", _llp),
                ("ti_current", _llp)]


TIMEBINS = 60


class Integ(C.Structure):
    """All.* parameters of the timestep loop."""
    _fields_ = [("courant_fac", C.c_double), ("err_tol_int_accuracy", C.c_double),
                ("max_size_timestep", C.c_double), ("min_size_timestep", C.c_double),
                ("time_begin", C.c_double), ("time_max", C.c_double), ("min_egy_spec", C.c_double),
                ("domain_build_active_fraction", C.c_double)]


class Timeline(C.Structure):
    """The integer timeline and the time bins (All.Ti_Current, TimeBinCount, TimeBinActive ...)."""
    _fields_ = [("ti_current", C.c_longlong), ("time", C.c_double), ("timebase_interval", C.c_double),
                ("num_steps", C.c_longlong), ("bin_count", C.c_longlong * TIMEBINS),
                ("bin_active", C.c_ubyte * TIMEBINS), ("need_wakeup", C.c_int), ("stop_code", C.c_int),
                ("n_active", C.c_longlong), ("n_woken", C.c_longlong), ("full_drift", C.c_int)]


# field name -> columns per particle
FIELDS = dict(pos=3, mass=1, h=1, numngb=1, drkern=1, divvel=1, vel=3, u=1, rho=1, press=1,
              nvt=6, nvt_fw=3, cn=1, fce=1, grad_rho=3, grad_p=3, grad_v=9, accel=3, dtu=1, maxsig=1,
              vel_kick=3, u_kick=1, soft=1)
INT_FIELDS = ("timebin", "wakeup")
I64_FIELDS = ("ti_begstep", "dt_step", "ti_current")


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB)
        L.orc_kernel.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_fd_alpha.argtypes = [C.c_int]
        L.orc_fd_alpha.restype = C.c_double
        L.orc_neighbors.argtypes = [C.POINTER(Params), C.c_long, _dp, _dp, C.c_int, _up, _lp, _ip, C.c_long]
        L.orc_neighbors.restype = C.c_long
        L.orc_matrix_invert.argtypes = [_dp, _dp]
        L.orc_matrix_invert.restype = C.c_double
        L.orc_density.argtypes = [C.POINTER(Params), C.POINTER(_CState), _up, _ip]
        L.orc_density_fixed_h.argtypes = [C.POINTER(Params), C.POINTER(_CState), _up, _dp]
        L.orc_slopelimiter.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_int, C.c_double, C.c_double]
        L.orc_gradients.argtypes = [C.POINTER(Params), C.POINTER(_CState), _up]
        L.orc_reconstruct.argtypes = [C.c_double, _dp, C.c_double, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_riemann.argtypes = [C.c_double, _dp, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_riemann_exact.argtypes = [C.c_double, _dp, _dp, _dp, _dp]
        L.orc_riemann_probe.argtypes = [C.c_int, C.c_double, C.c_long, _dp, _dp, C.POINTER(C.c_uint)]
        L.orc_rung_counts.argtypes = [_llp, C.c_int]
        L.orc_rung_counts.restype = None
        L.orc_force.argtypes = [C.POINTER(Params), C.POINTER(_CState), _up]
        L.orc_force_ex.argtypes = [C.POINTER(Params), C.POINTER(_CState), _up, _ip]
        L.orc_get_timestep.argtypes = [C.POINTER(Params), C.POINTER(Integ), C.POINTER(Timeline), C.POINTER(_CState), C.c_long]
        L.orc_get_timestep.restype = C.c_longlong
        for f in (L.orc_run_begin, L.orc_run_step, L.orc_drift_all):
            f.argtypes = [C.POINTER(Params), C.POINTER(Integ), C.POINTER(Timeline), C.POINTER(_CState)]
        L.orc_drift_all.restype = None
        L.orc_kick.argtypes = [C.POINTER(Params), C.POINTER(Integ), C.POINTER(Timeline), C.POINTER(_CState), C.c_long,
                               C.c_longlong, C.c_longlong, C.c_int]
        L.orc_drift.argtypes = [C.POINTER(Params), C.POINTER(Integ), C.POINTER(Timeline), C.POINTER(_CState), C.c_long,
                                C.c_longlong, C.c_longlong]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


PARAM_DEFAULTS = {"fd_correction": 0}
NACC = 17  # density accumulators per particle (OR_NACC)


def params(d):
    return Params(**{k: d.get(k, PARAM_DEFAULTS.get(k)) if k in PARAM_DEFAULTS else d[k] for k, _ in Params._fields_})


class State:
    """P[]/CellP[] hot-path fields of n particles (numpy, caller's order)."""

    def __init__(self, n):
        self.n = int(n)
        for k, c in FIELDS.items():
            setattr(self, k, np.zeros((self.n, c)) if c > 1 else np.zeros(self.n))
        self.timebin = np.zeros(self.n, dtype=np.int32)
        self.wakeup = np.zeros(self.n, dtype=np.int32)
        self.ti_begstep = np.zeros(self.n, dtype=np.int64)
        self.dt_step = np.zeros(self.n, dtype=np.int64)
        self.ti_current = np.zeros(self.n, dtype=np.int64)

    ALL = list(FIELDS) + list(INT_FIELDS) + list(I64_FIELDS)

    @classmethod
    def from_ic(cls, ic):
        s = cls(len(ic["mass"]))
        for k in ("pos", "vel", "mass", "u", "h"):
            getattr(s, k)[...] = ic[k]
        s.vel_kick[...] = ic["vel"]
        s.u_kick[...] = ic["u"]
        return s

    def copy(self):
        s = State(self.n)
        for k in self.ALL:
            getattr(s, k)[...] = getattr(self, k)
        return s

    def permuted(self, perm):
        """A copy with particle k taken from index perm[k]."""
        s = State(self.n)
        for k in self.ALL:
            getattr(s, k)[...] = getattr(self, k)[perm]
        return s

    def _c(self):
        for k in self.ALL:
            a = getattr(self, k)
            if not (a.flags.c_contiguous and a.dtype in (np.float64, np.int32, np.int64)):
                setattr(self, k, np.ascontiguousarray(a))
        cs = _CState()
        cs.n = self.n
        for k in FIELDS:
            setattr(cs, k, _ptr(getattr(self, k)))
        cs.timebin = _ptr(self.timebin, _ip)
        cs.wakeup = _ptr(self.wakeup, _ip)
        cs.ti_begstep = _ptr(self.ti_begstep, _llp)
        cs.dt_step = _ptr(self.dt_step, _llp)
        cs.ti_current = _ptr(self.ti_current, _llp)
        return cs


def _mask(active, n):
    if active is None:
        return None, None
    a = np.ascontiguousarray(np.asarray(active, dtype=np.uint8))
    assert a.shape == (n,)
    return a, _ptr(a, _up)


def kernel(kid, r, h):
    wk, dwk = C.c_double(), C.c_double()
    rc = lib().orc_kernel(kid, r, h, C.byref(wk), C.byref(dwk))
    if rc:
        raise ValueError("kernel")
    return wk.value, dwk.value


def fd_alpha(kid):
    return float(lib().orc_fd_alpha(int(kid)))


def neighbors(par, pos, h, symmetric, rowmask=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    n = len(h)
    p = params(par)
    off = np.zeros(n + 1, dtype=np.int64)
    m, mp = _mask(rowmask, n)
    tot = lib().orc_neighbors(C.byref(p), n, _ptr(pos), _ptr(h), int(symmetric), mp, _ptr(off, _lp), None, 0)
    if tot < 0:
        raise RuntimeError("orc_neighbors failed")
    nb = np.zeros(max(tot, 1), dtype=np.int32)
    tot2 = lib().orc_neighbors(C.byref(p), n, _ptr(pos), _ptr(h), int(symmetric), mp, _ptr(off, _lp), _ptr(nb, _ip), tot)
    assert tot2 == tot
    return off, nb[:tot]


def matrix_invert(T):
    T = np.ascontiguousarray(T, dtype=np.float64).reshape(9)
    Ti = np.zeros(9)
    cn = lib().orc_matrix_invert(_ptr(T), _ptr(Ti))
    return cn, Ti.reshape(3, 3)


def density(par, st, active=None, return_iters=False):
    p = params(par)
    cs = st._c()
    m, mp = _mask(active, st.n)
    it = np.zeros(st.n, dtype=np.int32)
    rc = lib().orc_density(C.byref(p), C.byref(cs), mp, _ptr(it, _ip))
    if rc:
        raise RuntimeError(f"orc_density stop code {rc}")
    return it if return_iters else None


def density_fixed_h(par, st, active=None):
    p = params(par)
    cs = st._c()
    m, mp = _mask(active, st.n)
    raw = np.zeros((st.n, NACC))
    rc = lib().orc_density_fixed_h(C.byref(p), C.byref(cs), mp, _ptr(raw))
    if rc:
        raise RuntimeError(f"orc_density_fixed_h failed {rc}")
    return raw


def gradients(par, st, active=None):
    p = params(par)
    cs = st._c()
    m, mp = _mask(active, st.n)
    rc = lib().orc_gradients(C.byref(p), C.byref(cs), mp)
    if rc:
        raise RuntimeError(f"orc_gradients failed {rc}")


def force(par, st, active=None):
    p = params(par)
    cs = st._c()
    m, mp = _mask(active, st.n)
    rc = lib().orc_force(C.byref(p), C.byref(cs), mp)
    if rc:
        raise RuntimeError(f"orc_force stop code {rc}")


def slopelimiter(g, valmax, valmin, alim, h, shoot_tol, pos_preserve, d_max, val_cen):
    g = np.array(g, dtype=np.float64)
    lib().orc_slopelimiter(_ptr(g), valmax, valmin, alim, h, shoot_tol, pos_preserve, d_max, val_cen)
    return g


def reconstruct(Qi, Gi, Qj, Gj, di, dj, mode=1):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (Gi, Gj, di, dj)]
    QL, QR = C.c_double(), C.c_double()
    lib().orc_reconstruct(Qi, _ptr(a[0]), Qj, _ptr(a[1]), _ptr(a[2]), _ptr(a[3]), C.byref(QL), C.byref(QR), mode)
    return QL.value, QR.value


def riemann(gamma, left, right, n, ptl):
    L = np.ascontiguousarray(left, dtype=np.float64)
    R = np.ascontiguousarray(right, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    PM, SM, fp = C.c_double(), C.c_double(), C.c_double()
    fv = np.zeros(3)
    rc = lib().orc_riemann(gamma, _ptr(L), _ptr(R), _ptr(nn), ptl, C.byref(PM), C.byref(SM), _ptr(fv), C.byref(fp))
    return rc, PM.value, SM.value, fv, fp.value


def riemann_exact(gamma, left, right):
    L = np.ascontiguousarray(left, dtype=np.float64)
    R = np.ascontiguousarray(right, dtype=np.float64)
    PM, SM = C.c_double(), C.c_double()
    lib().orc_riemann_exact(gamma, _ptr(L), _ptr(R), C.byref(PM), C.byref(SM))
    return PM.value, SM.value


# the rungs of the Riemann ladder, in bit order (the device's gz_stat names)
RUNGS = ("riemann_hllc", "riemann_vacuum", "riemann_roe", "riemann_pvrs", "riemann_kt", "riemann_exact",
         "riemann_exact_vacuum", "riemann_unphysical", "riemann_retry_first_order", "riemann_retry_zero_velocity",
         "riemann_fail", "riemann_faces", "entropic_considered", "entropic_applied")
PROBE_IN, PROBE_OUT = 24, 8


def riemann_probe(mode, gamma, cases):
    """The Riemann functions on face states (n x 24 cases, the layout of gz_riemann_probe):
    returns (out n x 8, rung bits n)."""
    c = np.ascontiguousarray(cases, dtype=np.float64).reshape(-1, PROBE_IN)
    n = len(c)
    out = np.zeros((n, PROBE_OUT))
    rung = np.zeros(n, dtype=np.uint32)
    rc = lib().orc_riemann_probe(int(mode), float(gamma), n, _ptr(c), _ptr(out), rung.ctypes.data_as(C.POINTER(C.c_uint)))
    if rc:
        raise ValueError("orc_riemann_probe mode")
    return out, rung


def rung_counts(reset=True):
    """The force passes' rung counts since the last reset, by name."""
    v = np.zeros(len(RUNGS), dtype=np.int64)
    lib().orc_rung_counts(_ptr(v, _llp), int(reset))
    return {k: int(x) for k, x in zip(RUNGS, v)}


def hydro_pass(par, st, active=None):
    """density() -> hydro_gradient_calc() -> hydro_force() on st, in place."""
    density(par, st, active)
    gradients(par, st, active)
    force(par, st, active)


def num_threads():
    return lib().orc_num_threads()


def integ(d):
    """orc_integ from a dict; domain_build_active_fraction defaults to 0 (every step drifts every particle)."""
    return Integ(**{k: float(d.get(k, 0.0) if k == "domain_build_active_fraction" else d[k]) for k, _ in Integ._fields_})


class Run:
    """The timestep loop of run() on an oracle State: begin() then step() per sync point."""

    def __init__(self, par, ip, st):
        self.par = params(par)
        self.ip = integ(ip)
        self.tl = Timeline()
        self.st = st

    def begin(self):
        cs = self.st._c()
        rc = lib().orc_run_begin(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs))
        if rc:
            raise RuntimeError(f"orc_run_begin stop code {rc}")

    def step(self):
        cs = self.st._c()
        rc = lib().orc_run_step(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs))
        if rc:
            raise RuntimeError(f"orc_run_step stop code {rc}")
        return int(self.tl.n_active)

    def drift_all(self):
        """gizmo_full_drift_to(All.Ti_Current): every particle behind the sync point drifted to it."""
        cs = self.st._c()
        lib().orc_drift_all(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs))

    def process_wake_ups(self):
        """process_wake_ups alone on the current timeline and state."""
        cs = self.st._c()
        lib().orc_process_wake_ups(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs))

    def get_timestep(self, i):
        cs = self.st._c()
        return int(lib().orc_get_timestep(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs), int(i)))

    def kick(self, i, tstart, tend, mode):
        cs = self.st._c()
        lib().orc_kick(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs), int(i), int(tstart), int(tend), int(mode))

    def drift(self, i, time0, time1):
        cs = self.st._c()
        lib().orc_drift(C.byref(self.par), C.byref(self.ip), C.byref(self.tl), C.byref(cs), int(i), int(time0), int(time1))

    @property
    def ti_current(self):
        return int(self.tl.ti_current)

    @property
    def time(self):
        return float(self.tl.time)
