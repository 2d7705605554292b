"""Python binding of the C ABI (include/gizmo_hydro.h) over ctypes.

``Hydro`` owns one device context: ``load`` the P[]/CellP[] fields, run ``density``,
``hydro_gradient_calc`` and ``hydro_force`` (the reference's entry points,
src/hydro/density_loop.cc:764, src/hydro/gradients_loop.cc:458,
src/hydro/hydro_force_loop.cc:423), ``store`` the results.  The binding never falls back
to another implementation: without the built library every call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgizmo_hip.so")

GZ_FP64 = 0
GZ_FP32_MIXED = 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_llp = C.POINTER(C.c_longlong)


class Params(C.Structure):
    _fields_ = [("box_size", C.c_double), ("kernel", C.c_int), ("gamma", C.c_double),
                ("des_num_ngb", C.c_double), ("max_num_ngb_dev", C.c_double),
                ("min_kernel_radius", C.c_double), ("max_kernel_radius", C.c_double),
                ("time", C.c_double), ("time_begin", C.c_double), ("precision", C.c_int)]


FIELD_WIDTH = dict(pos=3, vel=3, mass=1, h=1, u=1, numngb=1, drkern=1, divvel=1, rho=1, press=1,
                   nvt=6, nvt_fw=3, cn=1, fce=1, grad_rho=3, grad_p=3, grad_v=9, accel=3, dtu=1, maxsig=1)
INT_FIELDS = ("timebin", "wakeup")
LOOP_FIELD_WIDTH = dict(vel_kick=3, u_kick=1, soft=1)
I64_FIELDS = ("ti_begstep", "dt_step", "ti_current")
TIMEBINS = 60


class Fields(C.Structure):
    _fields_ = ([(k, _dp) for k in FIELD_WIDTH] + [(k, _ip) for k in INT_FIELDS] +
                [(k, _dp) for k in LOOP_FIELD_WIDTH] + [(k, _llp) for k in I64_FIELDS])


class IntegParams(C.Structure):
    """All.* parameters of the timestep loop (gz_integ_params)."""
    _fields_ = [("courant_fac", C.c_double), ("err_tol_int_accuracy", C.c_double),
                ("max_size_timestep", C.c_double), ("min_size_timestep", C.c_double),
                ("time_begin", C.c_double), ("time_max", C.c_double), ("min_egy_spec", C.c_double),
                ("domain_build_active_fraction", C.c_double)]


def integ_params(integ):
    """gz_integ_params from a dict; domain_build_active_fraction defaults to 0 (every step drifts
    every particle)."""
    return IntegParams(**{k: float(integ.get(k, 0.0) if k == "domain_build_active_fraction" else integ[k])
                          for k, _ in IntegParams._fields_})


class Timeline(C.Structure):
    """The integer timeline and the time bins (gz_timeline)."""
    _fields_ = [("ti_current", C.c_longlong), ("time", C.c_double), ("timebase_interval", C.c_double),
                ("num_steps", C.c_longlong), ("n_active", C.c_longlong), ("n_woken", C.c_longlong),
                ("bin_count", C.c_longlong * TIMEBINS), ("bin_active", C.c_ulonglong), ("stop_code", C.c_int),
                ("full_drift", C.c_int)]


class GizmoHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded libgizmo_hip.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GizmoHipError(-1, f"{LIB_PATH} is not built (run python -m gizmo_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.gz_create.argtypes = [C.POINTER(Params), C.c_long, C.c_int]
        L.gz_create.restype = C.c_void_p
        L.gz_destroy.argtypes = [C.c_void_p]
        L.gz_last_error.argtypes = [C.c_void_p]
        L.gz_last_error.restype = C.c_char_p
        L.gz_set_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        L.gz_load.argtypes = [C.c_void_p, C.c_long, C.POINTER(Fields)]
        L.gz_store.argtypes = [C.c_void_p, C.POINTER(Fields)]
        L.gz_set_active.argtypes = [C.c_void_p, C.POINTER(C.c_ubyte)]
        for f in ("gz_density", "gz_hydro_gradient_calc", "gz_hydro_force", "gz_synchronize", "gz_timing_reset"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.gz_permutation.argtypes = [C.c_void_p, _lp]
        L.gz_stat.argtypes = [C.c_void_p, C.c_char_p, _dp]
        L.gz_timing_enable.argtypes = [C.c_void_p, C.c_int]
        L.gz_timing.argtypes = [C.c_void_p, C.c_char_p, _lp, _dp]
        L.gz_neighbors.argtypes = [C.c_void_p, C.c_int, _lp, _ip, C.c_long]
        L.gz_neighbors.restype = C.c_long
        L.gz_run_begin.argtypes = [C.c_void_p, C.POINTER(IntegParams)]
        L.gz_run_step.argtypes = [C.c_void_p]
        L.gz_run_steps.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.gz_timeline_get.argtypes = [C.c_void_p, C.POINTER(Timeline)]
        L.gz_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.gz_riemann_probe.argtypes = [C.c_void_p, C.c_int, C.c_long, _dp, _dp, C.POINTER(C.c_uint)]
        L.gz_drift_all.argtypes = [C.c_void_p]
        L.gz_hydro_pass_noalloc.argtypes = [C.POINTER(Params), C.c_long, C.POINTER(Fields), C.POINTER(Fields), C.c_void_p,
                                            C.c_char_p, C.c_size_t]
        _lib = L
    return _lib


def fields_struct(fields, n):
    """A gz_fields of the given host arrays (dict name -> array, n particles in caller order);
    returns it with the contiguous copies it points into (keep them alive for the call)."""
    f = Fields()
    keep = []
    for k, w in list(FIELD_WIDTH.items()) + list(LOOP_FIELD_WIDTH.items()):
        if k in fields and fields[k] is not None:
            a = np.ascontiguousarray(fields[k], dtype=np.float64).reshape(n * w)
            keep.append(a)
            setattr(f, k, a.ctypes.data_as(_dp))
    for k in INT_FIELDS:
        if k in fields and fields[k] is not None:
            a = np.ascontiguousarray(fields[k], dtype=np.int32).reshape(n)
            keep.append(a)
            setattr(f, k, a.ctypes.data_as(_ip))
    for k in I64_FIELDS:
        if k in fields and fields[k] is not None:
            a = np.ascontiguousarray(fields[k], dtype=np.int64).reshape(n)
            keep.append(a)
            setattr(f, k, a.ctypes.data_as(_llp))
    return f, keep


def output_struct(names, n):
    """A gz_fields pointing at fresh host arrays for the named fields (all P[]/CellP[] fields
    and the two int fields when names is None); returns it and the dict of arrays."""
    names = list(FIELD_WIDTH) + list(INT_FIELDS) if names is None else list(names)
    f = Fields()
    out = {}
    widths = dict(FIELD_WIDTH, **LOOP_FIELD_WIDTH)
    for k in names:
        if k in widths:
            w = widths[k]
            a = np.zeros((n, w) if w > 1 else n)
            setattr(f, k, a.ctypes.data_as(_dp))
        elif k in I64_FIELDS:
            a = np.zeros(n, dtype=np.int64)
            setattr(f, k, a.ctypes.data_as(_llp))
        elif k in INT_FIELDS:
            a = np.zeros(n, dtype=np.int32)
            setattr(f, k, a.ctypes.data_as(_ip))
        else:
            raise KeyError(k)
        out[k] = a
    return f, out


def timeline_dict(t):
    return dict(ti_current=t.ti_current, time=t.time, timebase_interval=t.timebase_interval,
                num_steps=t.num_steps, n_active=t.n_active, n_woken=t.n_woken,
                bin_count=np.array(t.bin_count[:], dtype=np.int64), bin_active=t.bin_active,
                stop_code=t.stop_code, full_drift=t.full_drift)


def make_params(p, precision=GZ_FP64):
    return Params(box_size=p["box"], kernel=p["kernel"], gamma=p["gamma"], des_num_ngb=p["des_num_ngb"],
                  max_num_ngb_dev=p["max_num_ngb_dev"], min_kernel_radius=p["min_kernel_radius"],
                  max_kernel_radius=p["max_kernel_radius"], time=p["time"], time_begin=p["time_begin"],
                  precision=precision)


class Hydro:
    """One periodic box resident on one GPU."""

    def __init__(self, params, capacity, precision=GZ_FP64, device=0):
        self._p = make_params(params, precision)
        self.ctx = lib().gz_create(C.byref(self._p), int(capacity), int(device))
        if not self.ctx:
            raise GizmoHipError(-1, "gz_create failed (no device, or parameters rejected)")
        self.n = 0

    def close(self):
        if self.ctx:
            lib().gz_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise GizmoHipError(rc, lib().gz_last_error(self.ctx).decode())

    def set_params(self, params, precision=None):
        prec = self._p.precision if precision is None else precision
        self._p = make_params(params, prec)
        self._check(lib().gz_set_params(self.ctx, C.byref(self._p)))

    def load(self, fields):
        """fields: dict name -> numpy array (caller order). pos, vel, mass, u, h required."""
        n = len(fields["mass"])
        f, keep = fields_struct(fields, n)
        self._check(lib().gz_load(self.ctx, n, C.byref(f)))
        self.n = n

    def store(self, names=None):
        f, out = output_struct(names, self.n)
        self._check(lib().gz_store(self.ctx, C.byref(f)))
        return out

    def set_active(self, active=None):
        if active is None:
            self._check(lib().gz_set_active(self.ctx, None))
        else:
            a = np.ascontiguousarray(np.asarray(active, dtype=np.uint8))
            self._check(lib().gz_set_active(self.ctx, a.ctypes.data_as(C.POINTER(C.c_ubyte))))

    def density(self):
        self._check(lib().gz_density(self.ctx))

    def hydro_gradient_calc(self):
        self._check(lib().gz_hydro_gradient_calc(self.ctx))

    def hydro_force(self):
        self._check(lib().gz_hydro_force(self.ctx))

    def hydro_pass(self):
        self.density()
        self.hydro_gradient_calc()
        self.hydro_force()

    def synchronize(self):
        self._check(lib().gz_synchronize(self.ctx))

    def permutation(self):
        p = np.zeros(self.n, dtype=np.int64)
        self._check(lib().gz_permutation(self.ctx, p.ctypes.data_as(_lp)))
        return p

    def stat(self, name):
        v = C.c_double()
        self._check(lib().gz_stat(self.ctx, name.encode(), C.byref(v)))
        return v.value

    def timing_enable(self, on=True):
        self._check(lib().gz_timing_enable(self.ctx, 1 if on else 0))

    def timing_reset(self):
        self._check(lib().gz_timing_reset(self.ctx))

    def timing(self, name):
        cnt, ms = C.c_long(), C.c_double()
        self._check(lib().gz_timing(self.ctx, name.encode(), C.byref(cnt), C.byref(ms)))
        return cnt.value, ms.value

    # ---- the timestep loop (gz_run_begin / gz_run_step) ----
    def run_begin(self, integ):
        """integ: dict of the IntegParams fields (All.CourantFac, ErrTolIntAccuracy, ...)."""
        self._ip = integ_params(integ)
        self._check(lib().gz_run_begin(self.ctx, C.byref(self._ip)))

    def run_step(self):
        """One sync point of the run loop; returns the number of active particles updated."""
        self._check(lib().gz_run_step(self.ctx))
        return self.timeline()["n_active"]

    def run_steps(self, nsteps):
        """nsteps sync points in one call (no return to Python between them); returns the active
        and woken counts of each step."""
        act = np.zeros(max(int(nsteps), 1), dtype=np.int64)
        woken = np.zeros_like(act)
        self._check(lib().gz_run_steps(self.ctx, int(nsteps), act.ctypes.data_as(C.POINTER(C.c_longlong)),
                                       woken.ctypes.data_as(C.POINTER(C.c_longlong))))
        return act[:nsteps].tolist(), woken[:nsteps].tolist()

    def drift_all(self):
        """Every particle behind the current sync point drifted to it (gizmo_full_drift_to)."""
        self._check(lib().gz_drift_all(self.ctx))

    def set_option(self, name, value):
        self._check(lib().gz_set_option(self.ctx, name.encode(), float(value)))

    def timeline(self):
        t = Timeline()
        self._check(lib().gz_timeline_get(self.ctx, C.byref(t)))
        return timeline_dict(t)

    RUNGS = ("riemann_hllc", "riemann_vacuum", "riemann_roe", "riemann_pvrs", "riemann_kt", "riemann_exact",
             "riemann_exact_vacuum", "riemann_unphysical", "riemann_retry_first_order", "riemann_retry_zero_velocity",
             "riemann_fail", "riemann_faces", "entropic_considered", "entropic_applied")

    def rung_counts(self):
        """The Riemann rung counts of the force passes since the "riemann_stats" option was set."""
        return {k: int(self.stat(k)) for k in self.RUNGS}

    def riemann_probe(self, mode, cases):
        """The device's Riemann functions on face states (n x 24, see gz_riemann_probe): returns
        (out n x 8, rung bits n)."""
        c = np.ascontiguousarray(cases, dtype=np.float64).reshape(-1, 24)
        n = len(c)
        out = np.zeros((n, 8))
        rung = np.zeros(n, dtype=np.uint32)
        self._check(lib().gz_riemann_probe(self.ctx, int(mode), n, c.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                           rung.ctypes.data_as(C.POINTER(C.c_uint))))
        return out, rung

    def neighbors(self, symmetric):
        off = np.zeros(self.n + 1, dtype=np.int64)
        tot = lib().gz_neighbors(self.ctx, int(symmetric), off.ctypes.data_as(_lp), None, 0)
        if tot < 0:
            self._check(-tot)
        nb = np.zeros(max(tot, 1), dtype=np.int32)
        tot2 = lib().gz_neighbors(self.ctx, int(symmetric), off.ctypes.data_as(_lp), nb.ctypes.data_as(_ip), tot)
        if tot2 != tot:
            raise GizmoHipError(-1, "neighbour count changed between calls")
        return off, nb[:tot]


NOALLOC_MAX_N = 16384  # GZ_NOALLOC_MAX_N


def _device_fields(ptrs):
    """A gz_fields of device addresses (dict name -> int address, e.g. a tensor's data_ptr())."""
    f = Fields()
    kinds = dict({k: _dp for k in list(FIELD_WIDTH) + list(LOOP_FIELD_WIDTH)}, **{k: _ip for k in INT_FIELDS},
                 **{k: _llp for k in I64_FIELDS})
    for k, a in ptrs.items():
        if k not in kinds:
            raise KeyError(k)
        if a:
            setattr(f, k, C.cast(C.c_void_p(int(a)), kinds[k]))
    return f


def hydro_pass_noalloc(params, n, dev_in, dev_out, precision=GZ_FP64, stream=0):
    """One density -> gradients -> force pass over device buffers through gz_hydro_pass_noalloc
    (no device allocation: the library's static scratch).  dev_in / dev_out: dict field name ->
    device address of n values of the field's width in caller order (gz_load's and gz_store's
    layouts); stream: a hipStream_t as an int (0: the null stream)."""
    p = make_params(params, precision)
    fi = _device_fields(dev_in)
    fo = _device_fields(dev_out)
    err = C.create_string_buffer(512)
    rc = lib().gz_hydro_pass_noalloc(C.byref(p), int(n), C.byref(fi), C.byref(fo), C.c_void_p(int(stream)), err, len(err))
    if rc:
        raise GizmoHipError(rc, err.value.decode())
