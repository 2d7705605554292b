"""Python binding of the batch C ABI (include/gizmo_batch.h) over ctypes.

``Batch`` holds many independent periodic boxes on one GPU and advances all of them with
shared launches: ``add`` a box (its particles and its own box size, radius limits and
timestep-loop parameters), ``begin`` the boxes added since the last call (init() and the first
hot-path pass), ``step`` every started box to its own next sync point, ``store`` a box's fields,
``remove`` a box.  Each box follows the single-box run loop of ``Hydro.run_begin`` /
``Hydro.run_step`` and computes the same numbers alone or in any mix of boxes.  No fallback:
without the built library every call raises.
"""
import ctypes as C

from .hydro import (GZ_FP64, Fields, GizmoHipError, IntegParams, Params, Timeline, fields_struct, integ_params, lib, make_params,
                    output_struct, timeline_dict)


class BoxParams(C.Structure):
    _fields_ = [("box_size", C.c_double), ("min_kernel_radius", C.c_double), ("max_kernel_radius", C.c_double),
                ("integ", IntegParams)]


_bound = False


def blib():
    global _bound
    L = lib()
    if not _bound:
        L.gz_batch_create.argtypes = [C.POINTER(Params), C.c_long, C.c_int, C.c_int]
        L.gz_batch_create.restype = C.c_void_p
        L.gz_batch_destroy.argtypes = [C.c_void_p]
        L.gz_batch_last_error.argtypes = [C.c_void_p]
        L.gz_batch_last_error.restype = C.c_char_p
        L.gz_batch_add.argtypes = [C.c_void_p, C.c_long, C.POINTER(Fields), C.POINTER(BoxParams), C.POINTER(C.c_int)]
        for f in ("gz_batch_begin", "gz_batch_step", "gz_batch_synchronize", "gz_batch_timing_reset"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.gz_batch_remove.argtypes = [C.c_void_p, C.c_int]
        L.gz_batch_store.argtypes = [C.c_void_p, C.c_int, C.POINTER(Fields)]
        L.gz_batch_timeline.argtypes = [C.c_void_p, C.c_int, C.POINTER(Timeline)]
        L.gz_batch_particles.argtypes = [C.c_void_p]
        L.gz_batch_particles.restype = C.c_long
        L.gz_batch_boxes.argtypes = [C.c_void_p]
        L.gz_batch_stat.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double)]
        L.gz_batch_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        _bound = True
    return L


def box_params(params, integ):
    """gz_box_params from a parameter dict (box, min/max_kernel_radius) and an integ dict."""
    ip = integ_params(integ)
    return BoxParams(box_size=float(params["box"]), min_kernel_radius=float(params["min_kernel_radius"]),
                     max_kernel_radius=float(params["max_kernel_radius"]), integ=ip)


class Batch:
    """Independent periodic boxes resident on one GPU, stepped together."""

    def __init__(self, params, capacity, max_boxes, precision=GZ_FP64, device=0):
        self._p = make_params(params, precision)
        self.ctx = blib().gz_batch_create(C.byref(self._p), int(capacity), int(max_boxes), int(device))
        if not self.ctx:
            raise GizmoHipError(-1, "gz_batch_create failed (no device, no memory, or parameters rejected)")
        self.sizes = {}

    def close(self):
        if self.ctx:
            blib().gz_batch_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise GizmoHipError(rc, blib().gz_batch_last_error(self.ctx).decode())

    def add(self, fields, params, integ):
        """Load a box; returns its id.  fields as Hydro.load; params: box, min/max_kernel_radius."""
        n = len(fields["mass"])
        f, keep = fields_struct(fields, n)
        bp = box_params(params, integ)
        bid = C.c_int(-1)
        self._check(blib().gz_batch_add(self.ctx, n, C.byref(f), C.byref(bp), C.byref(bid)))
        self.sizes[bid.value] = n
        return bid.value

    def begin(self):
        self._check(blib().gz_batch_begin(self.ctx))

    def step(self):
        self._check(blib().gz_batch_step(self.ctx))

    def remove(self, box_id):
        self._check(blib().gz_batch_remove(self.ctx, int(box_id)))
        self.sizes.pop(box_id, None)

    def store(self, box_id, names=None):
        f, out = output_struct(names, self.sizes[box_id])
        self._check(blib().gz_batch_store(self.ctx, int(box_id), C.byref(f)))
        return out

    def timeline(self, box_id):
        t = Timeline()
        self._check(blib().gz_batch_timeline(self.ctx, int(box_id), C.byref(t)))
        return timeline_dict(t)

    def particles(self):
        return blib().gz_batch_particles(self.ctx)

    def boxes(self):
        return blib().gz_batch_boxes(self.ctx)

    def synchronize(self):
        self._check(blib().gz_batch_synchronize(self.ctx))

    def stat(self, name):
        v = C.c_double()
        self._check(blib().gz_batch_stat(self.ctx, name.encode(), C.byref(v)))
        return v.value

    def set_option(self, name, value):
        """Tuning options ("few_rows"): the same lists and sums whatever their value."""
        self._check(blib().gz_batch_set_option(self.ctx, name.encode(), float(value)))

    def timing_reset(self):
        self._check(blib().gz_batch_timing_reset(self.ctx))
