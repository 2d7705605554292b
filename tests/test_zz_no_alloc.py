"""gz_hydro_pass_noalloc, the pass that allocates nothing (its working memory is a static device
array), against the context path (gz_create + gz_load + the three stages + gz_store): the same
outputs bit for bit, in both arithmetic forms, on two calls in a row, the second with fewer
particles; and the clear error beyond its maximum size."""
import ctypes as C

import numpy as np
import pytest

from gizmo_amd import ics
from gizmo_amd.hydro import (FIELD_WIDTH, GZ_FP32_MIXED, GZ_FP64, NOALLOC_MAX_N, GizmoHipError, Hydro,
                             hydro_pass_noalloc)

pytestmark = pytest.mark.gpu

OUT = ["h", "rho", "press", "numngb", "drkern", "divvel", "nvt", "nvt_fw", "cn", "fce", "grad_rho", "grad_p", "grad_v",
       "accel", "dtu", "maxsig"]
IN = ["pos", "vel", "mass", "u", "h"]


def _context_pass(ic, prec):
    g = Hydro(dict(ic["params"]), len(ic["mass"]), prec)
    g.load({k: ic[k] for k in IN})
    g.hydro_pass()
    out = g.store(OUT)
    g.close()
    return out


class _Hip:
    """The caller's side: device buffers from the HIP runtime the library uses (the caller
    allocates its inputs and outputs; the entry point allocates nothing)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def download(self, ptr, shape):
        a = np.empty(shape, dtype=np.float64)
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return a

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc_pass(ic, prec):
    n = len(ic["mass"])
    hip = _Hip()
    try:
        din = {k: hip.upload(ic[k]) for k in IN}
        shapes = {k: (n, FIELD_WIDTH[k]) if FIELD_WIDTH[k] > 1 else (n,) for k in OUT}
        dout = {k: hip.upload(np.full(shapes[k], np.nan)) for k in OUT}
        hydro_pass_noalloc(dict(ic["params"]), n, din, dout, prec, 0)
        return {k: hip.download(dout[k], shapes[k]) for k in OUT}
    finally:
        hip.free()


@pytest.mark.parametrize("form", ["fp64", "fp32_mixed"])
def test_noalloc_pass_equals_context_pass(form):
    prec = GZ_FP64 if form == "fp64" else GZ_FP32_MIXED
    for n_side, seed in ((24, 5), (15, 6)):  # 13824 then 3375 particles
        ic = ics.lattice_box(n_side, seed, kernel=6 if n_side == 24 else 3, des_num_ngb=64.0 if n_side == 24 else 32.0,
                             mach_sigma=0.3)
        print(f"noalloc {form}: {n_side}^3 = {n_side ** 3} particles", flush=True)
        want = _context_pass(ic, prec)
        got = _noalloc_pass(ic, prec)
        for k in OUT:
            assert np.array_equal(got[k], want[k]), k


def test_noalloc_rejects_beyond_max():
    n = NOALLOC_MAX_N + 1
    ic = ics.lattice_box(4, 1, kernel=3, des_num_ngb=32.0)
    with pytest.raises(GizmoHipError) as e:
        hydro_pass_noalloc(dict(ic["params"]), n, {"pos": 1, "vel": 1, "mass": 1, "u": 1, "h": 1}, {"rho": 1})
    assert e.value.code == 1 and "outside" in str(e.value)
