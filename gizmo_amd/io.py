"""GIZMO HDF5 initial conditions and snapshots (include/gizmo_io.h through libgizmo_io.so).

The layout is the reference's: a Header group with its attributes (src/file_io/io.cc:5324-5440)
and a PartType0 group of gas datasets named as get_dataset_name (src/file_io/io.cc:4374-4800):
Coordinates, Velocities, ParticleIDs, Masses, InternalEnergy, Density, KernelMaxRadius (also
read as SmoothingLength), Pressure, HydroAcceleration, RateOfChangeOfInternalEnergy,
VelocityDivergence, DensityGradient, VelocityGradient.
"""
import ctypes as C
import os

import numpy as np

from .hydro import FIELD_WIDTH, INT_FIELDS, LOOP_FIELD_WIDTH, Fields

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class SnapHeader(C.Structure):
    _fields_ = [("num_gas", C.c_longlong), ("time", C.c_double), ("redshift", C.c_double), ("box_size", C.c_double),
                ("mass_table_gas", C.c_double), ("double_precision", C.c_int), ("num_files", C.c_int),
                ("legacy_names", C.c_int)]


class GizmoIOError(RuntimeError):
    pass


def iolib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libgizmo_io.so")
        if not os.path.exists(path):
            raise GizmoIOError(f"{path} not built (python -m gizmo_amd.build; needs the HDF5 C library)")
        L = C.CDLL(path)
        L.gz_ic_header.argtypes = [C.c_char_p, C.POINTER(SnapHeader)]
        L.gz_ic_read.argtypes = [C.c_char_p, C.c_long, C.POINTER(Fields), C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.gz_snapshot_write.argtypes = [C.c_char_p, C.c_long, C.POINTER(Fields), C.POINTER(C.c_longlong), C.POINTER(SnapHeader)]
        L.gz_io_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise GizmoIOError(f"error {rc}: {iolib().gz_io_last_error().decode()}")


def read_header(path):
    h = SnapHeader()
    _check(iolib().gz_ic_header(os.fsencode(path), C.byref(h)))
    return {k: getattr(h, k) for k, _ in SnapHeader._fields_ if k != "legacy_names"}


def read_ic(path):
    """The gas particles of an IC or snapshot file: dict of pos, vel, mass, u, rho (when
    present), h (when present), ids, and the header."""
    hdr = read_header(path)
    n = int(hdr["num_gas"])
    out = {"pos": np.zeros((n, 3)), "vel": np.zeros((n, 3)), "mass": np.zeros(n), "u": np.zeros(n), "h": np.zeros(n),
           "rho": np.full(n, np.nan)}
    f = Fields()
    for k, a in out.items():
        setattr(f, k, a.ctypes.data_as(C.POINTER(C.c_double)))
    ids = np.zeros(n, dtype=np.int64)
    has_h = C.c_int(0)
    _check(iolib().gz_ic_read(os.fsencode(path), n, C.byref(f), ids.ctypes.data_as(C.POINTER(C.c_longlong)), C.byref(has_h)))
    if not has_h.value:
        del out["h"]
    if np.isnan(out["rho"]).all():
        del out["rho"]
    out["ids"] = ids
    out["header"] = hdr
    return out


def write_snapshot(path, fields, *, time=0.0, box_size=1.0, redshift=0.0, ids=None, double_precision=True, legacy_names=False,
                   mass_table=0.0):
    """Write the gas fields of `fields` (names of gizmo_amd.hydro: pos, vel, mass, u, rho, h,
    press, accel, dtu, divvel, grad_rho, grad_v, and vel_kick / u_kick for P.Vel /
    CellP.InternalEnergy) as a GIZMO snapshot.  Without a mass field, mass_table is every gas
    particle's mass (Header/MassTable[0])."""
    n = len(fields["pos"])
    f = Fields()
    keep = []
    widths = dict(FIELD_WIDTH, **LOOP_FIELD_WIDTH)
    for k, a in fields.items():
        if k in INT_FIELDS or k not in widths:
            continue
        arr = np.ascontiguousarray(a, dtype=np.float64).reshape(n, -1)
        if arr.shape[1] != widths[k]:
            raise GizmoIOError(f"{k}: width {arr.shape[1]}, expected {widths[k]}")
        keep.append(arr)
        setattr(f, k, arr.ctypes.data_as(C.POINTER(C.c_double)))
    idp = None
    if ids is not None:
        ida = np.ascontiguousarray(ids, dtype=np.int64)
        keep.append(ida)
        idp = ida.ctypes.data_as(C.POINTER(C.c_longlong))
    h = SnapHeader(num_gas=n, time=time, redshift=redshift, box_size=box_size, mass_table_gas=mass_table,
                   double_precision=1 if double_precision else 0, num_files=1, legacy_names=1 if legacy_names else 0)
    _check(iolib().gz_snapshot_write(os.fsencode(path), n, C.byref(f), idp, C.byref(h)))
