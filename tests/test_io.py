"""GIZMO HDF5 ICs and snapshots (include/gizmo_io.h): the reference's layout, checked with the
HDF5 command-line tools (an independent reader) and by round trips.

Bars:
  * the datasets and header attributes carry the reference's names and shapes
    (src/file_io/io.cc:4374-4800, :5324-5440), as h5ls / h5dump list them;
  * FP64 snapshots read back bit for bit; FP32 snapshots read back as the FP32 rounding;
  * an IC in the form the reference's own test generators write (int32 particle counts, uint32
    IDs, no kernel radius; test/palpha_shock/make_ics.py:14-29) reads; the legacy kernel-radius
    name SmoothingLength (io.cc:4848-4866) reads; MassTable masses read;
  * wrong particle counts and missing groups fail with a message.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gizmo_amd import io

H5LS = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
tools = pytest.mark.skipif(not (os.path.exists(H5LS) and os.path.exists(H5DUMP)), reason="HDF5 command-line tools absent")


def fields(n, seed=7):
    rng = np.random.default_rng(seed)
    return dict(pos=rng.uniform(0, 1, (n, 3)), vel=rng.normal(size=(n, 3)), mass=rng.uniform(1, 2, n), u=rng.uniform(1, 3, n),
                rho=rng.uniform(0.5, 2, n), h=rng.uniform(0.01, 0.1, n), press=rng.uniform(0.1, 1, n), accel=rng.normal(size=(n, 3)),
                dtu=rng.normal(size=n), divvel=rng.normal(size=n), grad_rho=rng.normal(size=(n, 3)), grad_v=rng.normal(size=(n, 9)))


REFERENCE_NAMES = {"pos": "Coordinates", "vel": "Velocities", "mass": "Masses", "u": "InternalEnergy", "rho": "Density",
                   "h": "KernelMaxRadius", "press": "Pressure", "accel": "HydroAcceleration", "dtu": "RateOfChangeOfInternalEnergy",
                   "divvel": "VelocityDivergence", "grad_rho": "DensityGradient", "grad_v": "VelocityGradient"}


def h5ls(path):
    out = subprocess.run([H5LS, "-r", path], capture_output=True, text=True, check=True).stdout
    return {line.split()[0]: line.split(None, 1)[1].strip() for line in out.splitlines() if line.strip()}


@tools
def test_layout_names_and_shapes(tmp_path):
    n = 100
    p = str(tmp_path / "snap_000.hdf5")
    io.write_snapshot(p, fields(n), time=0.25, box_size=2.0, ids=np.arange(10, 10 + n))
    ls = h5ls(p)
    assert ls["/Header"].startswith("Group") and ls["/PartType0"].startswith("Group")
    for k, name in REFERENCE_NAMES.items():
        width = fields(1)[k].size
        shape = f"{{{n}}}" if width == 1 else f"{{{n}, {width}}}"
        assert ls[f"/PartType0/{name}"] == f"Dataset {shape}", (name, ls.get(f"/PartType0/{name}"))
    assert ls["/PartType0/ParticleIDs"] == f"Dataset {{{n}}}"
    attrs = subprocess.run([H5DUMP, "-A", "-g", "/Header", p], capture_output=True, text=True, check=True).stdout
    for a in ("NumPart_ThisFile", "NumPart_Total", "NumPart_Total_HighWord", "MassTable", "Time", "Redshift", "BoxSize",
              "NumFilesPerSnapshot", "Flag_DoublePrecision"):
        assert f'ATTRIBUTE "{a}"' in attrs, a
    time_block = attrs[attrs.index('ATTRIBUTE "Time"'):]
    assert "0.25" in time_block[:400]


def test_fp64_round_trip(tmp_path):
    n = 257
    p = str(tmp_path / "snap.hdf5")
    f = fields(n)
    io.write_snapshot(p, f, time=1.5, box_size=1.0, ids=np.arange(n)[::-1] + 1)
    r = io.read_ic(p)
    for k in ("pos", "vel", "mass", "u", "rho", "h"):
        assert np.array_equal(r[k], f[k]), k
    assert np.array_equal(r["ids"], np.arange(n)[::-1] + 1)
    hdr = r["header"]
    assert hdr["num_gas"] == n and hdr["time"] == 1.5 and hdr["box_size"] == 1.0 and hdr["double_precision"] == 1
    assert hdr["num_files"] == 1


def test_fp32_snapshot_reads_as_fp32_rounding(tmp_path):
    n = 64
    p = str(tmp_path / "snap32.hdf5")
    f = fields(n, seed=3)
    io.write_snapshot(p, f, double_precision=False)
    r = io.read_ic(p)
    for k in ("pos", "vel", "mass", "u", "rho", "h"):
        assert np.array_equal(r[k], f[k].astype(np.float32).astype(np.float64)), k
    assert r["header"]["double_precision"] == 0


def test_kick_fields_and_legacy_name_and_mass_table(tmp_path):
    n = 50
    p = str(tmp_path / "legacy.hdf5")
    f = fields(n, seed=11)
    vel_kick, u_kick = f["vel"] + 1.0, f["u"] * 2.0
    g = {k: f[k] for k in ("pos", "vel", "u", "h")}
    g.update(vel_kick=vel_kick, u_kick=u_kick)
    io.write_snapshot(p, g, legacy_names=True, mass_table=0.125)
    r = io.read_ic(p)
    assert np.array_equal(r["vel"], vel_kick) and np.array_equal(r["u"], u_kick)   # P.Vel, CellP.InternalEnergy
    assert np.array_equal(r["h"], f["h"])                                           # read under SmoothingLength
    assert np.array_equal(r["mass"], np.full(n, 0.125))                             # from MassTable[0]
    assert np.array_equal(r["ids"], np.arange(1, n + 1))
    if os.path.exists(H5LS):
        ls = h5ls(p)
        assert "/PartType0/SmoothingLength" in ls and "/PartType0/Masses" not in ls


LIBHDF5 = "/opt/conda/lib/libhdf5.so"


def write_reference_style_ic(path, pos, vel, mass, u, box):
    """An IC as the reference's own test generators write it with h5py
    (test/palpha_shock/make_ics.py:14-29): int32 NumPart_ThisFile, uint32 NumPart_Total, FP64
    datasets, uint32 ParticleIDs, no kernel radius, Flag_DoublePrecision 1.  Written through the
    HDF5 C library directly (ctypes), not through this project's writer."""
    import ctypes as C
    L = C.CDLL(LIBHDF5)
    L.H5open()
    hid = C.c_int64
    for fn in ("H5Fcreate", "H5Gcreate2", "H5Screate_simple", "H5Screate", "H5Acreate2", "H5Dcreate2"):
        getattr(L, fn).restype = hid
    T = {k: hid.in_dll(L, k + "_g").value for k in ("H5T_NATIVE_INT", "H5T_NATIVE_UINT", "H5T_NATIVE_DOUBLE", "H5T_STD_I32LE",
                                                    "H5T_STD_U32LE", "H5T_IEEE_F64LE")}
    f = L.H5Fcreate(os.fsencode(path), C.c_uint(2), hid(0), hid(0))
    assert f >= 0

    def attr(loc, name, ftype, mtype, arr):
        arr = np.ascontiguousarray(arr)
        if arr.ndim == 0:
            sp = L.H5Screate(C.c_int(0))
        else:
            dims = (C.c_uint64 * 1)(arr.size)
            sp = L.H5Screate_simple(C.c_int(1), dims, None)
        a = L.H5Acreate2(hid(loc), name.encode(), hid(T[ftype]), hid(sp), hid(0), hid(0))
        assert L.H5Awrite(hid(a), hid(T[mtype]), arr.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Aclose(hid(a)); L.H5Sclose(hid(sp))

    def dset(loc, name, ftype, mtype, arr):
        arr = np.ascontiguousarray(arr)
        dims = (C.c_uint64 * arr.ndim)(*arr.shape)
        sp = L.H5Screate_simple(C.c_int(arr.ndim), dims, None)
        d = L.H5Dcreate2(hid(loc), name.encode(), hid(T[ftype]), hid(sp), hid(0), hid(0), hid(0))
        assert L.H5Dwrite(hid(d), hid(T[mtype]), hid(0), hid(0), hid(0), arr.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Dclose(hid(d)); L.H5Sclose(hid(sp))

    n = len(mass)
    h = L.H5Gcreate2(hid(f), b"Header", hid(0), hid(0), hid(0))
    attr(h, "NumPart_ThisFile", "H5T_STD_I32LE", "H5T_NATIVE_INT", np.array([n, 0, 0, 0, 0, 0], np.int32))
    attr(h, "NumPart_Total", "H5T_STD_U32LE", "H5T_NATIVE_UINT", np.array([n, 0, 0, 0, 0, 0], np.uint32))
    attr(h, "NumPart_Total_HighWord", "H5T_STD_U32LE", "H5T_NATIVE_UINT", np.zeros(6, np.uint32))
    attr(h, "MassTable", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", np.zeros(6))
    attr(h, "Time", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", np.array(0.0))
    attr(h, "Redshift", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", np.array(0.0))
    attr(h, "BoxSize", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", np.array(box))
    attr(h, "NumFilesPerSnapshot", "H5T_STD_I32LE", "H5T_NATIVE_INT", np.array(1, np.int32))
    attr(h, "Flag_DoublePrecision", "H5T_STD_I32LE", "H5T_NATIVE_INT", np.array(1, np.int32))
    L.H5Gclose(hid(h))
    g = L.H5Gcreate2(hid(f), b"PartType0", hid(0), hid(0), hid(0))
    dset(g, "Coordinates", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", pos)
    dset(g, "Velocities", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", vel)
    dset(g, "Masses", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", mass)
    dset(g, "InternalEnergy", "H5T_IEEE_F64LE", "H5T_NATIVE_DOUBLE", u)
    dset(g, "ParticleIDs", "H5T_STD_U32LE", "H5T_NATIVE_UINT", np.arange(1, n + 1, dtype=np.uint32))
    L.H5Gclose(hid(g))
    L.H5Fclose(hid(f))


@pytest.mark.skipif(not os.path.exists(LIBHDF5), reason="HDF5 C library absent")
def test_reference_style_ic(tmp_path):
    n = 40
    p = str(tmp_path / "ic.hdf5")
    f = fields(n, seed=5)
    write_reference_style_ic(p, f["pos"], f["vel"], f["mass"], f["u"], 3.0)
    r = io.read_ic(p)
    assert "h" not in r and "rho" not in r
    for k in ("pos", "vel", "mass", "u"):
        assert np.array_equal(r[k], f[k]), k
    assert np.array_equal(r["ids"], np.arange(1, n + 1))
    assert r["header"]["box_size"] == 3.0 and r["header"]["num_gas"] == n and r["header"]["double_precision"] == 1


def test_errors(tmp_path):
    p = str(tmp_path / "s.hdf5")
    io.write_snapshot(p, fields(10))
    from gizmo_amd.hydro import Fields
    import ctypes as C
    f = Fields()
    pos = np.zeros((11, 3))
    f.pos = pos.ctypes.data_as(C.POINTER(C.c_double))
    rc = io.iolib().gz_ic_read(os.fsencode(p), 11, C.byref(f), None, None)
    assert rc != 0 and "10 gas particles" in io.iolib().gz_io_last_error().decode()
    with pytest.raises(io.GizmoIOError, match="cannot open"):
        io.read_header(str(tmp_path / "missing.hdf5"))
