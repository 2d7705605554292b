"""A GIZMO HDF5 IC through the device hot path and back out as a snapshot (include/gizmo_io.h
with include/gizmo_hydro.h): the snapshot holds exactly what the device computed."""
import numpy as np
import pytest

from gizmo_amd import ics, io
from gizmo_amd.hydro import GZ_FP64, Hydro

pytestmark = pytest.mark.gpu


def test_ic_pass_snapshot(tmp_path):
    ic = ics.sedov_box(12, 2001, kernel=6, des_num_ngb=64.0)
    par = dict(ic["params"])
    p_ic = str(tmp_path / "ic.hdf5")
    io.write_snapshot(p_ic, dict(pos=ic["pos"], vel=ic["vel"], mass=ic["mass"], u=ic["u"], h=ic["h"]), box_size=par["box"])
    r = io.read_ic(p_ic)
    n = r["header"]["num_gas"]
    print(f"ic -> pass -> snapshot: n={n}", flush=True)
    g = Hydro(par, n, GZ_FP64)
    g.load(dict(pos=r["pos"], vel=r["vel"], mass=r["mass"], u=r["u"], h=r["h"]))
    g.hydro_pass()
    names = ("pos", "vel", "mass", "u", "rho", "h", "press", "accel", "dtu", "divvel", "grad_rho", "grad_v")
    out = g.store(names)
    g.close()
    p_snap = str(tmp_path / "snapshot_000.hdf5")
    io.write_snapshot(p_snap, out, box_size=par["box"], ids=r["ids"])
    s = io.read_ic(p_snap)
    for k in ("pos", "vel", "mass", "u", "rho", "h"):
        assert np.array_equal(s[k], np.asarray(out[k]).reshape(s[k].shape)), k
    assert np.array_equal(s["ids"], r["ids"])
    assert np.isfinite(out["accel"]).all() and (out["rho"] > 0).all()
