"""The first-derivative volume correction of the density (option "fd_correction", the
reference's HYDRO_PARTITION_UNITY_IMPROVE_FD: src/hydro/density_loop.h:1281-1286 sums,
src/hydro/density_loop.cc:622-633 correction) on the device against the oracle.

Bars: at a fixed radius the device's density and pressure within 1e-15 of the oracle's per
particle in the FP64 form (the four extra sums run in the oracle's order; pow(h, 3) may differ
in its last bit between the device's and the host's libm); a whole pass (radius iteration,
gradients, force) within 1e-12 per particle (tests/parity.py) in FP64, through both kernel
shapes; the FP32-mixed form within 1e-5 per particle at the oracle's radii; a box of a batch
with the option on within 1e-12.  Workloads: particles whose number density grows as
exp(2x) (the correction is of order 1e-3 there) and the 24^3 Sedov blast.
"""
import numpy as np
import pytest

import parity
from gizmo_amd import ics
from gizmo_amd.batch import Batch
from gizmo_amd.hydro import GZ_FP32_MIXED, GZ_FP64, Hydro

pytestmark = pytest.mark.gpu

PATHS = {"lanes": 0, "waves": 2_000_000_000}
DENSITY_FIELDS = ("h", "numngb", "drkern", "divvel", "rho", "press", "nvt", "nvt_fw", "cn", "fce")
PASS_FIELDS = DENSITY_FIELDS + ("grad_rho", "grad_p", "grad_v", "accel", "dtu", "maxsig")


def workloads():
    return {"gradient24": ics.gradient_box(24, k=2.0, kernel=6, des_num_ngb=64.0),
            "sedov24": ics.sedov_box(24, 1002, kernel=6, des_num_ngb=64.0)}


@pytest.fixture(scope="module")
def fd_cases(oracle_lib):
    O = oracle_lib
    out = {}
    for name, ic in workloads().items():
        par = dict(ic["params"], fd_correction=1)
        s0 = O.State.from_ic(ic)
        s3 = s0.copy()
        O.density(par, s3)
        s1 = s3.copy()
        O.gradients(par, s3)
        O.force(par, s3)
        out[name] = dict(ic=ic, par=par, s0=s0, s1=s1, s3=s3)
    return out


def device(state, par, precision=GZ_FP64, path="lanes", fd=1):
    g = Hydro({k: v for k, v in par.items() if k != "fd_correction"}, state.n, precision)
    g.set_option("few_rows", PATHS[path])
    g.set_option("fd_correction", fd)
    g.load(dict(pos=state.pos, vel=state.vel, mass=state.mass, u=state.u, h=state.h))
    return g


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("name", ["gradient24", "sedov24"])
def test_fd_density_fixed_h(fd_cases, oracle_lib, name, path):
    c = fd_cases[name]
    par = dict(c["par"], time=1.0, time_begin=0.0, max_num_ngb_dev=1e30)
    s1 = c["s1"]
    print(f"fd fixed h: {name} {path} n={s1.n}", flush=True)
    g = device(s1, par, path=path)
    perm = g.permutation()
    g.density()
    assert g.stat("density_iterations") == 1
    out = g.store(DENSITY_FIELDS)
    g.close()
    so = s1.permuted(perm)
    so.h[:] = s1.h[perm]
    oracle_lib.density_fixed_h(par, so)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    for k in ("nvt_fw", "drkern", "divvel"):
        assert np.array_equal(out[k], getattr(so, k)[inv]), k
    # the density takes pow(h, 3) (the device's and the host's libm may differ in the last bit)
    for k in ("rho", "press"):
        a, b = out[k], getattr(so, k)[inv]
        assert float(np.max(np.abs(a - b) / np.abs(b))) <= 1e-15, k
    # the correction took effect
    off = s1.permuted(perm)
    off.h[:] = s1.h[perm]
    oracle_lib.density_fixed_h(dict(par, fd_correction=0), off)
    assert (out["rho"] < off.rho[inv]).any()


@pytest.mark.parametrize("path", list(PATHS))
@pytest.mark.parametrize("name", ["gradient24", "sedov24"])
def test_fd_pass_fp64(fd_cases, name, path):
    c = fd_cases[name]
    print(f"fd pass fp64: {name} {path}", flush=True)
    g = device(c["s0"], c["par"], path=path)
    g.hydro_pass()
    out = g.store(PASS_FIELDS)
    g.close()
    ref = c["s3"]
    errs = parity.errs(PASS_FIELDS, out, ref, parity.scales(ref, c["par"]["gamma"]))
    print("per-particle:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= 1e-12, (k, e)


@pytest.mark.parametrize("name", ["gradient24", "sedov24"])
def test_fd_fp32_mixed(fd_cases, oracle_lib, name):
    """FP32-mixed at the oracle's converged radii (a band wide enough that the first evaluation
    stands): per-particle errors within 1e-5 (du/dt 2e-5 on the blast, as without the option)."""
    O = oracle_lib
    c = fd_cases[name]
    par = dict(c["par"], max_num_ngb_dev=1e30)
    ref = c["s1"].copy()
    O.density(par, ref)
    O.gradients(par, ref)
    O.force(par, ref)
    g = device(c["s1"], par, precision=GZ_FP32_MIXED)
    g.hydro_pass()
    out = g.store(PASS_FIELDS)
    g.close()
    errs = parity.errs(PASS_FIELDS, out, ref, parity.scales(ref, par["gamma"]))
    print(f"fd fp32 {name}:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= (2e-5 if k == "dtu" else 1e-5), (k, e)


def test_fd_box_of_a_batch(fd_cases):
    """A batch with the option on: the box's first pass (every particle active) against the
    oracle's, at the same 1e-12 per-particle bar."""
    c = fd_cases["gradient24"]
    other = ics.lattice_box(14, 2002, kernel=6, des_num_ngb=64.0, mach_sigma=0.3)
    par = {k: v for k, v in c["par"].items() if k != "fd_correction"}
    integ = dict(courant_fac=0.05, err_tol_int_accuracy=0.002, max_size_timestep=0.01, min_size_timestep=0.0,
                 time_begin=0.0, time_max=1.0, min_egy_spec=0.0)
    s0 = c["s0"]
    b = Batch(par, s0.n + len(other["mass"]), 2, GZ_FP64)
    b.set_option("fd_correction", 1)
    b.add(dict(pos=other["pos"], vel=other["vel"], mass=other["mass"], u=other["u"], h=other["h"]), other["params"], integ)
    bid = b.add(dict(pos=s0.pos, vel=s0.vel, mass=s0.mass, u=s0.u, h=s0.h), par, integ)
    print(f"fd batch: n={s0.n}", flush=True)
    b.begin()
    out = b.store(bid, PASS_FIELDS)
    b.close()
    ref = c["s3"]
    errs = parity.errs(PASS_FIELDS, out, ref, parity.scales(ref, par["gamma"]))
    print("per-particle:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= 1e-12, (k, e)
