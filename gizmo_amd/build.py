"""Builds the HIP library gizmo_amd/libgizmo_hip.so for gfx950 (in-tree).

    python -m gizmo_amd.build

Flags:
  * --offload-arch=gfx950: MI355X only.
  * -ffp-contract=off for the FP64 kernels and the host code: the FP64 form keeps the
    reference's rounding order term by term, so it can be compared to the CPU oracle bit for
    bit.  The FP32-mixed kernels (gz_kernels_f32.hip) contract to FMA within each expression
    (so the kernel shapes round alike and agree bit for bit) and use the fast FP32 divide and
    square root.
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 instructions.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgizmo_hip.so")
SOURCES = ["gz_kernels.hip", "gz_kernels_f32.hip", "gz_api.hip", "gz_batch.hip"]
HEADERS = ["gz_device.h", "gz_kernels.h", "gz_groups.h", "gz_kernels_impl.h", "gz_rows_impl.h", "gz_time_impl.h", "gz_batch_impl.h",
           os.path.join("..", "..", "include", "gizmo_hydro.h"), os.path.join("..", "..", "include", "gizmo_batch.h")]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
# per source: the FP64 form and the host code keep strict rounding; the FP32-mixed kernels
# contract to FMA within each expression (not across statements: the kernel shapes, which
# spread one pair's terms over different statements, then round them alike and give the same
# bits) and use the fast FP32 divide/square root
SOURCE_FLAGS = {
    "gz_kernels.hip": ["-ffp-contract=off"],
    "gz_api.hip": ["-ffp-contract=off"],
    "gz_batch.hip": ["-ffp-contract=off"],
    "gz_kernels_f32.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
}


def build(verbose=False, jobs=None):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [hipcc()] + FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    bad = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            bad.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if bad:
        msg = "\n".join(f"--- {s} ---\n{o}" for s, o in bad)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_io(verbose)
    return LIB


# HDF5 ICs and snapshots (include/gizmo_io.h): host code against the HDF5 C library of the
# image (/opt/conda), its own library so the hot path does not depend on HDF5
IO_LIB = os.path.join(HERE, "libgizmo_io.so")
HDF5_ROOT = os.environ.get("GZ_HDF5_ROOT", "/opt/conda")


def build_io(verbose=False):
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        if verbose:
            print(f"no HDF5 under {HDF5_ROOT}: libgizmo_io.so not built", flush=True)
        return None
    tmp = IO_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", os.path.join(CSRC, "gz_io.cpp"), "-I" + inc, "-L" + lib,
           "-Wl,-rpath," + lib, "-lhdf5", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, IO_LIB)
    return IO_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
