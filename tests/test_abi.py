"""The C-ABI library loads and exports every entry point include/*.h declares (no GPU calls)."""
import ctypes
import glob
import os
import re

from gizmo_amd import hydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(headers=None):
    names = set()
    for h in headers or glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(gz_[a-z_0-9]+)\s*\(", src):
            names.add(m.group(1))
    return names


IO_HEADER = os.path.join(ROOT, "include", "gizmo_io.h")
# header-only C++ templates of the drop-in shim: compiled into the caller, not exported
DROPIN_HEADER = os.path.join(ROOT, "include", "gizmo_dropin.h")


def test_library_exports_declared_symbols():
    assert os.path.exists(hydro.LIB_PATH), "build with python -m gizmo_amd.build"
    L = ctypes.CDLL(hydro.LIB_PATH)
    names = declared_symbols([h for h in glob.glob(os.path.join(ROOT, "include", "*.h")) if h not in (IO_HEADER, DROPIN_HEADER)])
    assert {"gz_density", "gz_hydro_gradient_calc", "gz_hydro_force", "gz_load", "gz_store"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_matches_header_structs():
    # the ctypes mirrors follow the header's field order
    src = open(os.path.join(ROOT, "include", "gizmo_hydro.h")).read()
    body = src[src.index("typedef struct {\n    double *pos;"):]
    body = body[:body.index("} gz_fields;")]
    fields = re.findall(r"\*(\w+);", body)
    assert fields == [f for f, _ in hydro.Fields._fields_]
    body = src[src.index("typedef struct {\n    double box_size;"):]
    body = body[:body.index("} gz_params;")]
    params = re.findall(r"\b(?:double|int)\s+(\w+);", body)
    assert params == [f for f, _ in hydro.Params._fields_]


def test_io_library_exports_declared_symbols():
    from gizmo_amd import io
    L = io.iolib()
    names = declared_symbols([IO_HEADER])
    assert {"gz_ic_header", "gz_ic_read", "gz_snapshot_write", "gz_io_last_error"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_io_binding_matches_header_struct():
    from gizmo_amd import io
    src = open(IO_HEADER).read()
    body = src[src.index("typedef struct {"):]
    body = body[:body.index("} gz_snap_header;")]
    fields = re.findall(r"\b(?:long long|double|int)\s+(\w+);", body)
    assert fields == [f for f, _ in io.SnapHeader._fields_]
