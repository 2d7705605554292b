"""Writes tests/golden/ref_kernel_mat3.npz: input/output vectors of the reference's own
header-only kernel (src/mesh/kernel.h, cubic spline and Wendland C2) and 3x3 inverse
(src/math_types/mat3.h), from the libraries `make -C oracle ref` compiles from the reference
source tree where it lies (oracle/_ref/).  The fixture holds data only (inputs and outputs),
so the oracle stays pinned against the reference on machines without its source tree.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REFDIR = os.path.join(ROOT, "oracle", "_ref")
OUT = os.path.join(HERE, "ref_kernel_mat3.npz")


def load(kern):
    L = C.CDLL(os.path.join(REFDIR, f"libref_k{kern}.so"))
    dp = C.POINTER(C.c_double)
    L.ref_kernel.argtypes = [C.c_double, C.c_double, dp, dp]
    L.ref_mat3_invert.argtypes = [dp, dp]
    L.ref_mat3_invert.restype = C.c_double
    L.ref_mat3_frobenius_norm_sq.argtypes = [dp]
    L.ref_mat3_frobenius_norm_sq.restype = C.c_double
    return L


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for kern in (3, 6):
        L = load(kern)
        hs, rs, ws, dws = [], [], [], []
        for h in (0.0123, 0.25, 1.0, 3.7):
            r = np.concatenate([rng.uniform(0, h, 250), [0.0, 0.25 * h, 0.5 * h, 0.75 * h, h * (1 - 1e-15), h]])
            for x in r:
                w, d = C.c_double(), C.c_double()
                L.ref_kernel(float(x), h, C.byref(w), C.byref(d))
                hs.append(h); rs.append(x); ws.append(w.value); dws.append(d.value)
        out[f"k{kern}_h"] = np.array(hs)
        out[f"k{kern}_r"] = np.array(rs)
        out[f"k{kern}_w"] = np.array(ws)
        out[f"k{kern}_dw"] = np.array(dws)
    L = load(6)
    dp = C.POINTER(C.c_double)
    T, Ti, fro = [], [], []
    for k in range(300):
        A = rng.normal(size=(3, 3)) * (10.0 ** rng.uniform(-3, 3))
        M = A @ A.T + (0.1 if k % 3 else 1e-9) * np.eye(3)
        m = np.ascontiguousarray(M.reshape(9))
        mi = np.zeros(9)
        L.ref_mat3_invert(m.ctypes.data_as(dp), mi.ctypes.data_as(dp))
        T.append(m)
        Ti.append(mi)
        fro.append((L.ref_mat3_frobenius_norm_sq(m.ctypes.data_as(dp)), L.ref_mat3_frobenius_norm_sq(mi.ctypes.data_as(dp))))
    out["mat_T"] = np.array(T)
    out["mat_Tinv"] = np.array(Ti)
    out["mat_fro"] = np.array(fro)
    np.savez_compressed(OUT, **out)
    print(OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
