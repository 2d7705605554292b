"""Per-particle error measures of the parity tests.

err(field, a, b, ref) is the largest over particles of |a_i - b_i| / max(|b_i|, s_i): the error of
each particle relative to its own value, with a floor s_i at that particle's physical scale of
the field, so a quantity that cancels to nearly zero (the acceleration of a particle in a
uniform medium is a sum of pressure terms that cancel) is measured against the size of its
terms, not against zero.  Vectors and tensors count as one value per particle (their norm).
The scales come from the reference state (the oracle's): c_s = sqrt(gamma P / rho) and
  * h, rho, numngb, press, maxsig, cn: no floor (positive);
  * drkern, fce (dimensionless, O(1)): 1;
  * divvel: (c_s + |v|) / h;  grad_rho: rho / h;  grad_p: P / h;  grad_v: (c_s + |v|) / h;
  * nvt (the inverted second-moment matrix, ~ 1/h^2): |nvt_i|;  nvt_fw (~ h^-2): 1 / h^2;
  * accel: P / (rho h);  dtu: P c_s / (rho h);
  * vel, vel_kick: c_s + |v|;  u, u_kick: u;  soft: h;  pos: periodic differences over the box.
"""
import numpy as np


def _norm(x):
    x = np.asarray(x, dtype=np.float64)
    return np.abs(x) if x.ndim == 1 else np.sqrt((x.reshape(len(x), -1) ** 2).sum(1))


def scales(ref, gamma):
    """Per-particle floor of each field from a reference state (attributes or dict)."""
    g = (lambda k: ref[k]) if isinstance(ref, dict) else (lambda k: getattr(ref, k))
    rho = np.asarray(g("rho"), dtype=np.float64)
    p = np.asarray(g("press"), dtype=np.float64)
    h = np.asarray(g("h"), dtype=np.float64)
    v = _norm(g("vel"))
    safe_rho = np.where(rho > 0, rho, 1.0)
    cs = np.sqrt(np.maximum(gamma * p / safe_rho, 0.0))
    sv = cs + v
    s = dict(h=0.0, rho=0.0, numngb=0.0, press=0.0, maxsig=0.0, cn=0.0, drkern=1.0, fce=1.0,
             divvel=sv / h, grad_rho=rho / h, grad_p=p / h, grad_v=sv / h, nvt_fw=1.0 / h ** 2,
             accel=p / (safe_rho * h), dtu=p * cs / (safe_rho * h), vel=sv, vel_kick=sv, u=np.asarray(g("u"), dtype=np.float64),
             u_kick=np.asarray(g("u"), dtype=np.float64), soft=h)
    try:
        s["nvt"] = _norm(g("nvt"))
    except (KeyError, AttributeError):
        pass
    return s


def err(field, a, b, sc):
    """Largest per-particle relative error of field a against reference b (floors sc[field])."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = _norm(a - b)
    den = np.maximum(_norm(b), sc.get(field, 0.0) if isinstance(sc, dict) else sc)
    den = np.where(den > 0, den, 1e-300)
    return float((d / den).max()) if len(d) else 0.0


def errs(fields, out, ref, sc):
    g = (lambda k: ref[k]) if isinstance(ref, dict) else (lambda k: getattr(ref, k))
    return {k: err(k, out[k], g(k), sc) for k in fields}


def periodic_err(a, b, box):
    d = np.asarray(a) - np.asarray(b)
    return float(np.abs(d - box * np.round(d / box)).max())


def conservation(out, box_active=None):
    """Relative momentum and energy residuals of a force pass over every particle of a periodic
    box: |sum m a| / sum m |a| and |sum m (du/dt + v.a)| / sum m (|du/dt| + |v.a|)."""
    m = np.asarray(out["mass"], dtype=np.float64)
    acc = np.asarray(out["accel"], dtype=np.float64)
    du = np.asarray(out["dtu"], dtype=np.float64)
    v = np.asarray(out["vel"], dtype=np.float64)
    mom = np.abs((m[:, None] * acc).sum(0)) / np.maximum((m[:, None] * np.abs(acc)).sum(0), 1e-300)
    va = (v * acc).sum(1)
    e = abs((m * (du + va)).sum()) / max((m * (np.abs(du) + np.abs(va))).sum(), 1e-300)
    return float(mom.max()), float(e)
