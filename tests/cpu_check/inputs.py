"""Writes the binary input of pdx_cpu_check (see the header of pdx_cpu_check.cpp)."""
from __future__ import annotations

import numpy as np

MODELS = {"linear_elastic": 0, "neo_hookean": 1, "j2_plasticity": 2}


def write_input(path, X, vol, mass, bulk, shear, cst, fixed, fext=None, u0=None, *, horizon, G_s=0.1,
                model="linear_elastic", form=0, offsets=None, sy=None, hiso=None, hkin=None):
    X = np.ascontiguousarray(X, np.float64)
    n, dim = X.shape
    m = MODELS[model] if isinstance(model, str) else int(model)
    off = np.array([0, n] if offsets is None else offsets, np.int64)
    f64 = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))
    vec = lambda a: np.zeros((n, dim)) if a is None else np.ascontiguousarray(a, np.float64).reshape(n, dim)
    with open(path, "wb") as fh:
        np.array([n, dim, m, form, len(off) - 1], np.int32).tofile(fh)
        off.tofile(fh)
        np.array([horizon, G_s], np.float64).tofile(fh)
        X.tofile(fh)
        for a in (vol, mass, bulk, shear, cst):
            f64(a).tofile(fh)
        vec(fext).tofile(fh)
        vec(u0).tofile(fh)
        np.ascontiguousarray(fixed, np.uint8).tofile(fh)
        if m == 2:
            for a in (sy, hiso, 0.0 if hkin is None else hkin):
                f64(a).tofile(fh)
