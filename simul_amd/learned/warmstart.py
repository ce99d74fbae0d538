"""Warm-started dynamic relaxation: features of the specimen through the
bonds of its libpdx context, one inference pass, then DR from the predicted
field under the same criterion as the cold start (residual relative to the
cold-start residual).  The prescribed DOFs keep their prescribed values."""
from __future__ import annotations

import time

import numpy as np

from .features import feature_tensor, target_scale


def predict_u(net, body, spec, case_offsets=None) -> np.ndarray:
    y = net.predict(feature_tensor(body, spec, case_offsets))
    u = y.astype(np.float64) * target_scale(spec, case_offsets)[:, None]
    fixed = spec.fixed
    for q in range(3):
        m = ((fixed >> q) & 1).astype(bool)
        u[m, q] = spec.u0[m, q]
    return u


def warm_solve(body, spec, net, tol=1e-6, ref=None, case_offsets=None, max_iterations=200000, **kw):
    """Returns (DR result(s), timings dict).  ref: the cold-start residual (per
    case for a batch); computed here when None."""
    import torch
    if ref is None:
        ref = body.cold_residual()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = predict_u(net, body, spec, case_offsets)
    body.u = u
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    refs = np.atleast_1d(np.asarray(ref, np.float64))
    r = body.solve("adaptive", tol=tol, max_iterations=max_iterations, refs=refs, **kw)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    return r, {"inference_s": t_inf, "total_s": t_all}
