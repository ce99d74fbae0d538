"""Warm-started dynamic relaxation: features of the specimen through the
bonds of its libpdx context, one inference pass, then DR from the predicted
field under the same criterion as the cold start (residual relative to the
cold-start residual).  The prescribed DOFs keep their prescribed values."""
from __future__ import annotations

import time

import numpy as np

from .features import feature_tensor, hops_of, target_scale_tensor


def predict_u_device(net, body, spec, case_offsets=None):
    """The predicted displacement as a (n, 3) float64 CUDA tensor (features,
    inference and scaling on the GPU); prescribed DOFs keep spec.u0.  The
    network's input width gives its neighbour-mean hop count."""
    import torch
    y = net.predict_device(feature_tensor(body, spec, case_offsets, hops_of(net.dims[0])))
    u = y.to(torch.float64) * target_scale_tensor(spec, case_offsets)[:, None]
    fixed = torch.from_numpy(np.ascontiguousarray(spec.fixed, np.uint8)).cuda().to(torch.int32)
    u0 = torch.from_numpy(np.ascontiguousarray(spec.u0, np.float64)).cuda()
    for q in range(3):
        m = ((fixed >> q) & 1).bool()
        u[:, q] = torch.where(m, u0[:, q], u[:, q])
    return u


def predict_u(net, body, spec, case_offsets=None) -> np.ndarray:
    return predict_u_device(net, body, spec, case_offsets).cpu().numpy()


def warm_solve(body, spec, net, tol=1e-6, ref=None, case_offsets=None, max_iterations=200000, **kw):
    """Returns (DR result(s), timings dict).  ref: the cold-start residual (per
    case for a batch); computed here when None."""
    import torch
    if ref is None:
        ref = body.cold_residual()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    body.set_u_device(predict_u_device(net, body, spec, case_offsets))
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    refs = np.atleast_1d(np.asarray(ref, np.float64))
    r = body.solve("adaptive", tol=tol, max_iterations=max_iterations, refs=refs, **kw)
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    return r, {"inference_s": t_inf, "total_s": t_all}
