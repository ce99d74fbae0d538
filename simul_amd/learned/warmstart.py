"""Warm-started dynamic relaxation: features of the specimen through the
bonds of its libpdx context, one inference pass, then DR from the predicted
field under the same criterion as the cold start (residual relative to the
cold-start residual).  The prescribed DOFs keep their prescribed values."""
from __future__ import annotations

import time

import numpy as np

from .features import features_and_scale, hops_of


def predict_u_device(net, body, spec, case_offsets=None):
    """The predicted displacement as a (n, 3) float64 CUDA tensor (features,
    inference and scaling on the GPU); prescribed DOFs keep spec.u0.  The
    network's input width gives its neighbour-mean hop count; a case loaded
    through prescribed displacements is seen through features.load_view."""
    import torch
    F, s = features_and_scale(body, spec, case_offsets, hops_of(net.dims[0]))
    u = net.predict_device(F).to(torch.float64) * s[:, None]
    del F
    rows = np.flatnonzero(spec.fixed != 0)
    if len(rows):
        r = torch.from_numpy(rows).cuda()
        bits = torch.from_numpy(np.ascontiguousarray(spec.fixed[rows])).cuda().to(torch.int32)
        u0 = torch.from_numpy(np.ascontiguousarray(spec.u0[rows], np.float64)).cuda()
        sub = u[r]
        for q in range(3):
            sub[:, q] = torch.where(((bits >> q) & 1).bool(), u0[:, q], sub[:, q])
        u[r] = sub
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


def loading_start(net, body, spec, indices, u_final, n_increments):
    """The ``start`` of loading.IncrementalLoading from a network trained on
    prescribed-displacement cases: increment k's first guess is the network's
    prediction for the specimen with its prescribed rows ``indices`` at
    (k / n_increments) x u_final (the other prescribed DOFs at spec.u0)."""
    import dataclasses
    idx = np.asarray(indices, np.int64)
    uf = np.broadcast_to(np.asarray(u_final, np.float64), (len(idx), spec.X.shape[1]))

    def start(k):
        u0 = spec.u0.copy()
        u0[idx] = (k / n_increments) * uf
        return predict_u_device(net, body, dataclasses.replace(spec, u0=u0))
    return start
