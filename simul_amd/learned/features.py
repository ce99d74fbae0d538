"""Per-particle input features of the warm-start network.

Raw features (20 per particle), all dimensionless:
  0-2   position: (x - cx) / a, (y - cy) / b, (z - z_min) / H  (a, b lateral half-widths)
  3     log10(E / 1000 MPa)
  4     1 for the cortical phase, else 0
  5-7   fixed-DOF mask bits (x, y, z)
  8     1 on the loaded particles
  9-11  unit load direction (the specimen's resultant external force)
  12    log10(resultant pressure / 1 MPa)
  13-15 H / 25 mm, a / 20 mm, b / 20 mm
  16    specimen mean of log10(E / 1000 MPa)
  17    specimen cortical fraction
  18    1,   19  0
Two message-passing hops over the bonds (weighted neighbour means, libpdx
pdx_neighbor_mean) of the raw vector give 20 + 20 more; 4 zeros pad the
input to 64 columns (the GEMM K granularity).
"""
from __future__ import annotations

import numpy as np

FEAT = 64
RAW = 20


def _per_case(spec, off):
    if off is None:
        return [(0, spec.n)]
    return [(int(off[k]), int(off[k + 1])) for k in range(len(off) - 1)]


def raw_features(spec, case_offsets=None) -> np.ndarray:
    X, n = spec.X, spec.n
    F = np.zeros((n, RAW))
    logE = np.log10(spec.E / 1000.0)
    cort = (spec.material_id == 0).astype(np.float64) if spec.material_id is not None else np.zeros(n)
    for a0, b0 in _per_case(spec, case_offsets):
        Xc = X[a0:b0]
        lo, hi = Xc.min(0), Xc.max(0)
        c = 0.5 * (lo + hi)
        ax = max(0.5 * (hi[0] - lo[0]), 1e-9)
        by = max(0.5 * (hi[1] - lo[1]), 1e-9)
        H = max(hi[2] - lo[2], 1e-9)
        F[a0:b0, 0] = (Xc[:, 0] - c[0]) / ax
        F[a0:b0, 1] = (Xc[:, 1] - c[1]) / by
        F[a0:b0, 2] = (Xc[:, 2] - lo[2]) / H
        fe = spec.fext[a0:b0]
        tot = fe.sum(0)
        loaded = np.abs(fe).sum(1) > 0
        nrm = np.linalg.norm(tot)
        F[a0:b0, 8] = loaded
        if nrm > 0:
            F[a0:b0, 9:12] = tot / nrm
            # resultant pressure: |sum f| V / loaded area (area = n_loaded dx^2)
            p = nrm * spec.dx ** 3 / max(loaded.sum() * spec.dx ** 2, 1e-30)
            F[a0:b0, 12] = np.log10(max(p, 1e-12))
        F[a0:b0, 13] = H / 25.0
        F[a0:b0, 14] = ax / 20.0
        F[a0:b0, 15] = by / 20.0
        F[a0:b0, 16] = logE[a0:b0].mean()
        F[a0:b0, 17] = cort[a0:b0].mean()
    F[:, 3] = logE
    F[:, 4] = cort
    for q in range(3):
        F[:, 5 + q] = (spec.fixed >> q) & 1
    F[:, 18] = 1.0
    return F


def feature_matrix(body, spec, case_offsets=None) -> np.ndarray:
    """(n, 64) float32: raw features and two neighbour-mean hops through the
    bonds of ``body`` (a libpdx context of the same particles)."""
    raw = raw_features(spec, case_offsets)
    m1 = body.neighbor_mean(raw)
    m2 = body.neighbor_mean(m1)
    out = np.zeros((spec.n, FEAT), np.float32)
    out[:, :RAW] = raw
    out[:, RAW:2 * RAW] = m1
    out[:, 2 * RAW:3 * RAW] = m2
    return out


def target_scale(spec, case_offsets=None) -> np.ndarray:
    """Per-particle displacement scale s = p H / 1000 MPa of its case: the
    network predicts u / s."""
    s = np.zeros(spec.n)
    for a, b in _per_case(spec, case_offsets):
        Xc = spec.X[a:b]
        H = np.ptp(Xc[:, 2])
        fe = spec.fext[a:b]
        loaded = np.abs(fe).sum(1) > 0
        nrm = np.linalg.norm(fe.sum(0))
        p = nrm * spec.dx ** 3 / max(loaded.sum() * spec.dx ** 2, 1e-30)
        s[a:b] = max(p, 1e-12) * H / 1000.0
    return s
