"""Per-particle input features of the warm-start network.

Raw features (16 per particle), all dimensionless:
  0-2   position: (x - cx) / R, (y - cy) / R, (z - z_min) / H  (R = max lateral half-width)
  3     log10(E / 1000 MPa)
  4     1 for the cortical phase, else 0
  5-7   fixed-DOF mask bits (x, y, z)
  8     1 on the loaded particles
  9-11  unit load direction (the specimen's resultant external force)
  12    log10(resultant pressure / 1 MPa)
  13    H / 25 mm,  14  R / 20 mm,  15  1
Two message-passing hops over the bonds (weighted neighbour means, libpdx
pdx_neighbor_mean) of the raw vector give 16 + 16 more; 16 zeros pad the
input to 64 columns (the GEMM K granularity).
"""
from __future__ import annotations

import numpy as np

FEAT = 64
RAW = 16


def _per_case(spec, off):
    if off is None:
        return [(0, spec.n)]
    return [(int(off[k]), int(off[k + 1])) for k in range(len(off) - 1)]


def raw_features(spec, case_offsets=None) -> np.ndarray:
    X, n = spec.X, spec.n
    F = np.zeros((n, RAW))
    for a, b in _per_case(spec, case_offsets):
        Xc = X[a:b]
        lo, hi = Xc.min(0), Xc.max(0)
        c = 0.5 * (lo + hi)
        R = max(0.5 * (hi[0] - lo[0]), 0.5 * (hi[1] - lo[1]), 1e-9)
        H = max(hi[2] - lo[2], 1e-9)
        F[a:b, 0] = (Xc[:, 0] - c[0]) / R
        F[a:b, 1] = (Xc[:, 1] - c[1]) / R
        F[a:b, 2] = (Xc[:, 2] - lo[2]) / H
        fe = spec.fext[a:b]
        tot = fe.sum(0)
        loaded = np.abs(fe).sum(1) > 0
        nrm = np.linalg.norm(tot)
        F[a:b, 8] = loaded
        if nrm > 0:
            F[a:b, 9:12] = tot / nrm
            # resultant pressure: |sum f| * V / loaded area (area = n_loaded dx^2)
            p = nrm * spec.dx ** 3 / max(loaded.sum() * spec.dx ** 2, 1e-30)
            F[a:b, 12] = np.log10(max(p, 1e-12))
        F[a:b, 13] = H / 25.0
        F[a:b, 14] = R / 20.0
    F[:, 3] = np.log10(spec.E / 1000.0)
    F[:, 4] = (spec.material_id == 0) if spec.material_id is not None else 0.0
    for q in range(3):
        F[:, 5 + q] = (spec.fixed >> q) & 1
    F[:, 15] = 1.0
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
