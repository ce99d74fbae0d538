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
  18-19 the specimen's stiffness eccentricity: the E-weighted centroid minus
        the geometric centre, over the lateral half-widths (x, y); an axially
        loaded column bends in proportion to it
A case loaded through prescribed displacements is seen through load_view
(columns 8-12 from the mean prescribed displacement, 12 the per-mille nominal
strain).
``hops`` message-passing hops over the bonds (weighted neighbour means,
libpdx pdx_neighbor_mean; 2 by default, up to 4) of the raw vector give 20
more columns each; zeros pad the input to a multiple of 16 columns (the GEMM
K granularity): 64 columns for 2 hops, 112 for 4.  Columns 9-19 are constant
over a case, so their neighbour means are the columns themselves (every bond
joins two particles of one case): the device hops run over columns 0-9 only
and the hop blocks repeat columns 10-19.  Column 3 enters the network
only through the two hops: the particle's own modulus is i.i.d. noise in the
cancellous core, and a prediction that follows it is rough at the particle
scale.
"""
from __future__ import annotations

import numpy as np

FEAT = 64
RAW = 20
HOPPED = 10  # leading raw columns that the device hops average (the rest are case constants)


def feature_width(hops: int) -> int:
    """Input columns of the network for ``hops`` neighbour-mean hops."""
    return -(-RAW * (hops + 1) // 16) * 16


def hops_of(width: int) -> int:
    """The hop count whose feature_width is ``width``."""
    for h in range(0, 5):
        if feature_width(h) == width:
            return h
    raise ValueError(f"no hop count gives {width} input columns")


def _per_case(spec, off):
    if off is None:
        return [(0, spec.n)]
    return [(int(off[k]), int(off[k + 1])) for k in range(len(off) - 1)]


def _offsets(spec, case_offsets):
    return np.asarray([0, spec.n] if case_offsets is None else case_offsets, np.int64)


def _platens(spec, off):
    """The moving prescribed rows (fixed, nonzero u0), their cases, and each
    case's count of them and mean prescribed displacement."""
    ncase = len(off) - 1
    rows = np.flatnonzero(spec.fixed != 0)
    u0r = spec.u0[rows]
    moving = u0r[:, 0] != 0
    for q in range(1, u0r.shape[1]):
        moving |= u0r[:, q] != 0
    rows, u0r = rows[moving], u0r[moving]
    case = np.searchsorted(off, rows, side="right") - 1
    cnt = np.bincount(case, minlength=ncase)
    um = np.stack([np.bincount(case, weights=u0r[:, q], minlength=ncase) for q in range(u0r.shape[1])], 1)
    return rows, case, cnt, um / np.maximum(cnt, 1)[:, None]


def load_view(spec, case_offsets=None):
    """The specimen as the features see its load.  A case loaded only through
    prescribed displacements (platens: fixed particles with a nonzero u0, no
    external force) gets, for the features and the target scale alone, a force
    density along the mean prescribed displacement on the moving particles,
    sized so that the "pressure" is 1000 x |u| / H: the scale s = p H / 1000 MPa
    is then the prescribed |u| and column 12 the log of the per-mille nominal
    strain.  Force-loaded cases are left as they are; so is ``spec`` itself
    when no case needs the view.  (raw_view_tensor gives the raw features of
    this view without building it.)"""
    import dataclasses
    off = _offsets(spec, case_offsets)
    rows, case, cnt, um = _platens(spec, off)
    if len(rows) == 0:
        return spec
    forced = np.zeros(len(off) - 1, bool)
    forced[np.searchsorted(off, np.flatnonzero((spec.fext != 0).any(1)), side="right") - 1] = True
    load = np.zeros_like(um)
    for k in np.flatnonzero((cnt > 0) & ~forced):
        nrm = np.linalg.norm(um[k])
        if nrm > 0:
            H = max(np.ptp(spec.X[off[k]:off[k + 1], 2]), 1e-9)
            load[k] = um[k] / nrm * (1000.0 * nrm / H) / spec.dx
    keep = np.abs(load[case]).sum(1) > 0
    if not keep.any():
        return spec
    fext = spec.fext.copy()
    fext[rows[keep]] = load[case[keep]]
    return dataclasses.replace(spec, fext=fext)


def raw_features(spec, case_offsets=None) -> np.ndarray:
    X, n = spec.X, spec.n
    # column-major while filling (each column is one contiguous write), one
    # transposing copy at the end
    F = np.zeros((RAW, n))
    logE = np.log10(spec.E / 1000.0)
    cort = (spec.material_id == 0).astype(np.float64) if spec.material_id is not None else np.zeros(n)
    loaded_all = np.abs(spec.fext).sum(1) > 0
    for a0, b0 in _per_case(spec, case_offsets):
        Xc = X[a0:b0]
        lo, hi = Xc.min(0), Xc.max(0)
        c = 0.5 * (lo + hi)
        ax = max(0.5 * (hi[0] - lo[0]), 1e-9)
        by = max(0.5 * (hi[1] - lo[1]), 1e-9)
        H = max(hi[2] - lo[2], 1e-9)
        F[0, a0:b0] = (Xc[:, 0] - c[0]) / ax
        F[1, a0:b0] = (Xc[:, 1] - c[1]) / by
        F[2, a0:b0] = (Xc[:, 2] - lo[2]) / H
        tot = spec.fext[a0:b0].sum(0)
        loaded = loaded_all[a0:b0]
        nrm = np.linalg.norm(tot)
        F[8, a0:b0] = loaded
        if nrm > 0:
            F[9:12, a0:b0] = (tot / nrm)[:, None]
            # resultant pressure: |sum f| V / loaded area (area = n_loaded dx^2)
            p = nrm * spec.dx ** 3 / max(loaded.sum() * spec.dx ** 2, 1e-30)
            F[12, a0:b0] = np.log10(max(p, 1e-12))
        F[13, a0:b0] = H / 25.0
        F[14, a0:b0] = ax / 20.0
        F[15, a0:b0] = by / 20.0
        F[16, a0:b0] = logE[a0:b0].mean()
        F[17, a0:b0] = cort[a0:b0].mean()
        Ec = spec.E[a0:b0] if np.ndim(spec.E) else np.full(b0 - a0, float(spec.E))
        ce = (Ec @ Xc[:, :2]) / max(Ec.sum(), 1e-300)
        F[18, a0:b0] = (ce[0] - c[0]) / ax
        F[19, a0:b0] = (ce[1] - c[1]) / by
    F[3] = logE
    F[4] = cort
    for q in range(3):
        F[5 + q] = (spec.fixed >> q) & 1
    return np.ascontiguousarray(F.T)


def raw_features_tensor(body, spec, from_context=False):
    """raw_features computed on the GPU (libpdx pdx_warm_raw_features over the
    cases of ``body``, a libpdx context of the same particles): (n, 20)
    float64 CUDA tensor, the same values to FP64 rounding.  from_context:
    positions, external forces and DOF masks are the context's own (no copy
    of them from the host; only the moduli and phases are uploaded)."""
    import ctypes as C

    import torch
    from .. import _native as N
    n = spec.n
    E = np.ascontiguousarray(np.broadcast_to(np.asarray(spec.E, np.float64), (n,)))
    X = fext = fixed = None
    if not from_context:
        X = np.ascontiguousarray(spec.X, np.float64)
        fext = np.ascontiguousarray(spec.fext, np.float64)
        fixed = np.ascontiguousarray(spec.fixed, np.uint8)
    mat = None if spec.material_id is None else np.ascontiguousarray(spec.material_id, np.int32)
    out = torch.empty((n, RAW), dtype=torch.float64, device="cuda")
    torch.cuda.current_stream().synchronize()
    body.ctx.call("pdx_warm_raw_features", N.ptr(X, N._dp), N.ptr(E, N._dp), N.ptr(fext, N._dp),
                  N.ptr(fixed, N._u8p), N.ptr(mat, N._i32p), float(spec.dx), C.c_void_p(out.data_ptr()))
    return out


def _hop_columns(body, raw, hops):
    """(n, feature_width(hops)) float32 network input from the (n, 20) FP64
    raw tensor: [raw | hop 1 | ... | hop hops | zeros], the hops of columns
    0-9 on the device (pdx_neighbor_features_device), columns 10-19 repeated
    (case constants), column 3 zeroed."""
    import ctypes as C

    import torch
    n = raw.shape[0]
    out = torch.zeros((n, feature_width(hops)), dtype=torch.float32, device="cuda")
    out[:, :RAW] = raw
    if hops > 0:
        src = raw[:, :HOPPED].contiguous()
        hv = torch.empty((n, HOPPED * (hops + 1)), dtype=torch.float32, device="cuda")
        torch.cuda.current_stream().synchronize()
        body.ctx.call("pdx_neighbor_features_device", HOPPED, int(hops), C.c_void_p(src.data_ptr()),
                      C.c_void_p(hv.data_ptr()), HOPPED * (hops + 1))
        for k in range(1, hops + 1):
            out[:, k * RAW:k * RAW + HOPPED] = hv[:, k * HOPPED:(k + 1) * HOPPED]
            out[:, k * RAW + HOPPED:(k + 1) * RAW] = out[:, HOPPED:RAW]
    out[:, 3] = 0.0  # the local modulus only through its neighbour means
    return out


def scale_from_raw(raw):
    """The target scale s = p H / 1000 MPa of each row from its raw features
    (columns 12 and 13), p = 1e-12 MPa for an unloaded case (zero direction,
    columns 9-11): target_scale to FP64 rounding, on the device."""
    import torch
    H = 25.0 * raw[:, 13]
    loaded = (raw[:, 9:12] != 0).any(1)
    return torch.where(loaded, torch.pow(10.0, raw[:, 12]), torch.full_like(H, 1e-12)) * H / 1000.0


def raw_view_tensor(body, spec, case_offsets=None):
    """raw_features_tensor of load_view(spec) without building the view: the
    raw features of the context's positions, loads and DOF masks (those of
    ``spec``) with the specimen's moduli and phases, then, on the device, columns
    8-12 of each case that load_view would load through its platens (no force
    resultant, moving prescribed rows): 1 on the moving rows, the unit mean
    prescribed displacement and log10 of 1000 |u| / H.  Equal to the view's
    features to FP64 rounding."""
    import torch
    raw = raw_features_tensor(body, spec, from_context=True)
    off = _offsets(spec, case_offsets)
    rows, case, cnt, um = _platens(spec, off)
    if len(rows) == 0:
        return raw
    cases = np.flatnonzero(cnt > 0)  # non-empty cases with moving rows
    first = torch.from_numpy(off[cases]).cuda()
    head = raw[first][:, 9:14].cpu().numpy()  # direction, log p, H / 25 of those cases
    tab = np.zeros((len(off) - 1, 5))
    for k, h in zip(cases, head):
        nrm = np.linalg.norm(um[k])
        if np.any(h[:3] != 0) or nrm == 0:
            continue  # force-loaded (or no net displacement): its own load
        tab[k, :3] = um[k] / nrm
        tab[k, 3] = np.log10(max(1000.0 * nrm / (25.0 * h[4]), 1e-12))
        tab[k, 4] = 1.0
    if not tab[:, 4].any():
        return raw
    sizes = torch.from_numpy(np.diff(off)).cuda()
    t = torch.from_numpy(tab).cuda()[torch.repeat_interleave(torch.arange(len(off) - 1, device="cuda"), sizes)]
    m = t[:, 4] > 0
    raw[:, 9:12] = torch.where(m[:, None], t[:, :3], raw[:, 9:12])
    raw[:, 12] = torch.where(m, t[:, 3], raw[:, 12])
    moving = rows[tab[case, 4] > 0]
    raw[torch.from_numpy(moving).cuda(), 8] = 1.0
    return raw


def features_and_scale(body, spec, case_offsets=None, hops=2):
    """The network input (feature_tensor) and the per-row target scale
    (target_scale), both on the device, from one upload of the specimen
    arrays; the specimen is seen through load_view (raw_view_tensor)."""
    raw = raw_view_tensor(body, spec, case_offsets)
    return _hop_columns(body, raw, hops), scale_from_raw(raw)


def feature_tensor(body, spec, case_offsets=None, hops=2):
    """(n, feature_width(hops)) float32 device tensor: raw features of the
    specimen seen through load_view and ``hops`` neighbour-mean hops through
    the bonds of ``body`` (a libpdx context of the same particles), built on
    the device (pdx_warm_raw_features, pdx_neighbor_features_device) from one
    upload of the specimen arrays."""
    return features_and_scale(body, spec, case_offsets, hops)[0]


def feature_matrix(body, spec, case_offsets=None, hops=2) -> np.ndarray:
    """feature_tensor on the host (the training data)."""
    return feature_tensor(body, spec, case_offsets, hops).cpu().numpy()


def target_scale_tensor(spec, case_offsets=None):
    """target_scale computed on the GPU: (n,) float64 CUDA tensor."""
    import torch
    X = torch.from_numpy(np.ascontiguousarray(spec.X, np.float64)).cuda()
    fext = torch.from_numpy(np.ascontiguousarray(spec.fext, np.float64)).cuda()
    s = torch.zeros(spec.n, dtype=torch.float64, device="cuda")
    for a, b in _per_case(spec, case_offsets):
        z = X[a:b, 2]
        H = z.max() - z.min()
        fe = fext[a:b]
        loaded = (fe.abs().sum(1) > 0).to(torch.float64).sum()
        nrm = torch.linalg.norm(fe.sum(0))
        p = nrm * spec.dx ** 3 / torch.clamp(loaded * spec.dx ** 2, min=1e-30)
        s[a:b] = torch.clamp(p, min=1e-12) * H / 1000.0
    return s


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
