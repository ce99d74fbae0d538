"""Shared builders: the same seeded inputs for the HIP path and the oracle."""
from __future__ import annotations

import numpy as np

from oracle import pdoracle as po
from simul_amd import workloads as W
from simul_amd.body import PDBody


def lattice(nx, ny, nz, dx=1.0, jitter=0.0, seed=0):
    g = [np.arange(k, dtype=np.float64) * dx for k in (nx, ny, nz)]
    X = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3)
    if jitter:
        X = X + np.random.default_rng(seed).uniform(-jitter, jitter, X.shape) * dx
    return X


def smooth_u(X, amp=1e-3, seed=1):
    """A smooth non-affine displacement field (a few sine modes)."""
    rng = np.random.default_rng(seed)
    L = np.ptp(X, axis=0) + 1.0
    u = np.zeros_like(X)
    for d in range(X.shape[1]):
        k = rng.uniform(0.5, 1.5, size=X.shape[1]) * np.pi / L
        ph = rng.uniform(0, 2 * np.pi)
        u[:, d] = amp * L.mean() * np.sin(X @ k + ph)
    return u


def pair(X, horizon, E=15000.0, nu=0.3, model="linear_elastic", formulation="nosb", G_s=0.1,
         fixed=None, fext=None, dx=1.0, hetero=False, seed=3, **kw):
    """Builds (gpu_body, oracle_problem) over identical inputs."""
    n = len(X)
    Ev = np.full(n, E)
    if hetero:
        Ev = E * np.exp(0.4 * np.random.default_rng(seed).standard_normal(n))
    bulk, shear = W.bulk_shear(Ev, nu)
    vol = np.full(n, dx ** X.shape[1])
    cst = 18.0 * bulk / (np.pi * horizon ** 4)
    mass = vol * W.RHO_BONE
    j2 = {}
    if model == "j2_plasticity":
        j2 = dict(yield_stress=kw.pop("yield_stress"), h_iso=kw.pop("h_iso"), h_kin=kw.pop("h_kin", None))
    body = PDBody(X, vol, horizon, bulk, shear, mass=mass, model=model, formulation=formulation,
                  G_s=G_s, cstab=cst, fixed=fixed, fext=fext, **j2, **kw)
    oj = {}
    if model == "j2_plasticity":
        oj = dict(sy0=j2["yield_stress"], h_iso=j2["h_iso"], h_kin=j2["h_kin"])
    prob = po.Problem(X, vol, horizon, bulk, shear, mass=mass, model=model, formulation=formulation,
                      G_s=G_s, cstab=cst, fixed=fixed, fext=fext, **oj)
    if formulation == "nosb":
        prob.shape_tensor("relative", 1e-12)
    return body, prob


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
