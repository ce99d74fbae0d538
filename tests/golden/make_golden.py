"""Writes the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

Each fixture is one small seeded body: its inputs and the FP64 oracle's outputs
(neighbour CSR, bond data, K, K^-1, F, P, force, J2 state, broken flags), plus
closed-form values that do not come from the oracle (the interior shape tensor
of a regular lattice, the affine patch-test gradient).  The reference ships no
golden files for this path; these pin the oracle against regressions and give
the HIP path fixed vectors to meet.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pdoracle as po  # noqa: E402

H = 3.015


def lattice(nx, ny, nz, jitter, seed):
    g = [np.arange(k, dtype=np.float64) for k in (nx, ny, nz)]
    X = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3)
    if jitter:
        X = X + np.random.default_rng(seed).uniform(-jitter, jitter, X.shape)
    return X


def smooth_u(X, amp, seed):
    rng = np.random.default_rng(seed)
    L = np.ptp(X, axis=0) + 1.0
    u = np.zeros_like(X)
    for d in range(3):
        k = rng.uniform(0.5, 1.5, size=3) * np.pi / L
        ph = rng.uniform(0, 2 * np.pi)
        u[:, d] = amp * L.mean() * np.sin(X @ k + ph)
    return u


def moduli(n, E, nu, seed=None):
    Ev = np.full(n, E)
    if seed is not None:
        Ev = E * np.exp(0.3 * np.random.default_rng(seed).standard_normal(n))
    return Ev / (3 * (1 - 2 * nu)), Ev / (2 * (1 + nu))


def field_case(model, formulation, seed):
    X = lattice(6, 5, 5, 0.15, seed)
    n = len(X)
    bulk, shear = moduli(n, 15000.0, 0.3, seed + 1)
    cst = 18.0 * bulk / (np.pi * H ** 4)
    fixed = np.where(X[:, 2] < 0.5, 7, 0).astype(np.uint8)
    fext = np.zeros((n, 3))
    fext[X[:, 2] > 3.5, 2] = -2.5
    u = smooth_u(X, 2e-3 if model == "linear_elastic" else 1e-2, seed + 2)
    p = po.Problem(X, 1.0, H, bulk, shear, model=model, formulation=formulation, G_s=0.1, cstab=cst,
                   fixed=fixed, fext=fext, nthreads=1)
    if formulation == "nosb":
        p.shape_tensor("relative", 1e-12)
    p.set_u(u)
    p.compute_forces()
    out = dict(X=X, vol=np.ones(n), bulk=bulk, shear=shear, cstab=cst, fixed=fixed, fext=fext, u=u,
               horizon=H, G_s=0.1, row=p.row, col=p.col[: p.nb], xlen=p.xlen[: p.nb],
               omega=p.omega[: p.nb], f=p.f, residual=p.residual_norm())
    if formulation == "nosb":
        out.update(K=p.K, Kinv=p.Kinv, F=p.F, P=p.P)
    return out


def j2_case():
    X = lattice(5, 5, 5, 0.05, 31)
    n = len(X)
    E = 1000.0
    bulk, shear = moduli(n, E, 0.3)
    sy, hi = np.full(n, 2.0), np.full(n, 0.05 * E)
    G = np.diag([-0.004, 0.001, 0.0015]) + 1e-3 * np.random.default_rng(32).uniform(-1, 1, (3, 3))
    u = X @ G.T + smooth_u(X, 5e-4, 33)
    p = po.Problem(X, 1.0, H, bulk, shear, model="j2_plasticity", sy0=sy, h_iso=hi, nthreads=1)
    p.shape_tensor("relative", 1e-12)
    p.set_u(u)
    p.deformation_gradient()
    p.stress()
    P_trial = p.P.copy()
    p.commit()
    return dict(X=X, vol=np.ones(n), bulk=bulk, shear=shear, yield_stress=sy, h_iso=hi, u=u, horizon=H,
                F=p.F, P=P_trial, ep=p.ep, epbar=p.epbar)


def damage_case():
    X = lattice(6, 6, 5, 0.1, 41)
    n = len(X)
    bulk, shear = moduli(n, 15000.0, 0.3)
    u = smooth_u(X, 0.03, 42)
    p = po.Problem(X, 1.0, H, bulk, shear, nthreads=1)
    p.set_u(u)
    nbk = p.update_damage(0.01)
    return dict(X=X, vol=np.ones(n), bulk=bulk, shear=shear, u=u, horizon=H, s_c=0.01, row=p.row,
                col=p.col[: p.nb], broken=p.broken[: p.nb].copy(), n_broken=nbk,
                damage=p.local_damage())


def closed_form_case():
    """Regular lattice, spacing 1, delta = 3.015: interior K = k I with
    k = sum over offsets |d| < delta of w(|d|) d_x^2 (V = 1); u = G X gives
    F = I + G exactly at every particle."""
    X = lattice(9, 9, 9, 0.0, 0)
    r = np.arange(-3, 4)
    d = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    L = np.linalg.norm(d, axis=1)
    keep = (L > 0) & (L * L < H * H)
    w = 1.0 - L[keep] / H
    k = float(np.sum(w * d[keep, 0] ** 2))
    G = np.array([[1e-3, -2e-4, 5e-4], [3e-4, -8e-4, 1e-4], [-6e-4, 2e-4, 4e-4]])
    return dict(X=X, horizon=H, k_interior=k, n_interior_bonds=int(keep.sum()), G=G)


def main():
    cases = {
        "lattice_le": field_case("linear_elastic", "nosb", 11),
        "lattice_nh": field_case("neo_hookean", "nosb", 12),
        "lattice_bond": field_case("linear_elastic", "bond", 13),
        "j2": j2_case(),
        "damage": damage_case(),
        "closed_form": closed_form_case(),
    }
    for name, arrs in cases.items():
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrs.items()})
        print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
