"""Pins the FP64 oracle (oracle/pd_oracle.c) before it is trusted as the checker.

Each function of the oracle is checked against the reference's own known-answer
tests (file:line cited), closed forms, invariants, or an independent
pure-Python brute force on tiny inputs -- never against a retyping of its own
formula.
"""
import math

import numpy as np
import pytest
import scipy.optimize as so

from helpers import lattice
from oracle import pdoracle as po
from simul_amd import workloads as W


def grid2(nx, ny, dx=1.0):
    g = [np.arange(k, dtype=float) * dx for k in (nx, ny)]
    return np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 2)


# -- a1 neighbour sets ------------------------------------------------------
def test_neighbor_kats_3d():
    # backend/fea/peridynamics/tests/test_3d.py:67-138
    X = lattice(3, 3, 3)
    c = np.diff(po.neighbors(X, 1.1)[0])
    assert c[13] == 6 and c[0] == 3
    c = np.diff(po.neighbors(X, 1.5)[0])
    assert c[13] == 18 and c[0] == 6


def test_neighbor_kats_2d():
    # backend/fea/peridynamics/tests/test_particles.py:84-146
    c = np.diff(po.neighbors(grid2(3, 3), 1.5)[0])
    assert c[4] == 8 and all(c[k] == 3 for k in (0, 2, 6, 8))
    assert np.all(np.diff(po.neighbors(grid2(2, 2), 1.1)[0]) == 2)


def test_config1_bond_count_closed_form():
    # 20^3 cube, horizon 3.015 dx: interior stencil = lattice points with
    # 0 < |v|^2 < 3.015^2 -> 122; total over the cube by direct enumeration
    r = np.arange(-3, 4)
    V = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    d2 = (V ** 2).sum(1)
    stencil = V[(d2 > 0) & (d2 < 3.015 ** 2)]
    assert len(stencil) == 122
    n = 20
    per_axis = lambda off: np.array([n - abs(o) for o in off])
    total = int(np.sum(per_axis(stencil[:, 0]) * per_axis(stencil[:, 1]) * per_axis(stencil[:, 2])))
    row, _ = po.neighbors(W.cube().X, 3.015)
    cnt = np.diff(row)
    assert row[-1] == total == 816728
    assert cnt.min() == 28 and cnt.max() == 122


def test_cell_list_equals_brute_force_on_jittered_points():
    X = lattice(7, 6, 5, jitter=0.3, seed=3)
    row, col = po.neighbors(X, 2.2)
    brute = po.neighbor_counts_brute(X, 2.2)
    assert np.array_equal(np.diff(row), brute)
    # independent pure-Python set check on a subset
    for i in range(0, len(X), 17):
        s = {j for j in range(len(X)) if j != i and float(((X[j] - X[i]) ** 2).sum()) < 2.2 ** 2}
        assert s == set(col[row[i]:row[i + 1]].tolist())


# -- a2 bond data ------------------------------------------------------------
def test_bond_lengths_kat():
    # test_3d.py:179-213: horizon 1.1 on a unit lattice -> every |xi| = spacing
    X = lattice(2, 2, 2)
    p = po.Problem(X, 1.0, 1.1, 1.0, 1.0)
    assert np.allclose(p.xlen[: p.nb], 1.0, atol=1e-12)
    assert np.allclose(p.omega[: p.nb], 1.0 - 1.0 / 1.1)


# -- a3 shape tensor ------------------------------------------------------------
def test_shape_tensor_symmetric_spd_and_interior_closed_form():
    # test_3d.py:219-291 (symmetric everywhere, SPD at the centre)
    X = lattice(9, 9, 9)
    p = po.Problem(X, 1.0, 3.015, 1.0, 1.0)
    p.shape_tensor()
    assert np.allclose(p.K, np.transpose(p.K, (0, 2, 1)), atol=1e-12)
    c = 4 * 81 + 4 * 9 + 4
    assert np.all(np.linalg.eigvalsh(p.K[c]) > 0)
    # full neighbourhood: K = k I with k = sum over the stencil of w v_x^2
    r = np.arange(-3, 4)
    V = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3).astype(float)
    L = np.linalg.norm(V, axis=1)
    m = (L > 0) & (L < 3.015)
    k = np.sum((1 - L[m] / 3.015) * V[m, 0] ** 2)
    assert np.allclose(p.K[c], k * np.eye(3), rtol=1e-13, atol=1e-12)
    assert np.allclose(p.Kinv[c] @ p.K[c], np.eye(3), atol=1e-12)


def test_singular_tests():
    # one isolated particle: K = 0 -> identity inverse under both tests
    X = np.vstack([lattice(4, 4, 4), [[40.0, 40.0, 40.0]]])
    p = po.Problem(X, 1e-9, 3.015, 1.0, 1.0)  # tiny volumes: det K ~ 1e-40
    p.shape_tensor("absolute", 1e-12)
    assert p.sing[:-1].all()  # the unit-dependent reference test rejects them
    p.shape_tensor("relative", 1e-12)
    assert not p.sing[:-1].any() and p.sing[-1] == 1
    assert np.allclose(p.Kinv[-1], np.eye(3))


# -- a4 deformation gradient ---------------------------------------------------
@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_patch_test_affine(jitter):
    X = lattice(6, 6, 6, jitter=jitter, seed=1)
    p = po.Problem(X, 1.0, 3.015, 1.0, 1.0)
    p.shape_tensor()
    p.deformation_gradient()
    assert np.allclose(p.F, np.eye(3), atol=1e-12)  # test_3d.py:293-332 (every particle here)
    G = np.random.default_rng(0).uniform(-0.01, 0.01, (3, 3))
    p.set_u(X @ G.T)
    p.deformation_gradient()
    assert np.allclose(p.F, np.eye(3) + G, atol=1e-12)


# -- a5 stress ----------------------------------------------------------------------
def _one_particle(model, F, E=200e3, nu=0.3, **kw):
    Kb, mu = W.bulk_shear(E, nu)
    X = np.zeros((1, 3))
    p = po.Problem(X, 1.0, 1.0, Kb, mu, model=model, **kw)
    p.F[0] = F
    p.Kinv[0] = np.eye(3)
    p.stress()
    return p


def test_linear_elastic_closed_form_uniaxial():
    # sigma = lam tr(eps) I + 2 mu eps; uniaxial strain e: s11 = (lam + 2mu) e, s22 = lam e
    e = 1e-3
    Kb, mu = W.bulk_shear(1000.0, 0.25)
    lam = Kb - 2 * mu / 3
    p = _one_particle("linear_elastic", np.diag([1 + e, 1, 1]), E=1000.0, nu=0.25)
    assert np.allclose(p.P[0], np.diag([(lam + 2 * mu) * e, lam * e, lam * e]), rtol=1e-12)


def test_neo_hookean_rotation_and_tangent():
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    assert np.abs(_one_particle("neo_hookean", R).P[0]).max() < 1e-9  # rigid rotation: no stress
    # initial tangent = linear elastic C (backend/fea/fem/tests/test_fem.py:83-102)
    H = np.random.default_rng(2).uniform(-1, 1, (3, 3))
    h = 1e-7
    Pn = _one_particle("neo_hookean", np.eye(3) + h * H, E=1000.0).P[0] / h
    Pl = _one_particle("linear_elastic", np.eye(3) + h * H, E=1000.0).P[0] / h
    assert np.allclose(Pn, Pl, rtol=1e-5, atol=1e-5 * np.abs(Pl).max())


def test_neo_hookean_volumetric_closed_form():
    # F = s I: B - I = (s^2 - 1) I, J = s^3; P = J sigma F^-T = (mu (s^2-1) + 3 lam ln s) / s I
    s = 0.97
    E, nu = 1000.0, 0.3
    Kb, mu = W.bulk_shear(E, nu)
    lam = Kb - 2 * mu / 3
    P = _one_particle("neo_hookean", s * np.eye(3), E=E, nu=nu).P[0]
    assert np.allclose(P, (mu * (s * s - 1) + 3 * lam * math.log(s)) / s * np.eye(3), rtol=1e-12)


def _j2(F, sy, H, E=200e3, nu=0.3, hk=0.0, state=None):
    n = 1
    p = _one_particle("j2_plasticity", F, E=E, nu=nu, sy0=np.full(n, sy), h_iso=np.full(n, H),
                      h_kin=np.full(n, hk))
    return p


def _mises(s):
    d = s - np.trace(s) / 3 * np.eye(3)
    return math.sqrt(1.5 * np.sum(d * d))


def test_j2_kats():
    # backend/fea/fem/tests/test_j2_plasticity.py:149-288
    F = np.array([[1.001, 0.0002, 0], [0.0001, 0.999, 0], [0, 0, 1.0]])
    pe = _j2(F, 1e12, 0.0)
    pl = _one_particle("linear_elastic", F)
    assert np.allclose(pe.P[0], pl.P[0], rtol=1e-10)         # elastic == linear
    assert pe.epbar_t[0] == 0.0
    Fy = np.diag([1.01, 1.0, 1.0])
    pp = _j2(Fy, 250.0, 0.0)
    assert pp.epbar_t[0] > 1e-12                               # yields
    assert _mises(pp.P[0]) <= 250.0 * (1 + 1e-12)              # stress cap (perfect plasticity)
    assert abs(_mises(pp.P[0]) - 250.0) < 1e-9                 # consistency: on the yield surface
    H = 1e3
    ph = _j2(Fy, 250.0, H)
    assert abs(_mises(ph.P[0]) - (250.0 + H * ph.epbar_t[0])) < 1e-9  # hardened surface
    assert ph.epbar_t[0] < pp.epbar_t[0]


def test_j2_uniaxial_onset_closed_form():
    # uniaxial strain e11: von Mises of the trial stress = 2 mu e (deviator of diag(e,0,0))
    E, nu, sy = 200e3, 0.3, 250.0
    _, mu = W.bulk_shear(E, nu)
    e_y = sy / (2 * mu)
    assert _j2(np.diag([1 + 0.999 * e_y, 1, 1]), sy, 0.0).epbar_t[0] == 0.0
    assert _j2(np.diag([1 + 1.001 * e_y, 1, 1]), sy, 0.0).epbar_t[0] > 0.0


_DEV_D = np.diag([2.0, -1.0, -1.0]) / 3.0  # deviator of diag(1, 0, 0); |dev D| = sqrt(2/3)


@pytest.mark.parametrize("hk", [0.0, 4e3, 2.5e4])
def test_j2_kinematic_uniaxial_closed_form(hk):
    # uniaxial strain e past yield from the virgin state: the trial deviator is
    # 2 mu e dev D with Mises value q = 2 mu e, and one radial-return step with
    # linear isotropic (Hi) and kinematic (hk) hardening gives
    #   dgamma = (2 mu e - sy) / (3 mu + Hi + hk),   n = dev D / |dev D|
    #   eps_p = sqrt(3/2) dgamma n = 1.5 dgamma dev D
    #   beta  = sqrt(2/3) hk dgamma n = hk dgamma dev D
    #   s     = 2 mu dev(eps - eps_p) = (2 mu e - 3 mu dgamma) dev D   (elasticity)
    # and the relative stress s - beta, Mises value q - (3 mu + hk) dgamma, lies on
    # the isotropically hardened surface sy + Hi dgamma
    E, nu, sy, Hi = 200e3, 0.3, 250.0, 1.5e3
    Kb, mu = W.bulk_shear(E, nu)
    e = 4.0 * sy / (2 * mu)
    p = _j2(np.diag([1 + e, 1, 1]), sy, Hi, hk=hk)
    dg = (2 * mu * e - sy) / (3 * mu + Hi + hk)
    assert abs(p.epbar_t[0] - dg) < 1e-12 * dg
    assert np.allclose(p.ep_t[0].reshape(3, 3), 1.5 * dg * _DEV_D, rtol=1e-11, atol=1e-18)
    assert np.allclose(p.beta_t[0].reshape(3, 3), hk * dg * _DEV_D, rtol=1e-11, atol=1e-12)
    # the pressure stays elastic (tr eps_p = 0): tr P / 3 = Kb e
    assert abs(np.trace(p.P[0]) / 3 - Kb * e) < 1e-10 * Kb * e
    s = p.P[0] - np.trace(p.P[0]) / 3 * np.eye(3)
    assert np.allclose(s, 2 * mu * (e * _DEV_D - p.ep_t[0].reshape(3, 3)), rtol=1e-10, atol=1e-9)
    assert np.allclose(s, (2 * mu * e - 3 * mu * dg) * _DEV_D, rtol=1e-10, atol=1e-9)
    assert abs(_mises(s - p.beta_t[0].reshape(3, 3)) - (sy + Hi * dg)) < 1e-9 * sy


def test_j2_kinematic_bauschinger_reverse_onset():
    # forward uniaxial strain e1 (one plastic step, committed: eps_p = 1.5 g dev D,
    # beta = hk g dev D, epbar = g), then the strain is taken back down.  The trial
    # relative stress at strain e is [2 mu (e - 1.5 g) - hk g] dev D, Mises value
    # |2 mu (e - 1.5 g) - hk g|, so reverse yield starts at
    #   e_rev = 1.5 g + (hk g - sy - Hi g) / (2 mu),
    # above the onset e_iso = 1.5 g - (sy + Hi g) / (2 mu) that the same plastic
    # strain gives without the back stress.  The onset therefore pins beta.
    E, nu, sy, Hi, hk = 200e3, 0.3, 250.0, 1.5e3, 2.0e4
    Kb, mu = W.bulk_shear(E, nu)
    e1 = 3.0 * sy / (2 * mu)
    p = _j2(np.diag([1 + e1, 1, 1]), sy, Hi, hk=hk)
    g = float(p.epbar_t[0])
    assert g > 0
    p.commit()
    e_rev = 1.5 * g + (hk * g - sy - Hi * g) / (2 * mu)
    e_iso = 1.5 * g - (sy + Hi * g) / (2 * mu)
    assert e_rev - e_iso > 0.1 * sy / (2 * mu)
    for side, plastic in ((+1.0, False), (-1.0, True)):
        p.F[0] = np.diag([1 + e_rev + side * 5e-3 * abs(e_rev), 1, 1])
        p.stress()
        assert (p.epbar_t[0] > g) == plastic, (side, p.epbar_t[0], g)
    # between the two onsets: plastic with the back stress (elastic without it)
    p.F[0] = np.diag([1 + 0.5 * (e_rev + e_iso), 1, 1])
    p.stress()
    assert p.epbar_t[0] > g
    # the reverse step moves the back stress against the first one
    assert p.beta_t[0][0] < p.beta[0][0]


# -- a6 forces ---------------------------------------------------------------------
def _brute_force_f(X, u, h, Kb, mu, Gs, c):
    """Independent pure-Python restatement of nosb.py:59-275 for tiny inputs."""
    n = len(X)
    x = X + u
    nb = [[j for j in range(n) if j != i and np.sum((X[j] - X[i]) ** 2) < h * h] for i in range(n)]
    K, Ki, P = [], [], []
    for i in range(n):
        k = np.zeros((3, 3))
        for j in nb[i]:
            xi = X[j] - X[i]
            k += (1 - np.linalg.norm(xi) / h) * np.outer(xi, xi)
        K.append(k)
        Ki.append(np.linalg.inv(k))
    for i in range(n):
        N = np.zeros((3, 3))
        for j in nb[i]:
            xi = X[j] - X[i]
            N += (1 - np.linalg.norm(xi) / h) * np.outer(x[j] - x[i], xi)
        F = N @ Ki[i]
        eps = 0.5 * (F + F.T) - np.eye(3)
        P.append((Kb - 2 * mu / 3) * np.trace(eps) * np.eye(3) + 2 * mu * eps)
    f = np.zeros((n, 3))
    for i in range(n):
        for j in nb[i]:
            xi = X[j] - X[i]
            w = 1 - np.linalg.norm(xi) / h
            f[i] += w * (P[i] @ Ki[i] + P[j] @ Ki[j]) @ xi
            eta = x[j] - x[i]
            s = (np.linalg.norm(eta) - np.linalg.norm(xi)) / np.linalg.norm(xi)
            f[i] += Gs * c * s * w * eta / np.linalg.norm(eta)
    return f


def test_force_matches_pure_python_brute_force():
    X = lattice(3, 3, 3, jitter=0.1, seed=4)
    u = np.random.default_rng(5).uniform(-0.01, 0.01, X.shape)
    Kb, mu = 100.0, 40.0
    c = 18 * Kb / (math.pi * 1.8 ** 4)
    p = po.Problem(X, 1.0, 1.8, Kb, mu, G_s=0.1, cstab=c)
    p.shape_tensor()
    p.set_u(u)
    p.compute_forces()
    assert np.allclose(p.f, _brute_force_f(X, u, 1.8, Kb, mu, 0.1, c), rtol=1e-10, atol=1e-10)


def test_force_invariants():
    X = lattice(7, 7, 7, jitter=0.15, seed=6)
    p = po.Problem(X, 1.0, 3.015, 1000.0, 500.0, model="neo_hookean")
    p.shape_tensor()
    p.set_u(np.random.default_rng(1).uniform(-0.02, 0.02, X.shape))
    p.compute_forces()
    assert np.abs(p.f.sum(0)).max() < 1e-9 * np.abs(p.f).sum()    # linear momentum
    # affine deformation: F uniform, P uniform -> interior force vanishes (LE, G_s = 0)
    q = po.Problem(lattice(9, 9, 9), 1.0, 3.015, 1000.0, 500.0, G_s=0.0)
    q.shape_tensor()
    q.set_u(q.X @ np.diag([1e-3, -2e-4, 5e-4]).T)
    q.compute_forces()
    c = 4 * 81 + 4 * 9 + 4
    assert np.abs(q.f[c]).max() < 1e-10


def test_bond_based_force_closed_form():
    # two particles, PMB (quasi_static.py:119-143): f = c s w V eta/|eta|
    X = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    p = po.Problem(X, 2.0, 1.5, 1.0, 1.0, formulation="bond", cstab=7.0)
    p.set_u(np.array([[0, 0, 0], [0.1, 0, 0]]))
    p.compute_forces()
    expect = 7.0 * 0.1 * (1 - 1 / 1.5) * 2.0
    assert np.allclose(p.f, [[expect, 0, 0], [-expect, 0, 0]], rtol=1e-14)


# -- a7 dynamic relaxation ------------------------------------------------------------
def _brute_dt(X, h, V, m, modulus):
    """Independent pure-Python restatement of nosb_solver.py:94-154 on tiny inputs:
    bonds by brute force, K_i = sum w xi xi^T V_j (nosb.py:59-94) inverted by LAPACK,
    dpsi_k = w K_i^-1 xi_k (no V_j, as the reference loop has it),
    k_eff = modulus V_i (|sum dpsi|^2 + sum |dpsi|^2), dt = 0.5 * 2 / sqrt(max k_eff / m)."""
    n = len(X)
    lmax = 0.0
    for i in range(n):
        if m[i] < 1e-30:
            continue
        nb = [j for j in range(n) if j != i and np.sum((X[j] - X[i]) ** 2) < h * h]
        K = sum(((1 - np.linalg.norm(X[j] - X[i]) / h) * np.outer(X[j] - X[i], X[j] - X[i]) * V[j]
                 for j in nb), np.zeros((3, 3)))
        Ki = np.linalg.inv(K)
        d = [(1 - np.linalg.norm(X[j] - X[i]) / h) * Ki @ (X[j] - X[i]) for j in nb]
        keff = modulus * V[i] * (np.sum(np.sum(d, axis=0) ** 2) + sum(float(v @ v) for v in d))
        lmax = max(lmax, keff / m[i])
    return 0.5 * 2.0 / math.sqrt(lmax)


def test_estimate_dt_matches_pure_python_brute_force():
    # jittered 4^3 lattice, non-uniform volumes and masses: a dropped V_i, a V_j
    # in dpsi, a wrong safety factor or a missing sum of |dpsi|^2 all change dt
    X = lattice(4, 4, 4, jitter=0.2, seed=11)
    rng = np.random.default_rng(12)
    V = rng.uniform(0.6, 1.4, len(X))
    m = V * rng.uniform(0.8, 1.2, len(X)) * 1.9e-3
    Kb, mu = 1300.0, 600.0
    p = po.Problem(X, V, 2.2, Kb, mu, mass=m)
    p.shape_tensor("absolute", 1e-12)
    assert not p.sing.any()
    modulus = (Kb - 2 * mu / 3) + 2 * mu
    assert abs(p.estimate_dt(Kb, mu, 1.0) - _brute_dt(X, 2.2, V, m, modulus)) < 1e-12 * p.estimate_dt(Kb, mu, 1.0)


def test_estimate_dt_interior_closed_form():
    # regular lattice, uniform volume V != 1, only the centre particle of a 9^3
    # block has a finite mass m0 (all others 1e30 -> negligible eigenvalue).  Its
    # neighbourhood is the full stencil: K = k I with k = V sum w |v|^2 / 3,
    # sum dpsi = 0 by symmetry, sum |dpsi|^2 = sum w^2 |v|^2 / k^2, so
    #   lambda = modulus V (sum w^2 |v|^2) / k^2 / m0,  dt = 0.5 * 2 / sqrt(lambda)
    h, V, m0 = 3.015, 0.7, 2.3e-3
    X = lattice(9, 9, 9)
    n = len(X)
    c = 4 * 81 + 4 * 9 + 4
    m = np.full(n, 1e30)
    m[c] = m0
    Kb, mu = 2000.0, 900.0
    p = po.Problem(X, V, h, Kb, mu, mass=m)
    p.shape_tensor("absolute", 1e-12)
    r = np.arange(-3, 4)
    S = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3).astype(float)
    L = np.linalg.norm(S, axis=1)
    sel = (L > 0) & (L < h)
    w = 1 - L[sel] / h
    k = V * np.sum(w * L[sel] ** 2) / 3
    lam = (Kb + 4 * mu / 3) * V * np.sum(w ** 2 * L[sel] ** 2) / k ** 2 / m0
    dt = p.estimate_dt(Kb, mu, 1.0)
    assert abs(dt - 1.0 / math.sqrt(lam)) < 1e-12 * dt
    # 2-D: modulus lam + 2 mu with lam = K - mu (nosb_solver.py:112-115)
    g = np.arange(7, dtype=float)
    X2 = np.stack(np.meshgrid(g, g, indexing="ij"), -1).reshape(-1, 2)
    m2 = np.full(len(X2), 1e30)
    c2 = 3 * 7 + 3
    m2[c2] = m0
    p2 = po.Problem(X2, V, h, Kb, mu, mass=m2)
    p2.shape_tensor("absolute", 1e-12)
    S2 = np.stack(np.meshgrid(r, r, indexing="ij"), -1).reshape(-1, 2).astype(float)
    L2 = np.linalg.norm(S2, axis=1)
    sel2 = (L2 > 0) & (L2 < h)
    w2 = 1 - L2[sel2] / h
    k2 = V * np.sum(w2 * L2[sel2] ** 2) / 2
    lam2 = (Kb + mu) * V * np.sum(w2 ** 2 * L2[sel2] ** 2) / k2 ** 2 / m0
    dt2 = p2.estimate_dt(Kb, mu, 1.0)
    assert abs(dt2 - 1.0 / math.sqrt(lam2)) < 1e-12 * dt2


def test_estimate_dt_fallback_cfl():
    # no particle with mass >= 1e-30: the conservative CFL branch,
    # 0.1 * spacing / sqrt((K + 4 mu / 3) / rho) with spacing = mean(V)^(1/3)
    X = lattice(3, 3, 3)
    V = np.linspace(0.5, 1.5, len(X))
    p = po.Problem(X, V, 1.5, 1000.0, 300.0, mass=np.zeros(len(X)))
    p.shape_tensor()
    rho = 1.9e-9
    expect = 0.1 * np.mean(V) ** (1 / 3) / math.sqrt((1000.0 + 400.0) / rho)
    assert abs(p.estimate_dt(1000.0, 300.0, rho) - expect) < 1e-9 * expect


def test_kinetic_dr_single_dof_closed_form():
    # two particles, one fixed; kinetic DR of nosb_solver.py:156-256 converges to
    # the equilibrium of the PMB spring: s = fext / (c w V) (+ geometric exactness)
    X = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    fixed = np.array([7, 6], np.uint8)  # particle 1 free in x only
    fext = np.array([[0, 0, 0], [0.01, 0, 0]])
    p = po.Problem(X, 1.0, 1.5, 1.0, 1.0, formulation="bond", cstab=3.0, fixed=fixed, fext=fext, density=1.0)
    r = p.kd_solve(0.1, 0.0, max_it=200000, tol=1e-10)
    assert r["converged"]
    assert abs(p.u[1, 0] - 0.01 / (3.0 * (1 - 1 / 1.5))) < 1e-9


def test_adaptive_dr_reaches_newton_equilibrium():
    spec = W.cube(n_side=8)
    p = po.Problem(spec.X, spec.vol, spec.horizon, spec.bulk, spec.shear, model="neo_hookean", fixed=spec.fixed)
    p.shape_tensor()
    p.set_u(spec.u0)
    p.compute_forces()
    cold = p.residual_norm()
    r = p.adr_solve(p.adr_mass(1.0), max_it=50000, tol=1e-11, ref=cold)
    assert r["converged"]
    u_adr = p.u.copy()
    free = p.free_mask().reshape(-1)

    def R(w):
        uu = spec.u0.reshape(-1).copy()
        uu[free] += w
        p.set_u(uu.reshape(-1, 3))
        p.compute_forces()
        return p.f.reshape(-1)[free].copy()

    w = so.newton_krylov(R, u_adr.reshape(-1)[free] * 0.0, f_tol=1e-11 * cold, method="lgmres")
    un = spec.u0.reshape(-1).copy()
    un[free] += w
    assert np.linalg.norm(u_adr.reshape(-1) - un) / np.linalg.norm(un) < 1e-8


# -- explicit dynamics (explicit.py:62-158) + damage schedule (damage.py:144-152) ----------
def _spring_pair(c=3.0, V=2.0, m=0.5, h=1.5, v0=1e-3):
    # particle 0 fixed, particle 1 at +1 on x, free; PMB bond along x: for motion
    # along the bond f_1 = -c w V u (|xi| = 1), a linear spring k = c w V / m
    X = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    fixed = np.array([7, 6], np.uint8)
    p = po.Problem(X, V, h, 1.0, 1.0, formulation="bond", cstab=c, fixed=fixed, mass=np.full(2, m))
    p.v[1, 0] = v0
    return p, c * (1 - 1 / h) * V / m


def test_explicit_verlet_spring_closed_form():
    # velocity Verlet from u = 0, v = v0, a = 0: u_n = v0 dt sin(n th) / sin(th),
    # cos(th) = 1 - k dt^2 / 2 (the exact discrete solution of the recurrence
    # u_(n+1) = 2 u_n - u_(n-1) - k dt^2 u_n)
    p, k = _spring_pair()
    dt, v0 = 0.05, 1e-3
    th = math.acos(1 - k * dt * dt / 2)
    for n_done, n in ((0, 7), (7, 40)):
        nb, ke = p.explicit_run(dt, 0.0, n - n_done)
        assert nb == 0
        assert abs(p.u[1, 0] - v0 * dt * math.sin(n * th) / math.sin(th)) < 1e-12 * v0
    assert np.all(p.u[0] == 0) and np.all(p.u[1, 1:] == 0)


def test_explicit_damping_free_particle_closed_form():
    # a particle without bonds: a = 0, v_n = (1 - d)^n v0, u_n = v0 dt sum_{k<n} (1-d)^k
    X = np.array([[0.0, 0, 0], [10.0, 0, 0]])
    p = po.Problem(X, 1.0, 1.5, 1.0, 1.0, formulation="bond", cstab=1.0, mass=np.ones(2))
    p.v[:] = [[2e-3, -1e-3, 5e-4], [0, 0, 0]]
    d, dt, n = 0.03, 0.1, 25
    _, ke = p.explicit_run(dt, d, n)
    q = (1 - d) ** np.arange(n)
    assert np.allclose(p.u[0], np.array([2e-3, -1e-3, 5e-4]) * dt * q.sum(), rtol=1e-13)
    v = np.array([2e-3, -1e-3, 5e-4]) * (1 - d) ** n
    assert np.allclose(p.v[0], v, rtol=1e-13)
    assert abs(ke - 0.5 * float(v @ v)) < 1e-13 * ke


def test_explicit_damage_schedule_breaks_at_the_closed_form_step():
    # the spring pair kicked hard: the bond breaks at the first step whose stretch
    # u_n / |xi| exceeds s_c (both directed entries), after which the particle
    # keeps the velocity of that step plus half a step of its last acceleration
    p, k = _spring_pair(v0=0.05)
    dt, v0, s_c = 0.05, 0.05, 0.015
    th = math.acos(1 - k * dt * dt / 2)
    un = [v0 * dt * math.sin(n * th) / math.sin(th) for n in range(200)]
    n_star = next(n for n in range(200) if un[n] > s_c)
    nb, _ = p.explicit_run(dt, 0.0, n_star - 1, critical_stretch=s_c)
    assert nb == 0 and not p.broken[:2].any()
    nb, _ = p.explicit_run(dt, 0.0, 1, critical_stretch=s_c)
    assert nb == 2 and p.broken[:2].all()
    a_last = -k * un[n_star]
    assert abs(p.a[1, 0] - a_last) < 1e-12
    v_last, u0 = p.v[1, 0], p.u[1, 0]
    # next step: moves with v and a, then no force (a = 0) and v + a dt / 2 from there on
    p.explicit_run(dt, 0.0, 5, critical_stretch=s_c)
    v_after = v_last + 0.5 * a_last * dt
    assert abs(p.v[1, 0] - v_after) < 1e-15
    assert abs(p.u[1, 0] - (u0 + v_last * dt + 0.5 * a_last * dt * dt + 4 * v_after * dt)) < 1e-14
    assert np.allclose(p.local_damage(), [1.0, 1.0])


# -- a9 damage ------------------------------------------------------------------------
def test_damage_closed_form_uniform_stretch():
    # u = e x1: a bond along direction n stretches sqrt((1+e)^2 n1^2 + 1 - n1^2) - 1
    X = lattice(6, 6, 6)
    e = 0.05
    p = po.Problem(X, 1.0, 3.015, 1.0, 1.0)
    p.set_u(X * np.array([e, 0, 0]))
    s_c = 0.03
    nbk = p.update_damage(s_c)
    i, j = p.bond_pairs()
    n1 = (X[j, 0] - X[i, 0]) / np.linalg.norm(X[j] - X[i], axis=1)
    s = np.sqrt((n1 * (1 + e)) ** 2 + 1 - n1 ** 2) - 1
    assert nbk == int((s > s_c).sum())
    assert np.array_equal(p.broken[: p.nb].astype(bool), s > s_c)
    dmg = p.local_damage()
    assert np.allclose(dmg, np.bincount(i, weights=(s > s_c), minlength=len(X)) / np.diff(p.row))
