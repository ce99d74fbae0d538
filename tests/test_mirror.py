"""The reference-API mirror (simul_amd.peridynamics, simul_amd.framework).

The tests follow the reference's own tests (backend/fea/peridynamics/tests/
test_3d.py and test_particles.py) with the same inputs and expectations, and
add parity of the framework Solver against the FP64 oracle on the cube of
BASELINE config 1 (20^3 particles, dx 1 mm, bottom fixed, top prescribed
-0.5% axial strain).  Host-only parts run on the CPU; everything that reaches
libpdx is marked gpu.
"""
import math

import numpy as np
import pytest

from helpers import rel_l2
from oracle import pdoracle as po
from simul_amd.framework import Material, Method, Solver, create_domain, create_particle_domain
from simul_amd.peridynamics import (BondSystem, DamageModel, LoadControl, NeighborSearch, NOSBCompute,
                                    NOSBMaterial, NOSBSolver, ParticleSystem, QuasiStaticSolver)
from simul_amd.validation import FEAValidationError


# ---------------------------------------------------------------------------
# host-only (CPU)
# ---------------------------------------------------------------------------
def test_particle_grid_3d():
    # test_3d.py:19-61
    ps = ParticleSystem(27, dim=3)
    ps.initialize_from_grid((0.0, 0.0, 0.0), 0.1, (3, 3, 3), density=1000.0)
    pos = ps.get_positions()
    assert pos.shape == (27, 3)
    assert np.allclose(pos[0], 0.0) and np.allclose(pos[-1], [0.2, 0.2, 0.2])
    ps2 = ParticleSystem(8, dim=3)
    ps2.initialize_from_grid((0, 0, 0), 0.1, (2, 2, 2), density=1000.0)
    assert np.allclose(ps2.volume.to_numpy(), 1e-3) and np.allclose(ps2.mass.to_numpy(), 1.0)


def test_particle_fixed_and_displacement_2d():
    # test_particles.py:17-78
    ps = ParticleSystem(20, dim=2)
    ps.initialize_from_grid((0, 0), 0.1, (5, 4), density=1000.0)
    ps.set_fixed_particles(np.array([0, 1, 2, 3, 4]))
    fixed = ps.fixed.to_numpy()
    assert fixed.sum() == 5 and np.all(fixed[:5] == 1) and np.all(fixed[5:] == 0)
    ps = ParticleSystem(4, dim=2)
    ps.initialize_from_grid((0, 0), 1.0, (2, 2), density=1000.0)
    x = ps.x.to_numpy()
    x[:, 0] += 0.1
    ps.x.from_numpy(x)
    d = ps.get_displacements()
    assert np.allclose(d[:, 0], 0.1) and np.allclose(d[:, 1], 0.0)


def test_particle_validation():
    ps = ParticleSystem(4, dim=2)
    with pytest.raises(FEAValidationError):
        ps.initialize_from_grid((0, 0), -1.0, (2, 2))
    with pytest.raises(FEAValidationError):
        ps.initialize_from_grid((0, 0), 1.0, (2, 2), density=0.0)
    with pytest.raises(ValueError):
        ps.initialize_from_grid((0, 0), 1.0, (3, 2))
    with pytest.raises(FEAValidationError):
        NeighborSearch((0, 0), (1, 1), horizon=0.0, max_particles=4, dim=2)


def test_critical_stretch_formulas():
    # test_particles.py:187-218
    E, G_c, delta = 70e9, 1000, 0.003
    assert np.isclose(DamageModel.compute_critical_stretch(E, G_c, delta, dim=3),
                      np.sqrt(5 * G_c / (9 * E * delta)))
    assert np.isclose(DamageModel.compute_critical_stretch(E, G_c, delta, dim=2),
                      np.sqrt(4 * np.pi * G_c / (9 * E * delta)))
    assert DamageModel.compute_critical_stretch_bone("cortical") > 0
    assert DamageModel.compute_critical_stretch_bone("cancellous") > 0


def test_nosb_material_constants():
    m = NOSBMaterial(1e6, 0.25, dim=3)
    assert np.isclose(m.K, 1e6 / 1.5) and np.isclose(m.mu, 1e6 / 2.5)
    m.set_stabilization_modulus(0.03)
    assert np.isclose(m.c_bond, 18 * m.K / (math.pi * 0.03 ** 4))
    with pytest.raises(FEAValidationError):
        NOSBMaterial(-1.0, 0.3)


def test_framework_domain_and_material():
    d = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(2, 2, 2), n_divisions=(3, 3, 3))
    pos = d.get_positions()
    assert pos.shape == (27, 3) and np.allclose(pos[1], [0, 0, 1])  # i-major, then j, then k
    assert set(d.select(axis=2, value=0.0)) == set(np.where(pos[:, 2] == 0)[0])
    assert len(d.select_boundary()) == 26
    p = create_particle_domain(pos[::2], Method.PD)
    assert p.n_points == 14 and np.array_equal(p.get_positions(), pos[::2])
    m = Material(E=15000.0, nu=0.3, density=1.9e-9, dim=3)
    assert np.isclose(m.K_bulk, 15000.0 / 1.2) and np.isclose(m.mu, 15000.0 / 2.6)
    for bad in (dict(E=-1.0, nu=0.3), dict(E=1.0, nu=0.5), dict(E=1.0, nu=0.3, density=0.0)):
        with pytest.raises(FEAValidationError):
            Material(**bad)
    with pytest.raises(FEAValidationError):
        Material(E=1.0, nu=0.3, constitutive_model="j2_plasticity")
    with pytest.raises(NotImplementedError):
        create_domain(Method.FEM, dim=2, origin=(0, 0), size=(1, 1), n_divisions=(2, 2))


# ---------------------------------------------------------------------------
# through libpdx (GPU)
# ---------------------------------------------------------------------------
def _grid(n_per, spacing=1.0, dim=3):
    ps = ParticleSystem(int(np.prod(n_per)), dim=dim)
    ps.initialize_from_grid((0,) * dim, spacing, n_per, density=1000.0)
    return ps


@pytest.mark.gpu
def test_mirror_neighbor_kats(gpu):
    # test_3d.py:67-138 and test_particles.py:84-146
    for h, centre, corner in ((1.1, 6, 3), (1.5, 18, 6)):
        ps = _grid((3, 3, 3))
        ns = NeighborSearch((-1, -1, -1), (3, 3, 3), horizon=h, max_particles=27, max_neighbors=30, dim=3)
        ns.build(ps.X, 27)
        c = ns.get_all_neighbor_counts()
        assert c[13] == centre and c[0] == corner
    ps = _grid((3, 3), dim=2)
    ns = NeighborSearch((-1, -1), (3, 3), horizon=1.5, max_particles=9, max_neighbors=10, dim=2)
    ns.build(ps.X, 9)
    c = ns.get_all_neighbor_counts()
    assert c[4] == 8 and all(c[k] == 3 for k in (0, 2, 6, 8))
    ps = _grid((2, 2), dim=2)
    ns = NeighborSearch((-1, -1), (2, 2), horizon=1.1, max_particles=4, max_neighbors=10, dim=2)
    ns.build(ps.X, 4)
    assert np.all(ns.get_all_neighbor_counts() == 2)


@pytest.mark.gpu
def test_mirror_bonds(gpu):
    # test_3d.py:143-213
    ps = _grid((2, 2, 2))
    ns = NeighborSearch((-1, -1, -1), (2, 2, 2), horizon=1.5, max_particles=8, dim=3)
    ns.build(ps.X, 8)
    bonds = BondSystem(8, max_bonds=30, dim=3)
    bonds.build_from_neighbor_search(ps, ns, 1.5)
    assert np.array_equal(bonds.get_neighbor_count(), ns.get_all_neighbor_counts())
    assert bonds.get_broken_bonds().sum() == 0
    ps = _grid((2, 2, 2))
    ns = NeighborSearch((-1, -1, -1), (2, 2, 2), horizon=1.1, max_particles=8, dim=3)
    ns.build(ps.X, 8)
    bonds = BondSystem(8, max_bonds=30, dim=3)
    bonds.build_from_neighbor_search(ps, ns, 1.1)
    L = bonds.xi_length.to_numpy()
    nn = bonds.n_neighbors.to_numpy()
    for i in range(8):
        assert np.allclose(L[i, :nn[i]], 1.0)
    nb = bonds.neighbors.to_numpy()
    xi = bonds.xi.to_numpy()
    X = ps.X.to_numpy()
    for i in range(8):
        assert np.allclose(xi[i, :nn[i]], X[nb[i, :nn[i]]] - X[i])


@pytest.mark.gpu
def test_mirror_shape_tensor_and_identity_F(gpu):
    # test_3d.py:219-332
    ps = _grid((3, 3, 3))
    ps.set_material_constants(1e6 / 1.5, 1e6 / 2.5)
    ns = NeighborSearch((-1, -1, -1), (4, 4, 4), horizon=1.8, max_particles=27, max_neighbors=30, dim=3)
    ns.build(ps.X, 27)
    bonds = BondSystem(27, max_bonds=30, dim=3)
    bonds.build_from_neighbor_search(ps, ns, 1.8)
    nosb = NOSBCompute(ps, bonds)
    nosb.compute_shape_tensor()
    K = ps.K.to_numpy()
    assert all(np.allclose(K[i], K[i].T) for i in range(27))
    assert np.all(np.linalg.eigvalsh(K[13]) > 0)
    nosb.compute_deformation_gradient()
    assert np.allclose(ps.F.to_numpy()[13], np.eye(3), atol=1e-4)


@pytest.mark.gpu
def test_mirror_compression_and_quasi_static_stability(gpu):
    # test_3d.py:338-444
    spacing = 0.01
    horizon = 3.015 * spacing
    ps = _grid((2, 2, 2), spacing)
    ns = NeighborSearch((-horizon,) * 3, (2 * spacing + horizon,) * 3, horizon=horizon, max_particles=8,
                        max_neighbors=64, dim=3)
    ns.build(ps.X, 8)
    bonds = BondSystem(8, max_bonds=64, dim=3)
    bonds.build_from_neighbor_search(ps, ns, horizon)
    X = ps.X.to_numpy()
    bottom, top = np.where(X[:, 2] < spacing * 0.5)[0], np.where(X[:, 2] > spacing * 0.5)[0]
    ps.set_fixed_particles(bottom)
    x = ps.x.to_numpy()
    x[top, 2] += -0.0001
    ps.x.from_numpy(x)
    d = ps.get_displacements()
    assert np.allclose(d[bottom, 2], 0.0) and np.all(d[top, 2] < 0)
    E = 1e6
    c = 18 * (E / (3 * (1 - 2 * 0.25))) / (math.pi * horizon ** 4)
    solver = QuasiStaticSolver(ps, bonds, micromodulus=c)
    ps.set_fixed_particles(np.array([0, 1, 2, 3]))
    x = ps.x.to_numpy()
    x[4:, 2] += 0.0001
    ps.x.from_numpy(x)
    for _ in range(50):
        solver.step()
    disp = ps.get_displacements()
    assert np.max(np.abs(disp)) < 0.1 and not np.isnan(disp).any()


@pytest.mark.gpu
def test_mirror_nosb_solver_matches_oracle(gpu):
    """NOSBSolver with a LoadControl, adaptive DR, against the oracle's solve."""
    n_side = 8
    ps = _grid((n_side,) * 3)
    n = ps.n_particles
    X = ps.X.to_numpy()
    mat = NOSBMaterial(15000.0, 0.3, dim=3)
    ns = NeighborSearch((-4,) * 3, (12,) * 3, horizon=3.015, max_particles=n, max_neighbors=130, dim=3)
    ns.build(ps.X, n)
    bonds = BondSystem(n, max_bonds=130, dim=3)
    bonds.build_from_neighbor_search(ps, ns, 3.015)
    bottom, top = np.where(X[:, 2] < 0.5)[0], np.where(X[:, 2] > n_side - 1.5)[0]
    ps.set_fixed_particles(bottom)
    load = LoadControl(ps, top)
    load.set_load((0.0, 0.0, -2.0))
    solver = NOSBSolver(ps, bonds, mat, horizon=3.015, stabilization=0.1)
    r = solver.solve(external_force_func=load.apply, tol=1e-6, max_iterations=20000, verbose=False,
                     method="adaptive")
    assert r["converged"]
    fixed = np.zeros(n, np.uint8)
    fixed[bottom] = 7
    fext = np.zeros((n, 3))
    fext[top, 2] = -2.0
    prob = po.Problem(X, 1.0, 3.015, mat.K, mat.mu, G_s=0.1, cstab=mat.c_bond, fixed=fixed, fext=fext,
                      mass=ps.mass.to_numpy())
    prob.shape_tensor("absolute", 1e-12)
    ro = prob.adr_solve(prob.adr_mass(1.0), max_it=20000, tol=1e-6)
    assert ro["converged"]
    assert rel_l2(ps.get_displacements(), prob.u) < 1e-5


@pytest.mark.gpu
def test_framework_solver_config1_cube_matches_oracle(gpu):
    """Solver(Method.PD) on BASELINE config 1: 20^3 particles, dx 1 mm, E 15 GPa,
    nu 0.3, bottom layer fixed, top layer prescribed u_z = -0.5% x 19 mm."""
    d = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(19, 19, 19), n_divisions=(20, 20, 20))
    bottom, top = d.select(axis=2, value=0.0), d.select(axis=2, value=19.0)
    idx = np.concatenate([bottom, top])
    vals = np.zeros((len(idx), 3))
    vals[len(bottom):, 2] = -0.005 * 19.0
    d.set_fixed(idx, values=vals)
    mat = Material(E=15000.0, nu=0.3, density=1.9e-9, dim=3)
    s = Solver(d, mat, tol=1e-6, max_iterations=40000)
    res = s.solve()
    assert res.converged and res.relative_residual < 1e-6
    u = s.get_displacements()
    assert np.allclose(u[top, 2], -0.095) and np.allclose(u[bottom], 0.0)
    X = d.get_positions()
    n = len(X)
    fixed = np.zeros(n, np.uint8)
    fixed[idx] = 7
    prob = po.Problem(X, 1.0, 3.015, mat.K_bulk, mat.mu, G_s=0.1, cstab=18 * mat.K_bulk / (np.pi * 3.015 ** 4),
                      fixed=fixed, mass=np.full(n, mat.density))
    prob.shape_tensor("relative", 1e-12)
    u0 = np.zeros((n, 3))
    u0[top, 2] = -0.095
    prob.set_u(u0)
    ro = prob.adr_solve(prob.adr_mass(1.0), max_it=40000, tol=1e-6)
    assert ro["converged"]
    assert rel_l2(u, prob.u) < 1e-5, rel_l2(u, prob.u)
    P = s.get_stress()
    assert P.shape == (n, 3, 3) and np.isfinite(P).all()
    assert np.allclose(s.get_damage(), 0.0)
