"""The callers' side of the drop-in: the orchestrator's result.npz and the
API's region dict (simul_amd.results), the scene/coupling hooks of the PD
adapter (pd_adapter.py:178-216, coupled_solver.py:231-233,466-504) and
Material.constitutive_model through Solver, against the oracle where there is
physics to compare."""
import json

import numpy as np
import pytest

from helpers import rel_l2
from oracle import pdoracle as po
from simul_amd import results as R
from simul_amd.framework import Material, Method, Solver
from simul_amd.framework.domain import create_domain, create_particle_domain
from simul_amd.framework.result import SolveResult


class _FakeSolver:
    def __init__(self, n):
        rng = np.random.default_rng(0)
        self.u = rng.normal(size=(n, 3))
        self.P = rng.normal(size=(n, 3, 3))
        self.d = rng.uniform(size=n)

    def get_displacements(self):
        return self.u

    def get_stress(self):
        return self.P

    def get_damage(self):
        return self.d


def test_result_npz_keys_and_round_trip(tmp_path):
    # solve.py:114-126: keys, one-element converged / iterations / residual
    n = 17
    pos = np.random.default_rng(1).normal(size=(n, 3))
    s = _FakeSolver(n)
    res = SolveResult(converged=True, iterations=123, residual=4.5e-7, relative_residual=1e-6, elapsed_time=0.1)
    p = R.save_result_npz(tmp_path / "result.npz", pos, s, res)
    z = R.load_result_npz(p)
    assert set(z) == set(R.RESULT_KEYS) | {"stress", "damage"}
    assert np.array_equal(z["positions"], pos) and np.array_equal(z["displacements"], s.u)
    assert z["converged"].tolist() == [True] and z["iterations"].tolist() == [123]
    assert z["residual"].tolist() == [4.5e-7]
    assert np.array_equal(z["stress"], s.P) and np.array_equal(z["damage"], s.d)
    # without stress and damage those keys are absent (solve.py:121-124)
    a = R.result_arrays(pos, None, res)
    assert set(a) == set(R.RESULT_KEYS) and a["displacements"].size == 0


def test_scalar_stress_and_region_dict():
    # analysis.py:358-364: norm of the last axis, twice for a tensor = Frobenius norm
    P = np.array([[[3.0, 0, 0], [0, 4.0, 0], [0, 0, 12.0]], [[1.0, 2.0, 2.0], [0, 0, 0], [0, 0, 0]]])
    assert np.allclose(R.scalar_stress(P), [13.0, 3.0])
    assert np.allclose(R.scalar_stress(np.array([[3.0, 4.0]])), [5.0])
    assert R.scalar_stress(None) is None
    s = _FakeSolver(5)
    res = SolveResult(converged=False, iterations=9, residual=1.0, relative_residual=1.0, elapsed_time=0.0)
    d = R.region_result("disc", np.zeros((5, 3)), s, res)
    assert list(d) == ["name", "converged", "iterations", "displacements", "stress", "damage", "positions"]
    assert d["name"] == "disc" and d["converged"] is False and d["iterations"] == 9
    assert np.allclose(d["stress"], np.sqrt((s.P ** 2).sum(axis=(1, 2))))
    json.dumps(d)  # plain lists: serialisable as the API sends it


def test_solve_stage_refuses_other_methods_and_bad_input(tmp_path):
    st = R.SolveStage(method="spg")
    bad = st.run(tmp_path / "missing.npz", tmp_path)
    assert not bad.success and "invalid input" in bad.message
    np.savez(tmp_path / "in.npz", positions=np.zeros((8, 3)))
    r = st.run(tmp_path / "in.npz", tmp_path / "out")
    assert not r.success and "peridynamics only" in r.message


@pytest.mark.gpu
def test_solve_stage_writes_the_solver_result(gpu, tmp_path):
    """The orchestrator's stage on a voxel block (mm-MPa units): bottom layer
    fixed and no load, as the stage sets it up (solve.py:88-100), so the field is
    zero; then a loaded region through the API's dict (analysis.py:285-376)."""
    g = (np.arange(8) + 0.5) * 1.0
    pos = np.stack(np.meshgrid(g, g, g[:6], indexing="ij"), -1).reshape(-1, 3)
    np.savez(tmp_path / "model.npz", positions=pos)
    st = R.SolveStage(method="pd", E=15000.0, nu=0.3, density=1.9e-9, max_iterations=20000)
    r = st.run(tmp_path / "model.npz", tmp_path / "out")
    assert r.success, r.message
    z = R.load_result_npz(r.output_path)
    assert set(z) == set(R.RESULT_KEYS) | {"stress", "damage"}
    assert np.array_equal(z["positions"], pos)
    assert z["displacements"].shape == pos.shape and z["stress"].shape == (len(pos), 3, 3)
    assert z["converged"].dtype == bool
    # a loaded region, solved through the framework and returned as the API's dict
    d = create_particle_domain(pos, method=Method.PD)
    d.set_fixed(d.select(axis=2, value=pos[:, 2].min()))
    d.set_force(d.select(axis=2, value=pos[:, 2].max()), [0.0, 0.0, -2.0])
    s = Solver(d, Material(E=15000.0, nu=0.3, density=1.9e-9, dim=3), tol=1e-6, max_iterations=40000)
    res = s.solve()
    out = R.region_result("vertebra", pos, s, res)
    assert out["converged"] and np.allclose(out["displacements"], s.get_displacements())
    assert np.allclose(out["stress"], R.scalar_stress(s.get_stress()))
    assert min(u[2] for u in out["displacements"]) < 0


@pytest.mark.gpu
def test_neo_hookean_through_solver_matches_oracle(gpu):
    """Material(constitutive_model="neo_hookean") reaches the PD solve (the
    reference's PD path ignores it, material.py:140-143): large compression of a
    block, against the oracle's Neo-Hookean solve of the same input, and
    distinct from the linear-elastic answer."""
    d = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(9, 9, 9), n_divisions=(10, 10, 10))
    bottom, top = d.select(axis=2, value=0.0), d.select(axis=2, value=9.0)
    idx = np.concatenate([bottom, top])
    vals = np.zeros((len(idx), 3))
    vals[len(bottom):, 2] = -0.08 * 9.0
    d.set_fixed(idx, values=vals)
    mat = Material(E=1000.0, nu=0.3, density=1e-9, dim=3, constitutive_model="neo_hookean")
    s = Solver(d, mat, tol=1e-6, max_iterations=40000)
    res = s.solve()
    assert res.converged
    X = d.get_positions()
    n = len(X)
    fixed = np.zeros(n, np.uint8)
    fixed[idx] = 7
    prob = po.Problem(X, 1.0, 3.015, mat.K_bulk, mat.mu, model="neo_hookean", G_s=0.1,
                      cstab=18 * mat.K_bulk / (np.pi * 3.015 ** 4), fixed=fixed, mass=np.full(n, mat.density))
    prob.shape_tensor("relative", 1e-12)
    u0 = np.zeros((n, 3))
    u0[top, 2] = -0.72
    prob.set_u(u0)
    ro = prob.adr_solve(prob.adr_mass(1.0), max_it=40000, tol=1e-6)
    assert ro["converged"]
    assert rel_l2(s.get_displacements(), prob.u) < 1e-5
    le = Solver(d, Material(E=1000.0, nu=0.3, density=1e-9, dim=3), tol=1e-6, max_iterations=40000)
    le.solve()
    assert rel_l2(le.get_displacements(), prob.u) > 1e-3


@pytest.mark.gpu
def test_scene_and_coupling_hooks(gpu):
    """pd_adapter.py:178-216: contact forces injected into the force density
    act like a Neumann load and are cleared again; step() advances one explicit
    kinetic-damping step with get_stable_dt(); the coupling code's direct
    position access (coupled_solver.py:231-233) reads X + u."""
    d = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(7, 7, 7), n_divisions=(8, 8, 8))
    bottom, top = d.select(axis=2, value=0.0), d.select(axis=2, value=7.0)
    d.set_fixed(bottom)
    mat = Material(E=15000.0, nu=0.3, density=1.9e-9, dim=3)
    s = Solver(d, mat, tol=1e-7, max_iterations=40000)
    ad = s._adapter
    X = ad.get_reference_positions()
    assert np.array_equal(X, d.get_positions())
    # injected contact forces == the same Neumann load
    f = np.tile([0.5, 0.0, -3.0], (len(top), 1))
    ad.inject_contact_forces(top, f)
    assert s.solve().converged
    u_contact = s.get_displacements().copy()
    assert np.allclose(ad.get_current_positions(), X + u_contact)
    d2 = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(7, 7, 7), n_divisions=(8, 8, 8))
    d2.set_fixed(d2.select(axis=2, value=0.0))
    d2.set_force(top, [0.5, 0.0, -3.0])
    s2 = Solver(d2, mat, tol=1e-7, max_iterations=40000)
    assert s2.solve().converged
    assert rel_l2(u_contact, s2.get_displacements()) < 1e-6
    # cleared: back to the unloaded body (zero field from the loaded state)
    ad.clear_contact_forces()
    s.solve()
    assert np.abs(s.get_displacements()).max() < 1e-5 * np.abs(u_contact).max()
    # step(): one explicit step moves a loaded body by a dt^2-sized amount
    ad.inject_contact_forces(top, f)
    dt = ad.get_stable_dt()
    assert dt > 0 and abs(dt - ad.body.estimate_dt(mat.K_bulk, mat.mu, mat.density)) < 1e-15
    u_before = s.get_displacements().copy()
    ad.step(dt)
    du = s.get_displacements() - u_before
    assert np.abs(du[bottom]).max() == 0.0 and np.abs(du[top]).max() > 0
