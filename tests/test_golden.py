"""Committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (FP64, 1e-12 relative), and the
closed-form fixtures agree with the oracle.
GPU: the HIP path meets the same fixtures through the C ABI -- bit-exact for
neighbour sets, bond data and broken flags, 1e-12 for FP64 K / K^-1, and the
FP32 field bounds of the parity tests (relative L2) for F, P and f.
"""
import os

import numpy as np
import pytest

from helpers import rel_l2
from oracle import pdoracle as po

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELD_CASES = {"lattice_le": ("linear_elastic", "nosb"), "lattice_nh": ("neo_hookean", "nosb"),
               "lattice_bond": ("linear_elastic", "bond")}


def load(name):
    with np.load(os.path.join(HERE, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def close(a, b, rtol=1e-12):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.allclose(a, b, rtol=rtol, atol=rtol * max(np.abs(b).max(), 1e-300))


def _oracle(g, model, form):
    p = po.Problem(g["X"], g["vol"], float(g["horizon"]), g["bulk"], g["shear"], model=model, formulation=form,
                   G_s=float(g["G_s"]), cstab=g["cstab"], fixed=g["fixed"], fext=g["fext"], nthreads=1)
    if form == "nosb":
        p.shape_tensor("relative", 1e-12)
    p.set_u(g["u"])
    p.compute_forces()
    return p


@pytest.mark.parametrize("name", sorted(FIELD_CASES))
def test_oracle_reproduces_field_fixture(name):
    g = load(name)
    p = _oracle(g, *FIELD_CASES[name])
    assert np.array_equal(p.row, g["row"]) and np.array_equal(p.col[: p.nb], g["col"])
    assert close(p.xlen[: p.nb], g["xlen"]) and close(p.omega[: p.nb], g["omega"])
    assert close(p.f, g["f"])
    assert close(p.residual_norm(), g["residual"])
    if "K" in g:
        for k in ("K", "Kinv", "F", "P"):
            assert close(getattr(p, k), g[k]), k


def test_oracle_reproduces_j2_fixture():
    g = load("j2")
    p = po.Problem(g["X"], g["vol"], float(g["horizon"]), g["bulk"], g["shear"], model="j2_plasticity",
                   sy0=g["yield_stress"], h_iso=g["h_iso"], nthreads=1)
    p.shape_tensor("relative", 1e-12)
    p.set_u(g["u"])
    p.deformation_gradient()
    p.stress()
    assert close(p.F, g["F"]) and close(p.P, g["P"])
    p.commit()
    assert close(p.ep, g["ep"]) and close(p.epbar, g["epbar"])
    assert g["epbar"].max() > 0  # the fixture really yields


def test_oracle_reproduces_damage_fixture():
    g = load("damage")
    p = po.Problem(g["X"], g["vol"], float(g["horizon"]), g["bulk"], g["shear"], nthreads=1)
    p.set_u(g["u"])
    assert p.update_damage(float(g["s_c"])) == int(g["n_broken"]) > 0
    assert np.array_equal(p.broken[: p.nb], g["broken"])
    assert close(p.local_damage(), g["damage"])


def test_closed_form_fixture_against_oracle():
    g = load("closed_form")
    X, H = g["X"], float(g["horizon"])
    p = po.Problem(X, 1.0, H, 1.0, 1.0, nthreads=1)
    p.shape_tensor("relative", 1e-12)
    centre = int(np.argmin(np.linalg.norm(X - X.mean(0), axis=1)))
    assert np.diff(p.row)[centre] == int(g["n_interior_bonds"])
    assert np.allclose(p.K[centre], float(g["k_interior"]) * np.eye(3), rtol=1e-13, atol=1e-12)
    G = g["G"]
    p.set_u(X @ G.T)
    p.deformation_gradient()
    assert np.abs(p.F - (np.eye(3) + G)[None]).max() < 1e-13


# ---------------------------------------------------------------------------
# the HIP path against the same fixtures
# ---------------------------------------------------------------------------
def _body(g, model, form, **kw):
    from simul_amd.body import PDBody
    return PDBody(g["X"], g["vol"], float(g["horizon"]), g["bulk"], g["shear"], model=model, formulation=form,
                  G_s=float(g.get("G_s", 0.1)), cstab=g.get("cstab"), fixed=g.get("fixed"), fext=g.get("fext"),
                  **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FIELD_CASES))
def test_hip_meets_field_fixture(gpu, name):
    g = load(name)
    model, form = FIELD_CASES[name]
    b = _body(g, model, form)
    row, col, br = b.bonds()
    assert np.array_equal(row, g["row"]) and np.array_equal(col, g["col"]) and not br.any()
    _, L, w = b.bond_data()
    assert np.array_equal(L, g["xlen"]) and np.array_equal(w, g["omega"])
    b.u = g["u"]
    f = b.compute_forces()
    assert rel_l2(f, g["f"]) < 5e-5, rel_l2(f, g["f"])
    if form == "nosb":
        assert rel_l2(b.get("K"), g["K"]) < 1e-12 and rel_l2(b.get("K_inv"), g["Kinv"]) < 1e-12
        I = np.eye(3)[None]
        assert rel_l2(b.get("F") - I, g["F"] - I) < 2e-5
        assert rel_l2(b.get("P"), g["P"]) < 2e-5


@pytest.mark.gpu
def test_hip_meets_j2_fixture(gpu):
    g = load("j2")
    b = _body(g, "j2_plasticity", "nosb", yield_stress=g["yield_stress"], h_iso=g["h_iso"])
    b.u = g["u"]
    b.compute_F_stress()
    assert rel_l2(b.get("P"), g["P"]) < 2e-5
    b.commit_state()
    assert rel_l2(b.get("epbar"), g["epbar"]) < 1e-4
    assert rel_l2(b.get("ep").reshape(-1, 9), g["ep"]) < 1e-4


@pytest.mark.gpu
def test_hip_meets_damage_fixture(gpu):
    g = load("damage")
    b = _body(g, "linear_elastic", "nosb")
    b.u = g["u"]
    assert b.break_bonds(float(g["s_c"])) == int(g["n_broken"])
    _, _, br = b.bonds()
    assert np.array_equal(br, g["broken"])
    assert np.array_equal(b.local_damage(), g["damage"])


@pytest.mark.gpu
def test_hip_meets_closed_form_fixture(gpu):
    g = load("closed_form")
    from simul_amd.body import PDBody
    X = g["X"]
    b = PDBody(X, 1.0, float(g["horizon"]), 1.0, 1.0)
    centre = int(np.argmin(np.linalg.norm(X - X.mean(0), axis=1)))
    assert b.get("nnb")[centre] == int(g["n_interior_bonds"])
    assert np.allclose(b.get("K")[centre], float(g["k_interior"]) * np.eye(3), rtol=1e-13, atol=1e-12)
    b.u = X @ g["G"].T
    b.compute_F_stress()
    err = np.abs(b.get("F") - (np.eye(3) + g["G"])[None]).max() / np.abs(g["G"]).max()
    assert err < 2e-5
