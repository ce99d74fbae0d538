"""Explicit dynamics (ExplicitSolver, explicit.py:62-158) with the per-step
damage schedule (DamageModel.step, damage.py:144-152): the HIP path
(pdx_explicit_run through the C ABI) against the FP64 oracle's
orc_explicit_run on the same inputs.

Tolerances: trajectories (u, v) and kinetic energy within 1e-5 relative-L2
(FP32 bond arithmetic over a few hundred steps); with breaking, the broken
flags equal the oracle's except for bonds whose stretch crosses the threshold
within the FP32 trajectory difference (at most 0.2 % of the broken bonds; none
differed when measured), and every bond over the threshold at the final
positions is broken.
"""
import math

import numpy as np
import pytest

from helpers import lattice, rel_l2, smooth_u
from oracle import pdoracle as po
from simul_amd import workloads as W
from simul_amd.body import PDBody

pytestmark = pytest.mark.gpu

E, NU, RHO, H = 15000.0, 0.3, W.RHO_BONE, 3.015


def _pair(X, formulation, model="linear_elastic", fixed=None):
    n = len(X)
    bulk, shear = W.bulk_shear(np.full(n, E), NU)
    c = 18 * bulk / (math.pi * H ** 4)
    mass = np.full(n, RHO)
    body = PDBody(X, 1.0, H, bulk, shear, mass=mass, formulation=formulation, model=model, cstab=c, fixed=fixed)
    prob = po.Problem(X, 1.0, H, bulk, shear, mass=mass, formulation=formulation, model=model, cstab=c,
                      fixed=fixed)
    if formulation == "nosb":
        prob.shape_tensor("relative", 1e-12)
    return body, prob


def _dt():
    # ExplicitSolver.estimate_stable_dt (safety 0.5), halved
    return 0.5 * 0.5 * 1.0 / math.sqrt(E / RHO)


@pytest.mark.parametrize("formulation,model", [("bond", "linear_elastic"), ("nosb", "neo_hookean")])
def test_explicit_trajectory_matches_oracle(gpu, formulation, model):
    X = lattice(10, 9, 8, jitter=0.1, seed=31)
    fixed = np.where(X[:, 2] < 0.5, 7, 0).astype(np.uint8)
    body, prob = _pair(X, formulation, model, fixed)
    v0 = smooth_u(X, amp=1.0, seed=32) * 200.0 * (fixed == 0)[:, None]   # mm/s, strains ~1e-4
    body.set("v", v0)
    prob.v[...] = v0
    dt = _dt()
    for n_done, n in ((0, 1), (1, 150), (150, 300)):
        nb, ke = body.explicit_run(dt, 0.0, n - n_done)
        nbo, keo = prob.explicit_run(dt, 0.0, n - n_done)
        assert nb == nbo == 0
        assert rel_l2(body.u, prob.u) < 1e-5, (n, rel_l2(body.u, prob.u))
        assert rel_l2(body.get("v"), prob.v) < 1e-5, (n, rel_l2(body.get("v"), prob.v))
        assert abs(ke - keo) < 1e-5 * keo
    # viscous damping factor, continuing from the same state
    body.explicit_run(dt, 0.02, 40)
    prob.explicit_run(dt, 0.02, 40)
    assert rel_l2(body.get("v"), prob.v) < 1e-5


def test_explicit_damage_schedule_matches_oracle(gpu):
    X = lattice(16, 10, 6, jitter=0.1, seed=41)
    body, prob = _pair(X, "bond")
    xc = X.mean(0)
    rate = 900.0  # 1/s: uniform extension along x; about 15 % of the bonds break
    v0 = np.zeros_like(X)
    v0[:, 0] = rate * (X[:, 0] - xc[0])
    v0 += smooth_u(X, amp=1.0, seed=42) * 100.0
    body.set("v", v0)
    prob.v[...] = v0
    dt, s_c = _dt(), 3e-3
    nb_g = nb_o = 0
    for chunk in (60, 60, 60):
        g, _ = body.explicit_run(dt, 0.0, chunk, critical_stretch=s_c)
        o, _ = prob.explicit_run(dt, 0.0, chunk, critical_stretch=s_c)
        nb_g += g
        nb_o += o
    assert nb_o > 5000
    _, _, br = body.bonds()
    diff = int(np.count_nonzero(br != prob.broken[: prob.nb]))
    print(f"explicit damage: GPU {nb_g}, oracle {nb_o} broken, {diff} flags differ", flush=True)
    assert diff <= max(4, 2e-3 * nb_o)
    assert abs(nb_g - nb_o) <= max(4, 2e-3 * nb_o)
    assert np.array_equal(br.astype(bool).sum(), nb_g)
    assert rel_l2(body.u, prob.u) < 1e-4
    # local damage is current at return and consistent with the flags
    row = prob.row
    i = np.repeat(np.arange(len(X)), np.diff(row))
    phi = np.bincount(i, weights=br.astype(float), minlength=len(X)) / np.diff(row)
    assert np.array_equal(body.get("damage"), phi)
    # breaking is irreversible and tested after every step: every bond over the
    # threshold at the final positions (the oracle's test at the GPU's state)
    # is among the broken ones
    q = po.Problem(X, 1.0, H, 1.0, 1.0, formulation="bond")
    q.x[...] = body.x
    assert q.update_damage(s_c) > 0
    assert not np.any(q.broken[: q.nb].astype(bool) & ~br.astype(bool))


def test_explicit_solver_mirror(gpu):
    """The reference's ExplicitSolver API: step, run with callbacks at multiples
    of the interval, the CFL estimate and the energies (explicit.py:160-259)."""
    from simul_amd.peridynamics import BondSystem, DamageModel, ExplicitSolver, NeighborSearch, ParticleSystem
    spacing = 0.01
    horizon = 3.015 * spacing
    ps = ParticleSystem(9, dim=2)
    ps.initialize_from_grid((0, 0), spacing, (3, 3), density=1000.0)
    ns = NeighborSearch(domain_min=(-0.05, -0.05), domain_max=(0.05, 0.05), horizon=horizon, max_particles=9, dim=2)
    ns.build(ps.X, 9)
    bonds = BondSystem(9, max_bonds=32, dim=2)
    bonds.build_from_neighbor_search(ps, ns, horizon)
    E2 = 1e6
    c = 9 * E2 / (math.pi * 1.0 * horizon ** 3)
    dt = ExplicitSolver.estimate_stable_dt(E2, 1000.0, horizon, spacing, 0.01)
    assert abs(dt - 0.01 * spacing / math.sqrt(E2 / 1000.0)) < 1e-18
    assert ExplicitSolver.estimate_stable_dt(E2, 1000.0, horizon, spacing / 2, 0.5) < \
        ExplicitSolver.estimate_stable_dt(E2, 1000.0, horizon, spacing, 0.5)
    solver = ExplicitSolver(ps, bonds, micromodulus=c, dt=dt, damping=0.0)
    v = ps.v.to_numpy()
    v[4, 0] = 0.0001
    ps.v.from_numpy(v)
    for _ in range(10):
        solver.step()
    ke, se = solver.get_kinetic_energy(), solver.get_strain_energy()
    assert np.isfinite(ke) and np.isfinite(se) and ke < 1e3 and se < 1e3
    assert ke > 0 and se > 0
    # a = f / m with f a force density: the conserved energy is V KE + SE
    # (uniform V), to the Verlet error (omega dt ~ 0.05 at a tenth of that dt)
    V = spacing ** 2
    e0 = V * ke + se
    seen = []
    solver.dt = dt / 10
    solver.run(250, callback=lambda s, i: seen.append(i), callback_interval=100)
    assert seen == [0, 100, 200] and solver.step_count == 260
    e1 = V * solver.get_kinetic_energy() + solver.get_strain_energy()
    assert abs(e1 - e0) < 1e-4 * e0, (e0, e1)
    # the oracle from the same start gives the same trajectory
    X = ps.X.to_numpy()
    p = po.Problem(X, spacing ** 2, horizon, 1.0, 1.0, formulation="bond", cstab=c,
                   mass=np.full(9, 1000.0 * spacing ** 2))
    p.v[4, 0] = 0.0001
    p.explicit_run(dt, 0.0, 10)
    p.explicit_run(dt / 10, 0.0, 250)
    assert rel_l2(ps.get_displacements(), p.u) < 1e-5
    # a damage model breaks bonds after every step
    dm = DamageModel(critical_stretch=1e-12, dim=2)
    solver.run(3, damage_model=dm)
    assert dm.get_statistics(bonds)["broken_bonds"] > 0
