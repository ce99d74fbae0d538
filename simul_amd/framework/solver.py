"""Solver (reference framework/solver.py:15-67) and its peridynamics adapter
(reference framework/_adapters/pd_adapter.py:17-216), on libpdx.

Adapter set-up follows the reference: spacing from the domain's divisions,
horizon = horizon_factor x spacing (3.015), V = spacing^dim, the domain's
particle order, moduli from the Material.  Options:

- ``formulation``: "nosb" (correspondence + stabilisation, default) or
  "bond" (the PMB force the reference's Solver(Method.PD) iterates,
  solver/quasi_static.py:119-143, with c = 18K/(pi delta^4)).
- ``dr``: "adaptive" (default) or "kinetic" (reference kinetic damping with
  dt from the spectral estimate and ``damping`` as the viscous factor).
- ``stabilization`` (G_s, 0.1), ``damping`` (0.1), ``max_iterations`` (50000),
  ``tol`` (1e-4), ``horizon_factor`` (3.015), ``singular_test`` ("relative").

Boundary conditions: ``set_fixed(indices, values, dofs)`` prescribes the given
displacement components (the reference fixes all DOFs and ignores values);
``set_force`` loads a force density on the indices (the reference's
LoadControl adds one vector, the first row for a per-index array; here every
row is used).
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from ..body import PDBody
from .domain import Domain, Method
from .material import PD_MODELS, Material
from .result import SolveResult


class Solver:
    def __init__(self, domain: Domain, material: Material, **options):
        self.domain = domain
        self.material = material
        self._options = options
        if domain.method == Method.PD:
            self._adapter = PDAdapter(domain, material, **options)
        elif domain.method in (Method.FEM, Method.SPG, Method.COUPLED, Method.RIGID):
            raise NotImplementedError(f"{domain.method} is outside this library (peridynamics only)")
        else:
            raise ValueError(f"unsupported analysis method: {domain.method}")
        domain._adapter = self._adapter

    def solve(self, **kwargs) -> SolveResult:
        return self._adapter.solve(**kwargs)

    def get_displacements(self) -> np.ndarray:
        return self._adapter.get_displacements()

    def get_stress(self) -> np.ndarray:
        return self._adapter.get_stress()

    def get_damage(self) -> Optional[np.ndarray]:
        return self._adapter.get_damage()


class PDAdapter:
    def __init__(self, domain: Domain, material: Material, **options):
        dim = domain.dim
        pos = domain.get_positions()
        n = len(pos)
        n0 = domain.n_divisions[0]
        spacing = domain.size[0] / (n0 - 1) if n0 > 1 else domain.size[0]
        horizon = options.get("horizon_factor", 3.015) * spacing
        vol = np.full(n, spacing ** dim)
        model = material.constitutive_model if material.constitutive_model in PD_MODELS else "linear_elastic"
        if dim == 2 and model == "j2_plasticity":
            raise NotImplementedError("J2 plasticity in the PD solver is 3-D only")
        self.formulation = options.get("formulation", "nosb")
        self.dr = options.get("dr", "adaptive")
        if self.dr not in ("adaptive", "kinetic"):
            raise ValueError(f"dr must be 'adaptive' or 'kinetic', got {self.dr!r}")
        bulk = material.K_bulk if dim == 3 else material.E / (2 * (1 - material.nu))
        c = 18.0 * material.K_bulk / (np.pi * horizon ** 4) if dim == 3 else \
            9.0 * material.E / (np.pi * horizon ** 3)
        j2 = {}
        if model == "j2_plasticity":
            j2 = dict(yield_stress=material.yield_stress, h_iso=material.hardening_modulus or 0.0)
        self.n, self.dim, self.horizon, self.spacing = n, dim, horizon, spacing
        self.fixed = np.zeros(n, np.uint8)
        u0 = np.zeros((n, dim))
        if domain._fixed_indices is not None:
            dofs = domain._fixed_dofs if domain._fixed_dofs is not None else range(dim)
            mask = 0
            for q in dofs:
                mask |= 1 << int(q)
            self.fixed[domain._fixed_indices] |= np.uint8(mask)
            if domain._fixed_values is not None:
                vals = np.broadcast_to(np.asarray(domain._fixed_values, np.float64),
                                       (len(domain._fixed_indices), dim))
                for q in dofs:
                    u0[domain._fixed_indices, q] = vals[:, q]
        self.fext = np.zeros((n, dim))
        if domain._force_indices is not None:
            fv = np.asarray(domain._force_values, np.float64)
            self.fext[domain._force_indices] = np.broadcast_to(fv, (len(domain._force_indices), dim))
        self._contact = np.zeros((n, dim))
        self.body = PDBody(pos, vol, horizon, bulk, material.mu, mass=vol * material.density, model=model,
                           formulation=self.formulation, G_s=options.get("stabilization", 0.1), cstab=c,
                           fixed=self.fixed, fext=self.fext, u0=u0,
                           singular_test=options.get("singular_test", "relative"), **j2)
        self.rho = material.density
        self.material = material
        self._options = options
        self._dt = None

    def _push_loads(self):
        self.body.set("fext", self.fext + self._contact)

    def solve(self, **kwargs) -> SolveResult:
        max_it = kwargs.get("max_iterations", self._options.get("max_iterations", 50000))
        tol = kwargs.get("tol", self._options.get("tol", 1e-4))
        t0 = time.time()
        self._push_loads()
        if self.dr == "kinetic":
            r = self.body.solve("kinetic", tol=tol, max_iterations=max_it, dt=self.get_stable_dt(),
                                viscous_damping=self._options.get("damping", 0.1))
        else:
            r = self.body.solve("adaptive", tol=tol, max_iterations=max_it)
        return SolveResult(converged=r.converged, iterations=r.iterations, residual=r.residual,
                           relative_residual=r.relative_residual, elapsed_time=time.time() - t0)

    def get_displacements(self) -> np.ndarray:
        return self.body.u

    def get_stress(self) -> np.ndarray:
        """First Piola-Kirchhoff stress P (n, dim, dim) at the current state."""
        if self.formulation != "nosb":
            raise NotImplementedError("stress is defined for the correspondence (nosb) formulation")
        self.body.compute_F_stress()
        return self.body.get("P")

    def get_damage(self) -> Optional[np.ndarray]:
        return self.body.local_damage()

    # -- contact / scene hooks (pd_adapter.py:178-216) ------------------------
    def get_current_positions(self) -> np.ndarray:
        return self.body.x

    def get_reference_positions(self) -> np.ndarray:
        return self.body.get("X")

    def inject_contact_forces(self, indices: np.ndarray, forces: np.ndarray):
        np.add.at(self._contact, np.asarray(indices, np.int64), np.asarray(forces, np.float64))

    def clear_contact_forces(self):
        self._contact[:] = 0.0

    def step(self, dt: float = None):
        """One explicit kinetic-damping step with the stable dt (pd_adapter.py:208-210)."""
        self._push_loads()
        self.body.solve("kinetic", tol=0.0, max_iterations=1, dt=self.get_stable_dt(),
                        viscous_damping=self._options.get("damping", 0.1), use_graph=False, resume=True)

    def get_stable_dt(self) -> float:
        if self._dt is None:
            self._dt = self.body.estimate_dt(self.material.K_bulk, self.material.mu, self.rho)
        return float(self._dt)
