"""NOSBSolver (reference solver/nosb_solver.py:23-332).

Correspondence (NOSB) forces with stabilisation, per-particle moduli, the
reference's spectral dt estimate and its kinetic-damping iteration, run on the
device.  ``solve(..., method="adaptive")`` selects the adaptive dynamic
relaxation of libpdx instead (no reference counterpart; same equilibrium).
"""
from __future__ import annotations

import numpy as np

from ..core.nosb import NOSBCompute
from .quasi_static import _KineticLoop, device_load


class NOSBSolver(_KineticLoop):
    def __init__(self, particles, bonds, material, horizon: float, stabilization: float = 0.1, dt: float = None,
                 viscous_damping: float = 0.0, model: str = "linear_elastic"):
        self.particles = particles
        self.bonds = bonds
        self.material = material
        self.dim = particles.dim
        self.n_particles = particles.n_particles
        self.horizon = float(horizon)
        self.nosb = NOSBCompute(particles, bonds, stabilization, model=model)
        self.K = material.get_bulk_modulus()
        self.mu = material.get_shear_modulus()
        particles.set_material_constants(self.K, self.mu)
        material.set_stabilization_modulus(self.horizon)
        self.c_bond = material.c_bond if material.c_bond else 0.0
        body = bonds.body
        body.set_stabilization(stabilization, self.c_bond)
        self.nosb._bond_constant = self.c_bond
        self.nosb.compute_shape_tensor()
        self.nosb._sync_materials()
        self.dt = float(dt) if dt is not None else self._estimate_stable_dt()
        self.viscous_damping = float(viscous_damping)
        self.prev_ke = 0.0
        self.ke_increasing = True
        self.iteration = 0

    def _estimate_stable_dt(self) -> float:
        """0.5 x 2/sqrt(lambda_max) of the per-particle stiffness/mass estimate (nosb_solver.py:94-154)."""
        rho = float(np.mean(self.particles.density.to_numpy()))
        return self.particles.body.estimate_dt(self.K, self.mu, rho)

    def compute_forces(self, external_force_func=None):
        """F, P and the internal force, then the external-force callback on ps.f (nosb_solver.py:200-216)."""
        self.nosb.compute_deformation_gradient()
        self.nosb.compute_force_state_with_stabilization(self.c_bond)
        if external_force_func is not None:
            external_force_func()

    def step(self, external_force_func=None) -> dict:
        return self._step(external_force_func, self.viscous_damping)

    def update_shape_tensor(self):
        self.nosb.compute_shape_tensor()

    def solve(self, external_force_func=None, max_iterations: int = 100000, tol: float = 1e-6,
              verbose: bool = True, print_interval: int = 1000, method: str = "kinetic") -> dict:
        """Iterates to |f_free| / |f_start| < tol (nosb_solver.py:262-332): velocities
        reset, accelerations kept from before, as in the reference."""
        self.nosb._sync_materials()
        self.prev_ke, self.ke_increasing = 0.0, True
        if method == "kinetic":
            r = self._run(external_force_func, max_iterations, tol, self.viscous_damping, verbose)
        elif method == "adaptive":
            body = self.particles.body
            fext = device_load(external_force_func)
            body.set("fext", np.zeros((self.n_particles, self.dim)) if fext is None else fext)
            try:
                r = body.solve("adaptive", tol=tol, max_iterations=max_iterations)
            finally:
                body.set("fext", np.zeros((self.n_particles, self.dim)))
            self.iteration = r.iterations
        else:
            raise ValueError(f"unknown method {method!r} (kinetic or adaptive)")
        return {"converged": r.converged, "iterations": r.iterations, "residual": r.residual,
                "relative_residual": r.relative_residual}
