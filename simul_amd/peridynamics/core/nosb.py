"""NOSBCompute and NOSBMaterial (reference core/nosb.py:28-352).

NOSBCompute drives the libpdx passes of the bound context: the FP64 shape
tensor with the reference's absolute singular test |det K| > 1e-12
(nosb.py:89-94), the F / stress pass and the force pass with stabilisation.
Material constants are the particle system's per-particle bulk / shear moduli
(particles.py:52-53), as in the reference.
"""
from __future__ import annotations

import math

from ...validation import validate_elastic_constants


class NOSBCompute:
    def __init__(self, particles, bonds, stabilization: float = 0.0, model: str = "linear_elastic",
                 singular_test: str = "absolute", singular_tol: float = 1e-12):
        self.particles = particles
        self.bonds = bonds
        self.dim = particles.dim
        self.n_particles = particles.n_particles
        self.stabilization = float(stabilization)
        self.model = model
        self.singular_test, self.singular_tol = singular_test, singular_tol
        self._bond_constant = None

    @property
    def body(self):
        return self.bonds.body

    def _sync_materials(self):
        ps = self.particles
        bulk, shear = ps.bulk_mod._host, ps.shear_mod._host
        if (bulk > 0).all() and (shear > 0).all():
            self.body.set_materials(self.model, bulk, shear)

    def compute_shape_tensor(self):
        """K = sum_intact w xi (x) xi V_j and its FP64 inverse (nosb.py:59-94)."""
        self.body.set_formulation("nosb")
        self.body.shape_tensor(self.singular_test, self.singular_tol)

    def compute_deformation_gradient(self):
        """F (and the stress of the particles' model) at the current x (nosb.py:96-126)."""
        self._sync_materials()
        self.body.compute_F_stress()

    def compute_force_state_with_stabilization(self, bond_constant: float):
        """f_i = sum_j V_j [w (P_i K_i^-1 + P_j K_j^-1) xi + G_s c s w eta/|eta|]
        (nosb.py:193-275), written into the particles' f."""
        self._sync_materials()
        if self._bond_constant != bond_constant:
            self.body.set_stabilization(self.stabilization, bond_constant)
            self._bond_constant = bond_constant
        self.body.compute_forces()

    def compute_force_state_linear_elastic(self, bond_constant: float = 0.0):
        """Linear-elastic correspondence force (nosb.py:129-191): the force pass
        with the stabilisation weight the system was made with."""
        self.compute_force_state_with_stabilization(bond_constant)

    def apply_force_antisymmetry(self):
        """No-op, as in the reference (nosb.py:278-288)."""


class NOSBMaterial:
    def __init__(self, youngs_modulus: float, poisson_ratio: float, dim: int = 3):
        validate_elastic_constants(youngs_modulus, poisson_ratio, "NOSBMaterial")
        self.E = youngs_modulus
        self.nu = poisson_ratio
        self.dim = dim
        # nosb.py:314-320: 2-D plane-stress bulk modulus, 3-D bulk modulus
        self.K = youngs_modulus / (2 * (1 - poisson_ratio)) if dim == 2 else \
            youngs_modulus / (3 * (1 - 2 * poisson_ratio))
        self.mu = youngs_modulus / (2 * (1 + poisson_ratio))
        self.c_bond = None

    def get_bulk_modulus(self) -> float:
        return self.K

    def get_shear_modulus(self) -> float:
        return self.mu

    def set_stabilization_modulus(self, horizon: float, thickness: float = 1.0):
        """c = 9E/(pi t delta^3) in 2-D, 18K/(pi delta^4) in 3-D (nosb.py:335-346)."""
        if self.dim == 2:
            self.c_bond = 9 * self.E / (math.pi * thickness * horizon ** 3)
        else:
            self.c_bond = 18 * self.K / (math.pi * horizon ** 4)

    def __repr__(self) -> str:
        return f"NOSBMaterial(E={self.E:.2e}, nu={self.nu:.3f}, K={self.K:.2e}, mu={self.mu:.2e})"
