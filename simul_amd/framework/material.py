"""Material (reference framework/material.py:18-150) for the peridynamics
method: E, nu, density and the constitutive model.  ``constitutive_model``
("linear_elastic", "neo_hookean", "j2_plasticity") is honoured by the PD
solver here (the reference's PD path ignores it, material.py:140-143)."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

from ..validation import (FEAValidationError, validate_density, validate_elastic_constants,
                          validate_hardening_modulus, validate_yield_stress)

PD_MODELS = ("linear_elastic", "neo_hookean", "j2_plasticity")


@dataclass
class Material:
    E: float
    nu: float
    density: float = 1000.0
    dim: int = 2
    plane_stress: bool = False
    constitutive_model: str = "linear_elastic"
    yield_stress: Optional[float] = None
    hardening_modulus: Optional[float] = None
    lam: float = field(init=False, repr=False)
    mu: float = field(init=False, repr=False)
    K_bulk: float = field(init=False, repr=False)

    def __post_init__(self):
        validate_elastic_constants(self.E, self.nu, "Material")
        validate_density(self.density, "Material")
        if self.constitutive_model == "j2_plasticity":
            if self.yield_stress is None:
                raise FEAValidationError("j2_plasticity needs yield_stress", parameter="yield_stress")
            validate_yield_stress(self.yield_stress, "Material")
            validate_hardening_modulus(self.hardening_modulus or 0.0, "Material")
        self.mu = self.E / (2.0 * (1.0 + self.nu))
        self.lam = self.E * self.nu / ((1.0 + self.nu) * (1.0 - 2.0 * self.nu))
        self.K_bulk = self.E / (3.0 * (1.0 - 2.0 * self.nu))

    def _create_nosb_material(self):
        from ..peridynamics.core.nosb import NOSBMaterial
        return NOSBMaterial(self.E, self.nu, dim=self.dim)
