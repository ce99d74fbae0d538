"""Input validation and the error types of the reference API.

Mirrors backend/fea/fem/validation.py:18-111,219-238 (FEAValidationError is a
ValueError carrying parameter/value/suggestion).
"""
from __future__ import annotations

import logging

logger = logging.getLogger("simul_amd")


class FEAValidationError(ValueError):
    def __init__(self, message: str, parameter: str = "", value=None, suggestion: str = ""):
        self.parameter = parameter
        self.value = value
        self.suggestion = suggestion
        full = f"[FEA validation] {message}"
        if suggestion:
            full += f" -> suggestion: {suggestion}"
        super().__init__(full)


class FEAConvergenceError(RuntimeError):
    def __init__(self, message: str, iterations: int = 0, residual: float = 0.0, reason: str = ""):
        self.iterations = iterations
        self.residual = residual
        self.reason = reason
        super().__init__(f"[FEA convergence] {message}")


def validate_elastic_constants(E: float, nu: float, name: str = "material"):
    if E <= 0:
        raise FEAValidationError(f"{name}: Young's modulus E = {E}; must be positive.",
                                 parameter="E", value=E, suggestion="bone: 10-20 GPa")
    if nu < -1.0 or nu >= 0.5:
        raise FEAValidationError(f"{name}: Poisson ratio nu = {nu}; must satisfy -1 < nu < 0.5.",
                                 parameter="nu", value=nu, suggestion="bone: 0.2-0.35")


def validate_density(density: float, name: str = "material"):
    if density <= 0:
        raise FEAValidationError(f"{name}: density = {density}; must be positive.",
                                 parameter="density", value=density)


def validate_horizon(horizon: float, spacing: float = 0.0):
    if horizon <= 0:
        raise FEAValidationError(f"horizon = {horizon}; must be positive.", parameter="horizon",
                                 value=horizon, suggestion="horizon = 3.015 x particle spacing")
    if spacing > 0 and horizon < spacing:
        logger.warning("horizon (%.4e) below particle spacing (%.4e)", horizon, spacing)


def validate_yield_stress(sigma_y: float, name: str = "material"):
    if sigma_y <= 0:
        raise FEAValidationError(f"{name}: yield stress = {sigma_y}; must be positive.",
                                 parameter="yield_stress", value=sigma_y)


def validate_hardening_modulus(H: float, name: str = "material"):
    if H < 0:
        raise FEAValidationError(f"{name}: hardening modulus = {H}; must be >= 0.",
                                 parameter="hardening_modulus", value=H)
