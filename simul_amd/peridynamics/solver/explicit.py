"""ExplicitSolver (reference solver/explicit.py:16-259): velocity-Verlet
explicit time integration of the bond-based (PMB) force
f_i = sum_j c s w V_j eta/|eta|, with an optional DamageModel applied after
every step (core/damage.py:144-152).  The steps run on the device
(pdx_explicit_run); run() hands the device every stretch of steps between two
callbacks in one call.

The reference force has no external load: the solver clears the particle
system's external force density when it is created.
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np


class ExplicitSolver:
    def __init__(self, particles, bonds, micromodulus: float, dt: float = 1e-6, damping: float = 0.0):
        self.particles = particles
        self.bonds = bonds
        self.dt = float(dt)
        self.damping = float(damping)
        self.dim = particles.dim
        self.n_particles = particles.n_particles
        self.c = float(micromodulus)
        body = bonds.body
        body.set_formulation("bond")
        body.set_stabilization(1.0, self.c)
        body.set("fext", np.zeros((self.n_particles, self.dim)))
        self.time = 0.0
        self.step_count = 0

    def _advance(self, n_steps: int, damage_model=None):
        if n_steps <= 0:
            return
        s_c = 0.0 if damage_model is None else float(damage_model.critical_stretch)
        if damage_model is not None and not s_c > 0.0:
            raise ValueError("the damage model's critical stretch must be positive")
        self.bonds.body.explicit_run(self.dt, self.damping, n_steps, s_c)
        self.time += self.dt * n_steps
        self.step_count += n_steps

    def step(self, damage_model=None):
        """One step: position update, forces, velocity update, then damage (explicit.py:140-158)."""
        self._advance(1, damage_model)

    def run(self, n_steps: int, damage_model=None, callback: Optional[Callable] = None,
            callback_interval: int = 100):
        """n_steps steps; callback(self, i) after step i for every i that is a
        multiple of callback_interval (explicit.py:160-181)."""
        i = 0
        while i < n_steps:
            if callback is None:
                self._advance(n_steps - i, damage_model)
                return
            nxt = -(-i // callback_interval) * callback_interval  # next multiple of the interval >= i
            last = min(nxt, n_steps - 1)
            self._advance(last - i + 1, damage_model)
            if last == nxt:
                callback(self, nxt)
            i = last + 1

    @staticmethod
    def estimate_stable_dt(youngs_modulus: float, density: float, horizon: float, spacing: float,
                           safety_factor: float = 0.5) -> float:
        """CFL estimate safety x spacing / sqrt(E / rho) (explicit.py:183-214)."""
        return safety_factor * spacing / math.sqrt(youngs_modulus / density)

    def get_kinetic_energy(self) -> float:
        """0.5 sum m v.v (explicit.py:216-228)."""
        m = self.particles.mass.to_numpy()
        v = self.particles.v.to_numpy()
        return float(0.5 * np.sum(m * np.einsum("ij,ij->i", v, v)))

    def get_strain_energy(self) -> float:
        """0.25 c sum over intact directed bonds of w s^2 |xi| V_i V_j
        (explicit.py:230-259)."""
        body = self.bonds.body
        row, col, br = body.bonds()
        _, L, w = body.bond_data()
        x = body.x
        vol = self.particles.volume.to_numpy()
        i = np.repeat(np.arange(body.n), np.diff(row))
        j = col[: len(i)].astype(np.int64)
        keep = (br[: len(i)] == 0) & (L > 1e-10)
        i, j, L, w = i[keep], j[keep], L[keep], w[keep]
        s = (np.linalg.norm(x[j] - x[i], axis=1) - L) / L
        return float(0.25 * self.c * np.sum(w * s * s * L * vol[i] * vol[j]))
