"""DamageModel (reference core/damage.py:14-170): critical-stretch bond
breaking with the FP64 stretch (|eta| - |xi|)/|xi| > s_c strictly, and local
damage phi = broken / initial bonds, both on the device."""
from __future__ import annotations

import math


class DamageModel:
    def __init__(self, critical_stretch: float = 0.01, dim: int = 2):
        self.critical_stretch = float(critical_stretch)
        self.dim = dim

    @staticmethod
    def compute_critical_stretch(youngs_modulus: float, fracture_energy: float, horizon: float,
                                 dim: int = 3) -> float:
        """s_c = sqrt(4 pi G_c / (9 E delta)) in 2-D, sqrt(5 G_c / (9 E delta)) in 3-D (damage.py:45-68)."""
        if dim == 2:
            return math.sqrt(4 * math.pi * fracture_energy / (9 * youngs_modulus * horizon))
        return math.sqrt(5 * fracture_energy / (9 * youngs_modulus * horizon))

    @staticmethod
    def compute_critical_stretch_bone(bone_type: str = "cortical", horizon: float = 0.003) -> float:
        """damage.py:70-92: cortical E 17 GPa, G_c 500 J/m^2; cancellous 0.5 GPa, 50 J/m^2."""
        E, G_c = (17e9, 500.0) if bone_type == "cortical" else (0.5e9, 50.0)
        return DamageModel.compute_critical_stretch(E, G_c, horizon, dim=3)

    def update_damage(self, particles, bonds, critical_stretch: float) -> int:
        """Breaks every intact bond with stretch > s_c (damage.py:95-117); returns
        the number broken now."""
        return bonds.body.break_bonds(float(critical_stretch))

    def compute_local_damage(self, particles, bonds):
        """phi_i = broken_i / initial_i into particles.damage (damage.py:119-142)."""
        bonds.body.local_damage()

    def step(self, particles, bonds):
        self.update_damage(particles, bonds, self.critical_stretch)
        self.compute_local_damage(particles, bonds)

    def get_statistics(self, bonds) -> dict:
        """Intact / broken / total bonds and their ratio (damage.py:154-170)."""
        intact = bonds.count_intact_bonds()
        broken = bonds.count_broken_bonds()
        total = intact + broken
        return {"intact_bonds": int(intact), "broken_bonds": int(broken), "total_bonds": int(total),
                "damage_ratio": broken / total if total > 0 else 0.0}
