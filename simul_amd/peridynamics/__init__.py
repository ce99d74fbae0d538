"""Host-side mirror of the reference's peridynamics package
(backend/fea/peridynamics): the same classes, methods, argument meanings and
errors, with every kernel running in libpdx on the GPU.

Objects keep the reference's field accessors (``ps.x.to_numpy()``,
``bonds.xi_length.to_numpy()`` ...).  A particle system is bound to one libpdx
context from its first neighbour search on; from then on its fields read and
write the device state.
"""
from .core.bonds import BondSystem
from .core.damage import DamageModel
from .core.neighbor import NeighborSearch
from .core.nosb import NOSBCompute, NOSBMaterial
from .core.particles import ParticleSystem
from .solver.explicit import ExplicitSolver
from .solver.nosb_solver import NOSBSolver
from .solver.quasi_static import LoadControl, QuasiStaticSolver

__all__ = ["ParticleSystem", "NeighborSearch", "BondSystem", "NOSBCompute", "NOSBMaterial", "DamageModel",
           "NOSBSolver", "QuasiStaticSolver", "LoadControl", "ExplicitSolver"]
