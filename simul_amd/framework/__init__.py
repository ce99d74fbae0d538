"""Mirror of the reference's unified framework API (backend/fea/framework)
for the peridynamics method:

    init()
    domain = create_domain(Method.PD, dim=3, origin=(0, 0, 0), size=(19, 19, 19), n_divisions=(20, 20, 20))
    domain.set_fixed(domain.select(axis=2, value=0.0))
    domain.set_force(domain.select(axis=2, value=19.0), [0.0, 0.0, -1.0])
    solver = Solver(domain, Material(E=15000.0, nu=0.3, density=1.9e-9, dim=3))
    result = solver.solve()
    u = solver.get_displacements()

Only Method.PD runs here (on libpdx); the reference's FEM / SPG / coupled
methods are outside this library and raise NotImplementedError.
"""
from .domain import Domain, Method, create_domain, create_particle_domain
from .material import Material
from .result import SolveResult
from .runtime import Backend, Precision, get_backend, get_precision, init
from .solver import Solver

__all__ = ["init", "Backend", "Precision", "get_backend", "get_precision", "create_domain",
           "create_particle_domain", "Domain", "Method", "Material", "Solver", "SolveResult"]
