"""Mirror of backend/fea/peridynamics/solver (NOSB kinetic-damping DR, bond-based quasi-static,
explicit dynamics)."""
