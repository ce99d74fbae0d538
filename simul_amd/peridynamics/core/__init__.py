"""Mirror of backend/fea/peridynamics/core (particles, neighbours, bonds, NOSB, damage)."""
