"""NeighborSearch (reference core/neighbor.py:9-319).

The neighbour set of particle i is {j != i : |X_j - X_i|^2 < horizon^2} in
FP64 (neighbor.py:179-183), built on the GPU by libpdx (cell grid of edge
1.01 horizon).  ``build`` binds the particle system to its device context.
The reference truncates rows at ``max_neighbors`` without notice; here every
neighbour is kept and a longer row is reported with a warning.
"""
from __future__ import annotations

import warnings

import numpy as np

from ...validation import validate_horizon


def _owner(positions):
    ps = getattr(positions, "_owner", None)
    if ps is None:
        raise TypeError("positions must be a ParticleSystem position field (ps.X or ps.x)")
    return ps


class NeighborSearch:
    def __init__(self, domain_min, domain_max, horizon: float, max_particles: int, max_neighbors: int = 64,
                 dim: int = 2):
        validate_horizon(horizon)
        self.dim = dim
        self.domain_min = tuple(domain_min)[:dim]
        self.domain_max = tuple(domain_max)[:dim]
        self.horizon = float(horizon)
        self.cell_size = self.horizon * 1.01
        self.max_particles = int(max_particles)
        self.max_neighbors = int(max_neighbors)
        self._ps = None
        self._counts = None
        self._row = self._col = None

    def build(self, positions, n_particles: int):
        """Neighbour rows of every particle (neighbor.py:208-218)."""
        ps = _owner(positions)
        if n_particles != ps.n_particles:
            raise ValueError(f"n_particles {n_particles} != particle system size {ps.n_particles}")
        body = ps.body if ps.body is not None else ps.bind(self.horizon)
        body.build_bonds(self.horizon)
        row, col, _ = body.bonds()
        self._ps, self._row, self._col = ps, row, col
        self._counts = np.diff(row).astype(np.int32)
        if self._counts.max(initial=0) > self.max_neighbors:
            warnings.warn(f"{int(self._counts.max())} neighbours exceed max_neighbors={self.max_neighbors}; "
                          "all are kept", RuntimeWarning, stacklevel=2)

    def count_neighbors_only(self, positions, n_particles: int) -> np.ndarray:
        """Neighbour counts without storing rows (neighbor.py:239-264)."""
        ps = _owner(positions)
        body = ps.body if ps.body is not None else ps.bind(self.horizon)
        return body.count_neighbors(self.horizon)

    def get_neighbors(self, particle_idx: int) -> np.ndarray:
        if self._row is None:
            raise RuntimeError("call build() first")
        return self._col[self._row[particle_idx]:self._row[particle_idx + 1]].astype(np.int32)

    def get_all_neighbor_counts(self) -> np.ndarray:
        if self._counts is None:
            raise RuntimeError("call build() first")
        return self._counts.copy()
