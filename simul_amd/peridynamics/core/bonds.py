"""BondSystem (reference core/bonds.py:13-208).

The bonds live on the device (SELL rows of the bound context); the fields here
are (n_particles, max_bonds) host views in the reference layout, each row
listing the neighbours by ascending index, padded with zeros.
"""
from __future__ import annotations

import numpy as np


class _BondField:
    """(n_particles, max_bonds[, dim]) view with the Taichi field host API."""

    def __init__(self, bonds, name):
        self._bonds, self._name = bonds, name

    def to_numpy(self) -> np.ndarray:
        return self._bonds._view(self._name)

    def from_numpy(self, arr):
        if self._name != "broken":
            raise NotImplementedError(f"bond field '{self._name}' is derived from the reference positions")
        self._bonds._set_broken(np.asarray(arr))

    def __getitem__(self, idx):
        return self.to_numpy()[idx]

    def __setitem__(self, idx, value):
        a = self.to_numpy()
        a[idx] = value
        self.from_numpy(a)


class BondSystem:
    @classmethod
    def from_neighbor_counts(cls, n_particles: int, counts: np.ndarray, dim: int = 2, margin: int = 8):
        return cls(n_particles, max_bonds=int(np.max(counts)) + margin, dim=dim)

    def __init__(self, n_particles: int, max_bonds: int = 64, dim: int = 2):
        self.n_particles = int(n_particles)
        self.max_bonds = int(max_bonds)
        self.dim = dim
        self._ps = None
        self._horizon = None
        for name in ("neighbors", "n_neighbors", "xi", "xi_length", "broken", "omega", "initial_bonds"):
            setattr(self, name, _BondField(self, name))

    @property
    def body(self):
        if self._ps is None:
            raise RuntimeError("bonds not built: call build_from_neighbor_search first")
        return self._ps.body

    def build_from_neighbor_search(self, particles, neighbor_search, horizon: float):
        """Bonds of the neighbour rows with xi, |xi|, omega (bonds.py:82-137)."""
        if neighbor_search._ps is not particles:
            raise ValueError("the neighbour search was built on another particle system")
        if particles.body.horizon != float(horizon):
            particles.body.build_bonds(float(horizon))
        self._ps = particles
        self._horizon = float(horizon)
        self._initial = np.diff(self.body.bonds()[0]).astype(np.int32)
        if self._initial.max(initial=0) > self.max_bonds:
            raise ValueError(f"{int(self._initial.max())} bonds per particle exceed max_bonds={self.max_bonds}")

    # -- reference-layout views ---------------------------------------------
    def _padded(self, row, vals, tail=()):
        n, mb = self.n_particles, self.max_bonds
        out = np.zeros((n, mb) + tail, vals.dtype)
        k = np.arange(len(vals)) - np.repeat(row[:-1], np.diff(row))
        i = np.repeat(np.arange(n), np.diff(row))
        out[i, k] = vals
        return out

    def _view(self, name):
        b = self.body
        row, col, br = b.bonds()
        if name == "n_neighbors":
            return np.diff(row).astype(np.int32)
        if name == "initial_bonds":
            return self._initial.copy()
        if name == "neighbors":
            return self._padded(row, col.astype(np.int32))
        if name == "broken":
            return self._padded(row, br.astype(np.int32))
        xi, L, w = b.bond_data()
        if name == "xi":
            return self._padded(row, xi, (self.dim,))
        return self._padded(row, L if name == "xi_length" else w)

    def _set_broken(self, arr):
        row, _, _ = self.body.bonds()
        i = np.repeat(np.arange(self.n_particles), np.diff(row))
        k = np.arange(row[-1]) - np.repeat(row[:-1], np.diff(row))
        self.body.set_broken(arr[i, k] != 0)

    # -- queries (bonds.py:140-208) -------------------------------------------
    def reset_bonds(self):
        self.body.set_broken(np.zeros(self.body.n_bonds, bool))

    def get_neighbor_count(self) -> np.ndarray:
        return self._view("n_neighbors")

    def get_broken_bonds(self) -> np.ndarray:
        return self._view("broken")

    def count_intact_bonds(self) -> int:
        return int(self.body.stats()["n_intact"])

    def count_broken_bonds(self) -> int:
        st = self.body.stats()
        return int(st["n_bonds"] - st["n_intact"])
