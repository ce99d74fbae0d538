"""Domain, Method, create_domain, create_particle_domain (reference
framework/domain.py:13-349).  Positions, selection and boundary-condition
bookkeeping are host-side; the particle grid of create_domain is the
reference's node grid in i-major, then j, then k order (domain.py:317-349)."""
from __future__ import annotations

import enum
from typing import List, Optional, Tuple, Union

import numpy as np


class Method(enum.Enum):
    FEM = "fem"
    PD = "pd"
    SPG = "spg"
    RIGID = "rigid"
    COUPLED = "coupled"


class Domain:
    def __init__(self, method: Method, dim: int, origin: Tuple[float, ...], size: Tuple[float, ...],
                 n_divisions: Tuple[int, ...], **kwargs):
        self.method = method
        self.dim = dim
        self.origin = origin
        self.size = size
        self.n_divisions = n_divisions
        self._kwargs = kwargs
        self._fixed_indices: Optional[np.ndarray] = None
        self._fixed_values: Optional[np.ndarray] = None
        self._fixed_dofs: Optional[list] = None
        self._force_indices: Optional[np.ndarray] = None
        self._force_values: Optional[np.ndarray] = None
        self._adapter = None
        self._positions: Optional[np.ndarray] = None

    def select(self, axis: int, value: float, tol: Optional[float] = None) -> np.ndarray:
        """Indices with |x_axis - value| < tol; tol defaults to half the smallest
        coordinate gap on that axis (domain.py:87-113)."""
        pos = self.get_positions()
        if tol is None:
            tol = _half_gap(pos[:, axis])
        return np.where(np.abs(pos[:, axis] - value) < tol)[0]

    def set_fixed(self, indices: np.ndarray, values: Optional[np.ndarray] = None, dofs: Optional[list] = None):
        """Dirichlet set: ``values`` are prescribed displacements (default 0) and
        ``dofs`` the fixed components (default all), both honoured by the PD
        solver here."""
        self._fixed_indices = np.asarray(indices, dtype=np.int64)
        self._fixed_values = values
        self._fixed_dofs = dofs

    def set_force(self, indices: np.ndarray, forces: Union[List[float], np.ndarray]):
        """Neumann set: one force-density vector for all indices or one per index."""
        self._force_indices = np.asarray(indices, dtype=np.int64)
        self._force_values = np.asarray(forces, dtype=np.float64)

    def get_positions(self) -> np.ndarray:
        custom = getattr(self, "_custom_positions", None)
        if custom is not None:
            return custom
        if self._positions is not None:
            return self._positions
        raise RuntimeError("domain not initialised")

    def select_boundary(self, tol: Optional[float] = None) -> np.ndarray:
        pos = self.get_positions()
        out = set()
        for ax in range(self.dim):
            c = pos[:, ax]
            t = _half_gap(c) if tol is None else tol
            out.update(np.where(np.abs(c - c.min()) < t)[0])
            out.update(np.where(np.abs(c - c.max()) < t)[0])
        return np.array(sorted(out), dtype=np.int64)

    @property
    def n_points(self) -> int:
        return len(self.get_positions())


def _half_gap(coords):
    u = np.unique(np.round(coords, decimals=10))
    return float(np.min(np.diff(u)) * 0.5) if len(u) > 1 else 1e-6


def _particle_positions(dim, origin, size, n_divisions):
    axes = []
    for q in range(dim):
        n = n_divisions[q]
        h = size[q] / (n - 1) if n > 1 else size[q]
        axes.append(origin[q] + h * np.arange(n))
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, dim).astype(np.float64)


def create_domain(method: Method, dim: int, origin: Tuple[float, ...], size: Tuple[float, ...],
                  n_divisions: Tuple[int, ...], **kwargs) -> Domain:
    d = Domain(method, dim, origin, size, n_divisions, **kwargs)
    if method == Method.FEM:
        raise NotImplementedError("Method.FEM is outside this library (peridynamics only)")
    d._positions = _particle_positions(dim, origin, size, n_divisions)
    return d


def create_particle_domain(positions: np.ndarray, method: Method, **kwargs) -> Domain:
    """A PD domain on given particle positions (domain.py:261-314)."""
    positions = np.asarray(positions, dtype=np.float64)
    n, dim = positions.shape
    lo, hi = positions.min(axis=0), positions.max(axis=0)
    size = np.maximum(hi - lo, 1e-3)
    n_axis = max(2, int(round(n ** (1.0 / dim))))
    d = create_domain(method=method, dim=dim, origin=tuple(lo.tolist()), size=tuple(size.tolist()),
                      n_divisions=tuple([n_axis] * dim), **kwargs)
    d._custom_positions = positions.copy()
    return d
