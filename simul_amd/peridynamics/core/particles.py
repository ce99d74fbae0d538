"""ParticleSystem (reference core/particles.py:9-231).

Fields are ``Field`` objects with the Taichi field's ``to_numpy`` /
``from_numpy`` / item access.  Before the system is bound to a libpdx context
they are host arrays; after ``bind`` (done by the first neighbour search) the
device-resident ones read and write the context.
"""
from __future__ import annotations

import numpy as np

from ...validation import FEAValidationError, validate_density

# field name -> libpdx field id name (device-resident once bound)
_DEVICE = {"x": "x", "v": "v", "a": "a", "f": "f", "volume": "volume", "density": "density",
           "mass": "mass", "damage": "damage", "bulk_mod": "bulk", "shear_mod": "shear", "K": "K",
           "K_inv": "K_inv", "F": "F", "P": "P", "fixed": "fixed"}


class Field:
    """One per-particle quantity with the reference Taichi field's host API."""

    def __init__(self, owner, name, tail, dtype):
        self._owner, self._name = owner, name
        self._host = np.zeros((owner.n_particles,) + tuple(tail), dtype)

    @property
    def shape(self):
        return self._host.shape

    def to_numpy(self) -> np.ndarray:
        return self._owner._pull(self._name).copy()

    def from_numpy(self, arr):
        arr = np.asarray(arr)
        if arr.shape != self._host.shape:
            raise ValueError(f"{self._name}: expected shape {self._host.shape}, got {arr.shape}")
        self._host[...] = arr
        self._owner._push(self._name)

    def fill(self, value):
        self._host[...] = value
        self._owner._push(self._name)

    def __getitem__(self, idx):
        return self.to_numpy()[idx]

    def __setitem__(self, idx, value):
        a = self.to_numpy()
        a[idx] = value
        self.from_numpy(a)


class ParticleSystem:
    """Particle data of a peridynamic body (particles.py:26-62)."""

    def __init__(self, n_particles: int, dim: int = 2):
        if dim not in (2, 3):
            raise ValueError(f"dim must be 2 or 3, got {dim}")
        self.n_particles = int(n_particles)
        self.dim = dim
        self._body = None
        d = (dim,)
        dd = (dim, dim)
        for name, tail, dt in (("X", d, np.float64), ("x", d, np.float64), ("v", d, np.float64),
                               ("a", d, np.float64), ("f", d, np.float64), ("volume", (), np.float64),
                               ("density", (), np.float64), ("mass", (), np.float64),
                               ("damage", (), np.float64), ("bulk_mod", (), np.float64),
                               ("shear_mod", (), np.float64), ("K", dd, np.float64), ("K_inv", dd, np.float64),
                               ("F", dd, np.float64), ("P", dd, np.float64), ("fixed", (), np.int32)):
            setattr(self, name, Field(self, name, tail, dt))

    # -- device binding ------------------------------------------------------
    @property
    def body(self):
        """The bound libpdx body (simul_amd.body.PDBody) or None."""
        return self._body

    def bind(self, horizon: float):
        """Creates the libpdx context for this system (first neighbour search).
        Host field values become the device's initial state."""
        from ...body import PDBody
        h = {k: f._host for k, f in vars(self).items() if isinstance(f, Field)}
        vol = h["volume"]
        if not np.all(vol > 0):
            raise FEAValidationError("particle volumes must be positive (initialise the system first)",
                                     parameter="volume")
        bulk = np.where(h["bulk_mod"] > 0, h["bulk_mod"], 1.0)
        shear = np.where(h["shear_mod"] > 0, h["shear_mod"], 1.0)
        self._body = PDBody(h["X"], vol, horizon, bulk, shear, mass=h["mass"], build=False)
        self._push("fixed")
        self._body.x = h["x"]
        for name in ("v", "a", "density"):
            self._push(name)
        return self._body

    def _pull(self, name):
        f = getattr(self, name)
        b = self._body
        if b is not None and name in _DEVICE:
            v = b.get(_DEVICE[name])
            if name == "fixed":
                v = (v != 0).astype(np.int32)
            f._host[...] = v
        return f._host

    def _push(self, name):
        b = self._body
        if b is None or name not in _DEVICE or name in ("K", "K_inv", "F", "P"):
            return
        v = getattr(self, name)._host
        if name == "fixed":
            v = np.where(v != 0, (1 << self.dim) - 1, 0).astype(np.uint8)
        b.set(_DEVICE[name], v)

    # -- initialisation (particles.py:64-173) --------------------------------
    def initialize_from_grid(self, origin: tuple, spacing: float, n_points: tuple, density: float = 1000.0):
        validate_density(density, "PD particle system")
        if spacing <= 0:
            raise FEAValidationError(f"particle spacing is {spacing}; must be positive.",
                                     parameter="spacing", value=spacing)
        axes = [origin[q] + spacing * np.arange(n_points[q]) for q in range(self.dim)]
        # i-major, then j, then k (particles.py:86-101)
        pos = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, self.dim)
        if len(pos) != self.n_particles:
            raise ValueError(f"Grid produces {len(pos)} particles but system was initialized for "
                             f"{self.n_particles}")
        self._set_positions(pos)
        self._set_uniform(density, spacing ** self.dim)

    def initialize_from_arrays(self, positions: np.ndarray, volumes: np.ndarray, density: float = 1000.0):
        positions = np.asarray(positions, np.float64)
        volumes = np.asarray(volumes, np.float64)
        if positions.shape != (self.n_particles, self.dim) or volumes.shape != (self.n_particles,):
            raise ValueError("positions must be (n_particles, dim) and volumes (n_particles,)")
        self._set_positions(positions)
        self.volume._host[...] = volumes
        self.density._host[...] = density
        self.mass._host[...] = volumes * density
        self._init_fields()

    def _set_positions(self, pos):
        if self._body is not None:
            raise RuntimeError("reference positions are fixed once the system is bound to a device context")
        self.X._host[...] = pos
        self.x._host[...] = pos

    def _set_uniform(self, density, volume):
        self.density._host[...] = density
        self.volume._host[...] = volume
        self.mass._host[...] = volume * density
        self._init_fields()

    def _init_fields(self):
        for name in ("v", "a", "f", "damage", "fixed", "P"):
            getattr(self, name)._host[...] = 0
        eye = np.eye(self.dim)
        for name in ("K", "K_inv", "F"):
            getattr(self, name)._host[...] = eye

    # -- state helpers (particles.py:176-231) ---------------------------------
    def reset_forces(self):
        self.f.fill(0.0)

    def compute_acceleration(self):
        m = self.mass.to_numpy()
        ok = (self.fixed.to_numpy() == 0) & (m > 0)
        a = np.zeros((self.n_particles, self.dim))
        a[ok] = self.f.to_numpy()[ok] / m[ok, None]
        self.a.from_numpy(a)

    def set_fixed_particles(self, indices: np.ndarray):
        fixed = np.zeros(self.n_particles, np.int32)
        fixed[np.asarray(indices, np.int64)] = 1
        self.fixed.from_numpy(fixed)

    def set_material_constants(self, bulk_modulus: float, shear_modulus: float):
        self.set_material_constants_per_particle(np.full(self.n_particles, bulk_modulus),
                                                 np.full(self.n_particles, shear_modulus))

    def set_material_constants_per_particle(self, bulk_arr, shear_arr):
        self.bulk_mod._host[...] = np.asarray(bulk_arr, np.float64)
        self.shear_mod._host[...] = np.asarray(shear_arr, np.float64)
        if self._body is not None:
            self._body.set_materials(self._body.model, self.bulk_mod._host, self.shear_mod._host)

    def get_positions(self) -> np.ndarray:
        return self.x.to_numpy()

    def get_displacements(self) -> np.ndarray:
        return self.x.to_numpy() - self.X.to_numpy()

    def get_damage(self) -> np.ndarray:
        return self.damage.to_numpy()
