"""QuasiStaticSolver and LoadControl (reference solver/quasi_static.py:17-377).

The bond-based (PMB) force f_i = sum_j c s w V_j eta/|eta| with kinetic
damping: velocity Verlet, the viscous factor (1 - damping), a velocity reset
at each kinetic-energy peak, convergence on |f_free| / |f_start| < tol.  The
iteration loop runs on the device (libpdx kinetic DR with hipGraph chunks);
the reference's NaN guards on dx / a / v are not reproduced.
"""
from __future__ import annotations

import numpy as np


class LoadControl:
    """Adds one load vector to the force of the loaded particles (quasi_static.py:344-377)."""

    def __init__(self, particles, loaded_indices):
        self.particles = particles
        self.dim = particles.dim
        self.n_loaded = len(loaded_indices)
        self.load_mask = np.zeros(particles.n_particles, np.int32)
        self.load_mask[np.asarray(loaded_indices, np.int64)] = 1
        self.load = np.zeros(self.dim)

    def set_load(self, force_per_particle):
        self.load = np.asarray(list(force_per_particle)[: self.dim], np.float64)

    def field(self) -> np.ndarray:
        """The load as a per-particle (n, dim) force density."""
        return self.load_mask[:, None] * self.load[None, :]

    def apply(self):
        f = self.particles.f.to_numpy()
        f += self.field()
        self.particles.f.from_numpy(f)


def device_load(external_force_func):
    """The (n, dim) load a solve can hand to the device for an external force
    callback, or None for no callback.  Only LoadControl.apply translates; any
    other callable would have to run between the force and velocity updates of
    every iteration, which the device loop does not do."""
    if external_force_func is None:
        return None
    owner = getattr(external_force_func, "__self__", None)
    if isinstance(owner, LoadControl) and external_force_func.__func__ is LoadControl.apply:
        return owner.field()
    raise NotImplementedError("external_force_func must be None or LoadControl.apply; set other loads as a "
                              "per-particle force density (ParticleSystem.body.set_bc(fext=...))")


class _KineticLoop:
    """Shared device kinetic-damping loop of the two solvers."""

    def _run(self, external_force_func, max_iterations, tol, visc, verbose):
        body = self.particles.body
        fext = device_load(external_force_func)
        body.set("fext", np.zeros((self.n_particles, self.dim)) if fext is None else fext)
        try:
            r = body.solve("kinetic", tol=tol, max_iterations=max_iterations, dt=self.dt,
                           viscous_damping=visc)
        finally:
            body.set("fext", np.zeros((self.n_particles, self.dim)))
        self.iteration = r.iterations
        if verbose:
            print(f"kinetic DR: {r.iterations} iterations, rel. residual {r.relative_residual:.3e}, "
                  f"{r.resets} velocity resets, converged={r.converged}")
        return r

    def _step(self, external_force_func, visc):
        body = self.particles.body
        fext = device_load(external_force_func)
        if fext is not None:
            body.set("fext", fext)
        try:
            info = body.kd_step(self.dt, visc, self.prev_ke, self.ke_increasing)
        finally:
            if fext is not None:
                body.set("fext", np.zeros((self.n_particles, self.dim)))
        self.prev_ke = float(info["prev_ke"][0])
        self.ke_increasing = bool(info["ke_increasing"][0])
        self.iteration += 1
        return {"kinetic_energy": float(info["kinetic_energy"][0]), "residual": float(info["residual"][0]),
                "velocity_reset": bool(info["velocity_reset"][0])}


class QuasiStaticSolver(_KineticLoop):
    def __init__(self, particles, bonds, micromodulus: float, dt: float = None, damping: float = 0.1):
        self.particles = particles
        self.bonds = bonds
        self.dim = particles.dim
        self.n_particles = particles.n_particles
        self.damping = float(damping)
        self.c = float(micromodulus)
        body = bonds.body
        body.set_formulation("bond")
        body.set_stabilization(1.0, self.c)
        self.dt = float(dt) if dt is not None else self._estimate_stable_dt()
        self.prev_ke = 0.0
        self.ke_increasing = True
        self.iteration = 0
        self._init_acceleration()

    def _estimate_stable_dt(self) -> float:
        """CFL-type estimate of quasi_static.py:68-109: spacing from the mean
        volume, horizon from the mean neighbour count, wave speed
        sqrt(c delta V / rho), dt = 0.01 spacing / c_wave clipped to [1e-12, 1e-6]."""
        ps = self.particles
        vol = ps.volume.to_numpy()
        rho = ps.density.to_numpy()
        avg_vol, avg_rho = float(np.mean(vol)), float(np.mean(rho))
        spacing = avg_vol ** (1.0 / self.dim)
        nn = float(np.mean(self.bonds.get_neighbor_count()))
        ratio = (3 * nn / (4 * np.pi)) ** (1.0 / 3.0) if self.dim == 3 else np.sqrt(nn / np.pi)
        horizon = ratio * spacing
        c_wave = np.sqrt(self.c * horizon * avg_vol / avg_rho + 1e-20)
        return float(max(1e-12, min(0.01 * spacing / c_wave, 1e-6)))

    def _init_acceleration(self):
        z = np.zeros((self.n_particles, self.dim))
        self.particles.a.from_numpy(z)
        self.particles.v.from_numpy(z)

    def step(self, external_force_func=None) -> dict:
        return self._step(external_force_func, self.damping)

    def solve(self, external_force_func=None, max_iterations: int = 100000, tol: float = 1e-6,
              verbose: bool = True, print_interval: int = 5000) -> dict:
        self._init_acceleration()
        self.prev_ke, self.ke_increasing = 0.0, True
        r = self._run(external_force_func, max_iterations, tol, self.damping, verbose)
        return {"converged": r.converged, "iterations": r.iterations, "residual": r.residual,
                "relative_residual": r.relative_residual, "velocity_resets": r.resets}
