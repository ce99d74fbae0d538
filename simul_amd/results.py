"""Result formats of the reference's two callers of the PD solve.

* The orchestrator's solve stage (backend/orchestrator/stages/solve.py:10-170,
  StageResult of stages/base.py:10-17): an npz voxel model in, the framework
  Solver run with the bottom layer fixed, and ``result.npz`` out with the keys
  positions / displacements / converged / iterations / residual, plus stress
  and damage when the solver has them (solve.py:114-128).
* The API's per-region result dict (backend/api/services/analysis.py:353-376):
  name, converged, iterations and the displacement / scalar stress / damage /
  position lists, the stress reduced to one value per particle by the norm of
  its last axis, twice for a tensor (analysis.py:358-364).

Only the peridynamics method is served here; a stage for another method ends
the way the reference's stage ends on a solver error, with success False and
the message.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Optional

import numpy as np

RESULT_KEYS = ("positions", "displacements", "converged", "iterations", "residual")


@dataclass
class StageResult:
    """Outcome of a pipeline stage (stages/base.py:10-17)."""
    success: bool
    output_path: Path
    elapsed_time: float
    message: str = ""
    cached: bool = False


def result_arrays(positions, displacements, result, stress=None, damage=None) -> dict:
    """The arrays of result.npz (solve.py:114-124): converged / iterations /
    residual as one-element arrays, stress and damage only when present."""
    out = {
        "positions": np.asarray(positions),
        "displacements": np.asarray(displacements) if displacements is not None else np.array([]),
        "converged": np.array([result.converged]),
        "iterations": np.array([result.iterations]),
        "residual": np.array([result.residual]),
    }
    if stress is not None:
        out["stress"] = np.asarray(stress)
    if damage is not None:
        out["damage"] = np.asarray(damage)
    return out


def save_result_npz(path, positions, solver, result) -> Path:
    """Writes the solver's displacements, stress and damage with the solve's
    result as a compressed npz (solve.py:105-126)."""
    path = Path(path)
    np.savez_compressed(path, **result_arrays(positions, solver.get_displacements(), result,
                                              solver.get_stress(), solver.get_damage()))
    return path


def load_result_npz(path) -> dict:
    """The arrays of a result.npz (plain arrays, no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def scalar_stress(stress):
    """One value per particle: the norm over the last axis, applied twice to a
    (n, d, d) tensor, i.e. its Frobenius norm (analysis.py:358-364)."""
    if stress is None:
        return None
    s = np.asarray(stress, np.float64)
    if s.ndim > 1:
        s = np.sqrt(np.sum(s ** 2, axis=-1))
        if s.ndim > 1:
            s = np.sqrt(np.sum(s ** 2, axis=-1))
    return s


def region_result(name, positions, solver, result) -> dict:
    """The API's result dict of one material region (analysis.py:351-376)."""
    disps = solver.get_displacements()
    stress = scalar_stress(solver.get_stress())
    damage = solver.get_damage()
    return {
        "name": name,
        "converged": result.converged,
        "iterations": result.iterations,
        "displacements": disps.tolist() if disps is not None else [],
        "stress": stress.tolist() if stress is not None else [],
        "damage": damage.tolist() if damage is not None else [],
        "positions": np.asarray(positions).tolist(),
    }


class SolveStage:
    """The orchestrator's solve stage (solve.py:10-170) on the peridynamics
    solver: positions from the input npz, the bottom layer (lowest z) fixed,
    Material(E, nu, density), Solver(max_iterations, tolerance), result.npz in
    output_dir.  Solver options beyond those two pass through ``options``.

    The tolerance reaches the solve as the adapter's ``tol`` option; the
    reference stage passes it as ``tolerance``, which its PD adapter does not
    read (pd_adapter.py:159), so the reference solves to the adapter's default
    1e-4 whatever the stage is given."""

    name = "solve"

    def __init__(self, method: str = "pd", E: float = 12e9, nu: float = 0.3, density: float = 1850.0,
                 max_iterations: int = 10000, tolerance: float = 1e-6, **options):
        self.method = method
        self.E, self.nu, self.density = E, nu, density
        self.max_iterations, self.tolerance = max_iterations, tolerance
        self.options = options

    def validate_input(self, input_path) -> bool:
        p = Path(input_path)
        return p.exists() and p.suffix == ".npz"

    def run(self, input_path, output_dir, progress_callback: Optional[Callable[[str, dict], None]] = None
            ) -> StageResult:
        input_path, output_dir = Path(input_path), Path(output_dir)
        output_dir.mkdir(parents=True, exist_ok=True)
        if not self.validate_input(input_path):
            return StageResult(False, output_dir, 0.0, f"invalid input file: {input_path}")
        start = time.time()
        try:
            if self.method != "pd":
                raise NotImplementedError(f"method {self.method!r}: this library solves peridynamics only")
            from .framework import Material, Method, Solver, init
            from .framework.domain import create_particle_domain

            with np.load(input_path, allow_pickle=False) as z:
                positions = np.asarray(z["positions"], np.float64)
            if progress_callback:
                progress_callback("solve", {"message": f"setting up {len(positions)} particles"})
            init()
            domain = create_particle_domain(positions, method=Method.PD)
            domain.set_fixed(domain.select(axis=2, value=positions.min(axis=0)[2]))
            material = Material(E=self.E, nu=self.nu, density=self.density, dim=3)
            solver = Solver(domain, material, max_iterations=self.max_iterations, tol=self.tolerance,
                            **self.options)
            result = solver.solve()
            out = save_result_npz(output_dir / "result.npz", positions, solver, result)
            elapsed = time.time() - start
            if progress_callback:
                progress_callback("solve", {"message": f"done: converged={result.converged}, "
                                                       f"iterations={result.iterations} ({elapsed:.1f}s)"})
            return StageResult(True, out, elapsed, f"converged={result.converged}, iterations={result.iterations}")
        except Exception as e:  # the reference's stage reports any solver error this way
            return StageResult(False, output_dir, time.time() - start, f"solve failed: {e}")
