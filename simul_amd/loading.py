"""Incremental quasi-static loading: a prescribed-displacement ramp solved by
dynamic relaxation increment by increment, with the J2 state committed and
critical-stretch bonds broken at each converged increment.

The reference applies prescribed displacements by writing the fixed
particles' positions (tests/test_3d.py:366-377), commits J2 state on every
stress evaluation (fem/material/j2_plasticity.py:242-243) and leaves bond
breaking to the caller (core/damage.py:144-152, NOSBSolver.update_shape_tensor
at solver/nosb_solver.py:258-260).  Here one increment is:

1. prescribed DOFs moved to (k / n) x their final displacement;
2. DR from the previous equilibrium to |f_free| / |f_start| < tol, the trial
   J2 state evaluated from the committed one in every pass;
3. bonds with stretch > s_c broken (FP64 stretch, strict), the shape tensor
   recomputed and the increment re-equilibrated to the same tolerance of the
   increment's start residual, up to ``max_break_rounds`` times;
4. the J2 state committed.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np


@dataclass
class IncrementResult:
    increment: int
    iterations: int
    converged: bool
    relative_residual: float
    broken_now: int
    broken_total: int
    break_rounds: int
    elapsed_s: float


class IncrementalLoading:
    def __init__(self, body, indices, u_final, n_increments: int = 20, critical_stretch: Optional[float] = None,
                 tol: float = 1e-6, max_iterations: int = 200000, method: str = "adaptive",
                 max_break_rounds: int = 3, dt: float = 0.0, viscous_damping: float = 0.0,
                 progress=None, slice_iterations: int = 4096, mass=None, predictor=None, start=None):
        if n_increments < 1:
            raise ValueError("n_increments must be >= 1")
        self.body = body
        self.indices = np.asarray(indices, np.int64)
        self.u_final = np.broadcast_to(np.asarray(u_final, np.float64), (len(self.indices), body.dim)).copy()
        self.n_increments = int(n_increments)
        self.critical_stretch = critical_stretch
        self.tol, self.max_iterations, self.method = float(tol), int(max_iterations), method
        self.max_break_rounds = int(max_break_rounds)
        self.dt, self.viscous_damping = float(dt), float(viscous_damping)
        self.progress, self.slice_iterations = progress, slice_iterations
        # adaptive-DR mass of the solves (PDBody.solve's mass=).  "spectral"
        # tunes the first solve's mass to the spectrum and reuses its beta for
        # the later solves (the elastic stiffness, which bounds the spectrum,
        # does not grow with plastic flow; a solve that diverges with the
        # reused beta tunes again)
        self.mass = mass
        self._beta = None
        # start of an increment: None = the previous state with the new
        # boundary values; "secant" = 2 u_(k-1) - u_(k-2) from the second
        # increment on, u_0 being the state the loading starts from; ``start``
        # = a callable k -> (n, dim) float64 CUDA tensor, the first guess of
        # an increment that has no secant guess (a learned warm start).  The
        # convergence reference stays the residual of the plain start, so
        # every start reaches the same tolerance.
        if predictor not in (None, "secant"):
            raise ValueError("predictor must be None or 'secant'")
        self.predictor = predictor
        self.start = start
        # displacement states stay on the GPU (pdx_get_u_device / pdx_set_u_device)
        import torch
        self._idx = torch.from_numpy(self.indices).cuda()
        self._u_final = torch.from_numpy(self.u_final).cuda()
        self._hist: list = [body.get_u_device()] if predictor == "secant" else []
        # broken bonds so far: counted once here, then from each break_bonds
        # call (pdx_stats reads the whole bond list back)
        st = body.stats()
        self._broken = int(st["n_bonds"] - st["n_intact"])
        self.results: List[IncrementResult] = []

    def _solve(self, ref=0.0):
        b = self.body
        mass = self.mass if self.method == "adaptive" else None
        reuse = mass == "spectral" and self._beta is not None
        r = b.solve(self.method, tol=self.tol, max_iterations=self.max_iterations, ref=ref, dt=self.dt,
                    viscous_damping=self.viscous_damping, progress=self.progress,
                    slice_iterations=self.slice_iterations if self.progress is not None else None,
                    mass=self._beta if reuse else mass)
        if mass == "spectral":
            self._beta = getattr(b, "last_beta", None)
        return r

    def step(self, k: int) -> IncrementResult:
        """Increment k (1-based)."""
        b = self.body
        t0 = time.perf_counter()
        u = b.get_u_device()
        u[self._idx] = (k / self.n_increments) * self._u_final
        b.set_u_device(u)
        guess = None
        if self.predictor == "secant" and len(self._hist) == 2:
            guess = 2.0 * self._hist[1] - self._hist[0]
        elif self.start is not None:
            guess = self.start(k)
        if guess is not None:
            b.compute_forces(download=False)
            ref0 = float(np.atleast_1d(b.residual_norm())[0])
            guess[self._idx] = u[self._idx]
            b.set_u_device(guess)
            r = self._solve(ref0 if ref0 > 0.0 else 0.0)
        else:
            r = self._solve()
        # re-equilibration after breaking is measured against the increment's
        # own start residual: the residual right after a few bonds break is
        # tiny, and 1e-6 of it lies below the FP32 force noise
        ref = r.reference_residual
        iters = r.iterations
        broken_now = 0
        rounds = 0
        if self.critical_stretch is not None:
            while rounds < self.max_break_rounds:
                nb = b.break_bonds(self.critical_stretch)
                if nb == 0:
                    break
                broken_now += nb
                rounds += 1
                if b.formulation == "nosb":
                    b.shape_tensor(b.singular_test, b.singular_tol)
                r = self._solve(ref)
                iters += r.iterations
        if b.model == "j2_plasticity":
            b.commit_state()
        if self.predictor is not None:
            self._hist = (self._hist + [b.get_u_device()])[-2:]
        self._broken += broken_now
        res = IncrementResult(k, iters, bool(r.converged), float(r.relative_residual), broken_now,
                              self._broken, rounds, time.perf_counter() - t0)
        self.results.append(res)
        return res

    def run(self, callback=None) -> List[IncrementResult]:
        for k in range(len(self.results) + 1, self.n_increments + 1):
            res = self.step(k)
            if callback is not None:
                callback(res)
        return self.results
