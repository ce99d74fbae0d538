"""SolveResult (reference framework/result.py:6-21)."""
from dataclasses import dataclass


@dataclass
class SolveResult:
    converged: bool
    iterations: int
    residual: float
    relative_residual: float
    elapsed_time: float
