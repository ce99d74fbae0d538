"""Runtime initialisation (reference framework/runtime.py:35-132).

The reference picks a Taichi backend (CUDA -> Vulkan -> CPU fallback).  Here
the only backend is HIP on the process's current GPU; there is no CPU
fallback, so ``init`` raises when no GPU is visible.  Idempotent per process.
"""
from __future__ import annotations

import enum
from typing import Optional


class Backend(enum.Enum):
    HIP = "hip"
    AUTO = "auto"


class Precision(enum.Enum):
    F32 = "f32"
    F64 = "f64"


_state = {"backend": None, "precision": None}


def init(backend: Backend = Backend.AUTO, precision: Precision = Precision.F64) -> dict:
    """Loads libpdx and checks for a GPU once per process.  ``precision`` is
    recorded for callers; the solver's precision policy is fixed (FP32 bond
    arithmetic, FP64 K^-1, tensor algebra and reductions)."""
    if _state["backend"] is not None:
        return {"backend": _state["backend"].value, "precision": _state["precision"].value,
                "already_initialized": True}
    if backend not in (Backend.AUTO, Backend.HIP):
        raise ValueError(f"unsupported backend {backend}")
    from .. import _native as N
    N.load()
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device available: the peridynamics solver runs on the GPU only")
    _state["backend"], _state["precision"] = Backend.HIP, precision
    return {"backend": Backend.HIP.value, "precision": precision.value, "already_initialized": False}


def get_backend() -> Optional[Backend]:
    return _state["backend"]


def get_precision() -> Optional[Precision]:
    return _state["precision"]
