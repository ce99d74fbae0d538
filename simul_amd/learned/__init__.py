"""Learned warm start for dynamic relaxation (no reference counterpart).

A per-particle network maps (normalised coordinates, material, boundary
conditions, load vector, two radius-delta message-passing hops of these) to a
displacement field; DR starts from it instead of from rest.  Training and
inference run on the BF16 MFMA kernels of libpdx (include/pdx_nn.h); the plain
cold-start solve stays the fallback and the judge: a warm-started solve uses
the same residual criterion, relative to the cold-start residual.
"""
from .features import FEAT, feature_matrix, target_scale
from .net import WarmStartNet
from .warmstart import warm_solve

__all__ = ["FEAT", "feature_matrix", "target_scale", "WarmStartNet", "warm_solve"]
