"""Shared pieces of the kernel adapters (orihime/ops/sw.py:25-130,
orihime/ops/sw_affine.py:26-109)."""

from __future__ import annotations

import warnings

import torch
from torch import Tensor

from .. import _checks
from ..operator import KernelCall, _validate_temperature


def warn_positive_gap(value, name: str) -> None:
    if isinstance(value, (int, float)) and value > 0:
        warnings.warn(
            f"{name}={value} is positive. Gap penalties are typically negative. "
            "If this is intentional, pass as tensor to suppress this warning.",
            UserWarning,
            stacklevel=5,
        )


def scalar(param) -> float:
    if isinstance(param, Tensor):
        return _checks.host_scalar(param, "scoring parameter")
    return float(param)


def _prefix_mask_lengths(mask: Tensor, name: str) -> Tensor:
    m = mask.to(torch.int32)
    has_hole = (m[:, 1:] - m[:, :-1] > 0).any(dim=1)
    # the hole check reads back one bool; it is skipped under graph capture
    if not _checks.deferred(has_hole) and has_hole.any():
        bad = has_hole.nonzero(as_tuple=True)[0].tolist()
        raise ValueError(
            f"{name} is not prefix-only (has holes) at batch indices: {bad}. Masks must have "
            "all True values at the beginning, followed by all False values."
        )
    return mask.sum(dim=1).to(torch.int32)


def resolve_lengths(call: KernelCall) -> Tensor | None:
    """lengths, or prefix masks mask1/mask2 converted to lengths; None means
    every pair uses the full shape."""
    mask1 = call.config.get("mask1")
    mask2 = call.config.get("mask2")
    if call.lengths is not None and (mask1 is not None or mask2 is not None):
        raise ValueError("Cannot specify both 'lengths' and 'mask1'/'mask2'")
    if mask1 is None and mask2 is None:
        return call.lengths
    B, L1, L2 = call.primary.shape
    lengths = torch.empty((B, 2), dtype=torch.int32, device=call.primary.device)
    lengths[:, 0] = _prefix_mask_lengths(mask1, "mask1") if mask1 is not None else L1
    lengths[:, 1] = _prefix_mask_lengths(mask2, "mask2") if mask2 is not None else L2
    return lengths


def check_temperature(temp) -> None:
    _validate_temperature(temp)
