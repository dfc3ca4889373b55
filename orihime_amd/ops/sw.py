"""Linear-gap soft Smith-Waterman kernel adapter (orihime/ops/sw.py).

Same structure as the affine adapter: one fused native call for value, map
and per-pair parameter gradients; a value-only primitive; a one-pass map VJP.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor

from .. import _ops
from ..operator import (
    BackwardPass,
    ForwardPass,
    KernelCall,
    OperatorKernels,
    _validate_derivative_vector,
)
from ._common import check_temperature, resolve_lengths, scalar, warn_positive_gap

_PARAM_NAMES = ("gap", "temp")


def _prepare(call: KernelCall):
    gap, temp = call.params
    warn_positive_gap(gap, "gap")
    check_temperature(temp)
    return (scalar(gap), scalar(temp)), resolve_lengths(call)


def _forward_kernel(call: KernelCall) -> ForwardPass:
    params, lengths = _prepare(call)
    value, mp, g_gap, g_temp = _ops.orihime.soft_sw_with_grads(call.primary, *params, lengths)
    return ForwardPass(value=value, saved_tensors=(mp, g_gap, g_temp))


def _backward_kernel(call: KernelCall, forward: ForwardPass) -> BackwardPass:
    mp, g_gap, g_temp = forward.saved_tensors
    return BackwardPass(marginals=mp, entropy=g_temp, param_grads=(g_gap, g_temp))


def _value_only(call: KernelCall) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sw_value_only(call.primary, *params, lengths)


def _map_backward_kernel(call: KernelCall, state, grad_map: Tensor) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sw_marginals_hvp(call.primary, grad_map.contiguous(), *params, lengths)


def _map_vjp(call: KernelCall, grad_map: Tensor):
    params, lengths = _prepare(call)
    gs, a, b = _ops.sw_marginals_backward(call.primary, grad_map.contiguous(), *params, lengths)
    return gs, (a, b)


def _param_backward_kernel(call: KernelCall, state, name: str, grad_map: Tensor | None) -> Tensor:
    params, lengths = _prepare(call)
    if name == "gap":
        field = _ops.sw_marginals_grad_gap(call.primary, *params, lengths)
        param = call.params[0]
    elif name == "temp":
        field = _ops.sw_marginals_grad_temp(call.primary, *params, lengths)
        param = call.params[1]
    else:
        raise ValueError(f"unknown sw parameter {name!r}")
    if grad_map is None:
        return field
    contracted = (field * grad_map).sum()
    if isinstance(param, Tensor):
        return contracted.to(device=param.device, dtype=param.dtype).reshape(param.shape)
    return contracted.reshape(1)


kernels = OperatorKernels(
    forward=_forward_kernel,
    backward=_backward_kernel,
    map_backward=_map_backward_kernel,
    param_backward=_param_backward_kernel,
    map_vjp=_map_vjp,
    value_only=_value_only,
    config={"mask1": None, "mask2": None},
)


def forward(scores: Tensor, gap: float, temp: float,
            lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Value [B] and alignment marginals [B, L1, L2] for scalar parameters."""
    return _ops.sw_forward(scores, gap, temp, lengths)


def forward_t(scores: Tensor, gap: Tensor, temp: Tensor,
              lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """As ``forward`` with tensor parameters (differentiable in them)."""
    if lengths is None:
        B, L1, L2 = scores.shape
        lengths = torch.tensor([[L1, L2]] * B, dtype=torch.int32, device=scores.device)
    return _ops.sw_forward_t(scores, gap, temp, lengths)


def value_grad_params(scores: Tensor, gap: float, temp: float,
                      lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Per-pair gradients of the value w.r.t. gap and temperature."""
    return _ops.sw_value_grad_params(scores, gap, temp, lengths)


def marginals_backward(scores: Tensor, grad_marginals: Tensor, gap: float, temp: float,
                       lengths: Optional[Tensor] = None):
    """Gradients of <grad_marginals, map> w.r.t. scores and (gap, temp) ([1] each)."""
    _validate_derivative_vector(scores, grad_marginals, name="cotangent", primary_name="scores")
    return _ops.sw_marginals_backward(scores, grad_marginals, gap, temp, lengths)


def marginals_hvp(scores: Tensor, v: Tensor, gap: float, temp: float,
                  lengths: Optional[Tensor] = None) -> Tensor:
    """Hessian-vector product of the value w.r.t. the scores: H . v."""
    _validate_derivative_vector(scores, v, name="tangent", primary_name="scores")
    return _ops.sw_marginals_hvp(scores, v, gap, temp, lengths)


def marginals_grad_gap(scores, gap, temp, lengths=None) -> Tensor:
    """d(map)/d(gap), [B, L1, L2]."""
    return _ops.sw_marginals_grad_gap(scores, gap, temp, lengths)


def marginals_grad_temp(scores, gap, temp, lengths=None) -> Tensor:
    """d(map)/d(temperature), [B, L1, L2]."""
    return _ops.sw_marginals_grad_temp(scores, gap, temp, lengths)


__all__ = [
    "forward",
    "forward_t",
    "value_grad_params",
    "marginals_backward",
    "marginals_hvp",
    "marginals_grad_gap",
    "marginals_grad_temp",
]
