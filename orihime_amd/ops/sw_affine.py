"""Affine-gap soft Smith-Waterman kernel adapter (orihime/ops/sw_affine.py).

Maps the engine roles onto the named primitives.  One fused native call
(``soft_sw_affine_with_grads``) returns value, map and the per-pair
parameter gradients, so the backward role reuses them instead of running the
forward and backward sweeps again (orihime/ops/sw_affine.py:133-146 reruns
them); value-only calls skip the map entirely; the map VJP gives the score
VJP and all three parameter VJPs in one forward-over-reverse pass.
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

_PARAM_NAMES = ("gap_open", "gap_ext", "temp")


def _prepare(call: KernelCall):
    go, ge, temp = call.params
    check_temperature(temp)
    warn_positive_gap(go, "gap_open")
    warn_positive_gap(ge, "gap_ext")
    return (scalar(go), scalar(ge), scalar(temp)), resolve_lengths(call)


def _forward_pass(call: KernelCall) -> ForwardPass:
    params, lengths = _prepare(call)
    value, mp, g_open, g_ext, g_temp = _ops.orihime.soft_sw_affine_with_grads(
        call.primary, *params, lengths)
    return ForwardPass(value=value, saved_tensors=(mp, g_open, g_ext, g_temp))


def _backward_pass(call: KernelCall, forward: ForwardPass) -> BackwardPass:
    mp, g_open, g_ext, g_temp = forward.saved_tensors
    return BackwardPass(marginals=mp, entropy=g_temp, param_grads=(g_open, g_ext, g_temp))


def _value_only(call: KernelCall) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sw_affine_value_only(call.primary, *params, lengths)


def _map_backward(call: KernelCall, state, grad_map: Tensor) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sw_affine_marginals_hvp(call.primary, grad_map.contiguous(), *params, lengths)


def _map_vjp(call: KernelCall, grad_map: Tensor):
    params, lengths = _prepare(call)
    gs, a, b, c = _ops.sw_affine_marginals_backward(call.primary, grad_map.contiguous(), *params,
                                                    lengths)
    return gs, (a, b, c)


_SENS = {
    "gap_open": "sw_affine_marginals_grad_gap_open",
    "gap_ext": "sw_affine_marginals_grad_gap_ext",
    "temp": "sw_affine_marginals_grad_temp",
}


def _param_backward(call: KernelCall, state, param_name: str, grad_map: Tensor | None) -> Tensor:
    params, lengths = _prepare(call)
    field = getattr(_ops, _SENS[param_name])(call.primary, *params, lengths)
    if grad_map is None:
        return field
    result = (grad_map * field).sum()
    param = call.params[_PARAM_NAMES.index(param_name)]
    if isinstance(param, Tensor):
        return result.reshape(param.shape).to(device=param.device, dtype=param.dtype)
    return result.reshape(1)


SW_AFFINE_KERNELS = OperatorKernels(
    forward=_forward_pass,
    backward=_backward_pass,
    map_backward=_map_backward,
    param_backward=_param_backward,
    map_vjp=_map_vjp,
    value_only=_value_only,
    config={"mask1": None, "mask2": None},
)


# Public low-level kernel bindings (orihime/ops/sw_affine.py:210-451).

def forward(scores: Tensor, gap_open: float, gap_ext: float, temp: float,
            lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Value [B] and alignment marginals [B, L1, L2] for scalar parameters."""
    return _ops.sw_affine_forward(scores, gap_open, gap_ext, temp, lengths)


def forward_t(scores: Tensor, gap_open: Tensor, gap_ext: Tensor, temp: Tensor,
              lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """As ``forward`` with tensor parameters (differentiable in them)."""
    if lengths is None:
        B, L1, L2 = scores.shape
        lengths = torch.tensor([[L1, L2]] * B, dtype=torch.int32, device=scores.device)
    return _ops.sw_affine_forward_t(scores, gap_open, gap_ext, temp, lengths)


def value_grad_params(scores: Tensor, gap_open: float, gap_ext: float, temp: float,
                      lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor, Tensor]:
    """Per-pair gradients of the value w.r.t. gap_open, gap_ext, temperature."""
    return _ops.sw_affine_value_grad_params(scores, gap_open, gap_ext, temp, lengths)


def marginals_backward(scores: Tensor, grad_marginals: Tensor, gap_open: float, gap_ext: float,
                       temp: float, lengths: Optional[Tensor] = None):
    """Gradients of <grad_marginals, map> w.r.t. scores [B,L1,L2] and the
    three parameters ([1] each)."""
    _validate_derivative_vector(scores, grad_marginals, name="cotangent", primary_name="scores")
    return _ops.sw_affine_marginals_backward(scores, grad_marginals, gap_open, gap_ext, temp,
                                             lengths)


def marginals_hvp(scores: Tensor, v: Tensor, gap_open: float, gap_ext: float, temp: float,
                  lengths: Optional[Tensor] = None) -> Tensor:
    """Hessian-vector product of the value w.r.t. the scores: H . v."""
    _validate_derivative_vector(scores, v, name="tangent", primary_name="scores")
    return _ops.sw_affine_marginals_hvp(scores, v, gap_open, gap_ext, temp, lengths)


def marginals_grad_gap_open(scores, gap_open, gap_ext, temp, lengths=None) -> Tensor:
    """d(map)/d(gap_open), [B, L1, L2]."""
    return _ops.sw_affine_marginals_grad_gap_open(scores, gap_open, gap_ext, temp, lengths)


def marginals_grad_gap_ext(scores, gap_open, gap_ext, temp, lengths=None) -> Tensor:
    """d(map)/d(gap_ext), [B, L1, L2]."""
    return _ops.sw_affine_marginals_grad_gap_ext(scores, gap_open, gap_ext, temp, lengths)


def marginals_grad_temp(scores, gap_open, gap_ext, temp, lengths=None) -> Tensor:
    """d(map)/d(temperature), [B, L1, L2]."""
    return _ops.sw_affine_marginals_grad_temp(scores, gap_open, gap_ext, temp, lengths)


__all__ = [
    "forward",
    "forward_t",
    "value_grad_params",
    "marginals_backward",
    "marginals_hvp",
    "marginals_grad_gap_open",
    "marginals_grad_gap_ext",
    "marginals_grad_temp",
]
