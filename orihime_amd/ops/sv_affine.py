"""Affine-gap soft Saigo-Vert kernel adapter
(orihime/ops/sv_affine.py).

One native call (``soft_sv_affine_with_grads``: forward and reverse sweep)
gives value, map and the per-pair parameter gradients, so the backward role
reuses them instead of running ``sv_affine_value_grad_params`` again
(orihime/ops/sv_affine.py:221-235 does); value-only calls run the forward
sweep alone.  Derivatives of the map (HVP, parameter sensitivities, map
VJP) run the Saigo-Vert forward-over-reverse tangent sweeps
(``ohm_sv_affine_tangent``).
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor

from .. import _ops
from ..operator import BackwardPass, ForwardPass, KernelCall, OperatorKernels
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
    value, mp, g_open, g_ext, g_temp = _ops.orihime.soft_sv_affine_with_grads(
        call.primary, *params, lengths)
    return ForwardPass(value=value, saved_tensors=(mp, g_open, g_ext, g_temp))


def _backward_pass(call: KernelCall, forward: ForwardPass) -> BackwardPass:
    mp, g_open, g_ext, g_temp = forward.saved_tensors
    return BackwardPass(marginals=mp, entropy=g_temp, param_grads=(g_open, g_ext, g_temp))


def _value_only(call: KernelCall) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sv_affine_value_only(call.primary, *params, lengths)


def _map_backward(call: KernelCall, state, grad_map: Tensor) -> Tensor:
    params, lengths = _prepare(call)
    return _ops.sv_affine_marginals_hvp(call.primary, grad_map.contiguous(), *params, lengths)


_SENS = {
    "gap_open": "sv_affine_marginals_grad_gap_open",
    "gap_ext": "sv_affine_marginals_grad_gap_ext",
    "temp": "sv_affine_marginals_grad_temp",
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


def _map_vjp(call: KernelCall, grad_map: Tensor):
    params, lengths = _prepare(call)
    gs, a, b, c = _ops.sv_affine_marginals_backward(call.primary, grad_map.contiguous(), *params,
                                                 lengths)
    return gs, (a, b, c)


SV_AFFINE_KERNELS = OperatorKernels(
    forward=_forward_pass,
    backward=_backward_pass,
    map_backward=_map_backward,
    param_backward=_param_backward,
    map_vjp=_map_vjp,
    value_only=_value_only,
    config={"mask1": None, "mask2": None},
)


# Public low-level kernel bindings (orihime/ops/sv_affine.py).

def forward(scores: Tensor, gap_open: float, gap_ext: float, temp: float,
            lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """Value [B] and alignment marginals [B, L1, L2] for scalar parameters."""
    return _ops.sv_affine_forward(scores, gap_open, gap_ext, temp, lengths)


def forward_t(scores: Tensor, gap_open: Tensor, gap_ext: Tensor, temp: Tensor,
              lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
    """As ``forward`` with tensor parameters (differentiable in them)."""
    if lengths is None:
        B, L1, L2 = scores.shape
        lengths = torch.tensor([[L1, L2]] * B, dtype=torch.int32, device=scores.device)
    return _ops.sv_affine_forward_t(scores, gap_open, gap_ext, temp, lengths)


def value_grad_params(scores: Tensor, gap_open: float, gap_ext: float, temp: float,
                      lengths: Optional[Tensor] = None):
    """(dS/dgap_open, dS/dgap_ext, dS/dtemperature), each [B]."""
    return _ops.sv_affine_value_grad_params(scores, gap_open, gap_ext, temp, lengths)


__all__ = ["forward", "forward_t", "value_grad_params"]
