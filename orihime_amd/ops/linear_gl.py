"""Linear-gap Needleman-Wunsch and Saigo-Vert kernel adapters
(orihime/ops/nw.py, orihime/ops/sv_linear.py).

One native call (``soft_nw_with_grads`` / ``soft_sv_linear_with_grads``)
gives value, map and the per-pair gap and temperature gradients, so the
backward role reuses them; value-only calls run the forward sweep alone.
Derivatives of the map (HVP, parameter sensitivities, map VJP) run the
family's affine forward-over-reverse tangent sweeps with one gap direction
moving both affine gap scores.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor

from .. import _ops
from ..operator import BackwardPass, ForwardPass, KernelCall, OperatorKernels
from ._common import check_temperature, resolve_lengths, scalar, warn_positive_gap


def _kernels(ns: str, soft: str) -> OperatorKernels:
    def prepare(call: KernelCall):
        gap, temp = call.params
        warn_positive_gap(gap, "gap")
        check_temperature(temp)
        return (scalar(gap), scalar(temp)), resolve_lengths(call)

    def forward(call: KernelCall) -> ForwardPass:
        params, lengths = prepare(call)
        value, mp, g_gap, g_temp = getattr(_ops.orihime, f"{soft}_with_grads")(
            call.primary, *params, lengths)
        return ForwardPass(value=value, saved_tensors=(mp, g_gap, g_temp))

    def backward(call: KernelCall, fwd: ForwardPass) -> BackwardPass:
        mp, g_gap, g_temp = fwd.saved_tensors
        return BackwardPass(marginals=mp, entropy=g_temp, param_grads=(g_gap, g_temp))

    def value_only(call: KernelCall) -> Tensor:
        params, lengths = prepare(call)
        return getattr(_ops, f"{ns}_value_only")(call.primary, *params, lengths)

    def map_backward(call: KernelCall, state, grad_map: Tensor) -> Tensor:
        params, lengths = prepare(call)
        return getattr(_ops, f"{ns}_marginals_hvp")(call.primary, grad_map.contiguous(),
                                                     *params, lengths)

    def param_backward(call: KernelCall, state, name: str, grad_map: Tensor | None) -> Tensor:
        params, lengths = prepare(call)
        field = getattr(_ops, f"{ns}_marginals_grad_{name}")(call.primary, *params, lengths)
        if grad_map is None:
            return field
        result = (grad_map * field).sum()
        param = call.params[0 if name == "gap" else 1]
        if isinstance(param, Tensor):
            return result.reshape(param.shape).to(device=param.device, dtype=param.dtype)
        return result.reshape(1)

    def map_vjp(call: KernelCall, grad_map: Tensor):
        params, lengths = prepare(call)
        gs, a, b = getattr(_ops, f"{ns}_marginals_backward")(call.primary, grad_map.contiguous(),
                                                             *params, lengths)
        return gs, (a, b)

    return OperatorKernels(forward=forward, backward=backward, map_backward=map_backward,
                           param_backward=param_backward, map_vjp=map_vjp,
                           value_only=value_only, config={"mask1": None, "mask2": None})


NW_KERNELS = _kernels("nw", "soft_nw")
SV_KERNELS = _kernels("sv_linear", "soft_sv_linear")


def _bindings(ns: str):
    def forward(scores: Tensor, gap: float, temp: float,
                lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
        """Value [B] and alignment marginals [B, L1, L2]."""
        return getattr(_ops, f"{ns}_forward")(scores, gap, temp, lengths)

    def forward_t(scores: Tensor, gap: Tensor, temp: Tensor,
                  lengths: Optional[Tensor] = None) -> Tuple[Tensor, Tensor]:
        """As ``forward`` with tensor parameters (differentiable in them)."""
        if lengths is None:
            B, L1, L2 = scores.shape
            lengths = torch.tensor([[L1, L2]] * B, dtype=torch.int32, device=scores.device)
        return getattr(_ops, f"{ns}_forward_t")(scores, gap, temp, lengths)

    def value_grad_params(scores: Tensor, gap: float, temp: float,
                          lengths: Optional[Tensor] = None):
        """(dS/dgap, dS/dtemperature), each [B]."""
        return getattr(_ops, f"{ns}_value_grad_params")(scores, gap, temp, lengths)

    return forward, forward_t, value_grad_params
