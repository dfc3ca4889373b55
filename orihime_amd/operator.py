"""Operator engine for the soft Smith-Waterman family.

Behaviour of the reference engine (orihime/operator.py) that the alignment
operators rely on: input validation and its messages (:1948-2074; the
shared argument checks live in ``_checks``), mask handling with the finite
exclusion sentinel (:2178-2206), autocast / dtype handling (:958-984), and
the value / map / entropy observables with their autograd rules
(:1288-1716), the map VJP and the parameter sensitivities (:2312-2399).

Kernel work goes through the per-family adapters in ``ops/`` (named
primitives of ``_ops``); this module never computes a DP itself.
"""

from __future__ import annotations

from collections.abc import Callable, Mapping, Sequence
from dataclasses import dataclass, field
from types import MappingProxyType
from typing import Any

import torch
from torch import Tensor

from . import _checks
from ._checks import Param

# the names the adapters and autograd Functions below use
_validate_scalar_parameter_shape = _checks.scalar_shape
_validate_temperature = _checks.temperature
_validate_derivative_vector = _checks.derivative_vector
_validate_numerical_domain = _checks.numerical_domain
_normalize_masked = _checks.mask_to_sentinel
_graph_safe_assert = _checks.require
_LENGTHS_RANGE = "lengths entries must lie within the padded input shape"


# --------------------------------------------------------------------------
# Kernel adapter contract (orihime/operator.py:389-459)
# --------------------------------------------------------------------------


@dataclass(frozen=True)
class KernelCall:
    """Inputs shared by the per-family kernel callables."""

    primary: Tensor
    params: tuple[Param, ...]
    lengths: Tensor | None
    config: Mapping[str, Any]


@dataclass(frozen=True)
class ForwardPass:
    value: Tensor
    saved_tensors: tuple[Tensor, ...] = ()


@dataclass(frozen=True)
class BackwardPass:
    marginals: Tensor
    entropy: Tensor
    param_grads: tuple[Tensor, ...]
    saved_tensors: tuple[Tensor, ...] = ()


@dataclass(frozen=True)
class OperatorKernels:
    """Callable roles of one operator family.

    ``forward`` -> value (and optional saved state); ``backward`` -> map,
    entropy and per-pair parameter gradients; ``map_backward`` -> the
    score-space VJP of the map; ``map_param_backward`` -> the parameter VJPs
    of the map, all parameters at once; ``param_backward`` -> one parameter's
    raw sensitivity field (cotangent None) or its contraction.
    ``value_only`` -> value without map (no backward work).
    """

    forward: Callable[[KernelCall], ForwardPass]
    backward: Callable[[KernelCall, ForwardPass], BackwardPass]
    map_backward: Callable[[KernelCall, Any, Tensor], Tensor]
    param_backward: Callable[[KernelCall, Any, str, Tensor | None], Tensor]
    map_vjp: Callable[[KernelCall, Tensor], tuple[Tensor, tuple[Tensor, ...]]] | None = None
    value_only: Callable[[KernelCall], Tensor] | None = None
    config: Mapping[str, Any] = field(default_factory=dict)


def _to_compute(t: Tensor) -> Tensor:
    return t.to(torch.float32) if t.dtype in (torch.float16, torch.bfloat16) else t


def _restore(grad: Tensor | None, like: Tensor) -> Tensor | None:
    if grad is None:
        return None
    return grad.to(device=like.device, dtype=like.dtype)


def _contract_param(grad_value: Tensor, per_pair: Tensor, param: Param) -> Tensor | None:
    if not isinstance(param, Tensor):
        return None
    return (grad_value * per_pair).sum().reshape(param.shape).to(param.dtype)


# --------------------------------------------------------------------------
# Autograd functions.  Inputs: (op, primary, lengths, config_items, *params)
# --------------------------------------------------------------------------


def _call(op: "Operator", primary, lengths, config, params) -> KernelCall:
    return KernelCall(primary=primary, params=tuple(params), lengths=lengths,
                      config=MappingProxyType(dict(config)))


class _ValueFn(torch.autograd.Function):
    """Value observable.  Backward: map * grad (scores), sum(grad * dS/dtheta)
    (tensor parameters) — orihime/operator.py:1324-1450."""

    @staticmethod
    def forward(ctx, op, primary, lengths, config, *params):
        call = _call(op, primary, lengths, config, params)
        fwd = op._kernels.forward(call)
        need_params = any(isinstance(p, Tensor) and p.requires_grad for p in params)
        if need_params:
            bwd = op._kernels.backward(call, fwd)
            saved = (bwd.marginals, *bwd.param_grads)
        else:
            saved = (fwd.saved_tensors[0],)
        ctx.op = op
        ctx.config = config
        ctx.lengths = lengths
        ctx.n_params = len(params)
        ctx.param_is_tensor = tuple(isinstance(p, Tensor) for p in params)
        ctx.save_for_backward(primary, *saved, *[p for p in params if isinstance(p, Tensor)])
        ctx.has_param_grads = need_params
        return fwd.value

    @staticmethod
    def backward(ctx, grad_value):
        saved = ctx.saved_tensors
        primary = saved[0]
        marginals = saved[1]
        grad_value = _to_compute(grad_value)
        grads: list[Tensor | None] = [None, None, None, None]
        g_scores = None
        if ctx.needs_input_grad[1]:
            g_scores = _restore(marginals * grad_value.reshape(-1, 1, 1), primary)
        grads[1] = g_scores
        tensor_params = list(saved[2 + (ctx.n_params if ctx.has_param_grads else 0):])
        param_grads = saved[2:2 + ctx.n_params] if ctx.has_param_grads else ()
        out = []
        ti = 0
        for k in range(ctx.n_params):
            if not ctx.param_is_tensor[k]:
                out.append(None)
                continue
            p = tensor_params[ti]
            ti += 1
            if ctx.needs_input_grad[4 + k] and ctx.has_param_grads:
                out.append(_restore(_contract_param(grad_value, param_grads[k], p), p))
            else:
                out.append(None)
        return (None, g_scores, None, None, *out)


class _MapFn(torch.autograd.Function):
    """Map observable.  Backward: one forward-over-reverse pass gives the
    score VJP (H . grad_map) and every parameter VJP at once
    (orihime/operator.py:1617-1716 runs one HVP plus one U/W pass per
    parameter)."""

    @staticmethod
    def forward(ctx, op, primary, lengths, config, *params):
        call = _call(op, primary, lengths, config, params)
        fwd = op._kernels.forward(call)
        bwd = op._kernels.backward(call, fwd)
        ctx.op = op
        ctx.config = config
        ctx.lengths = lengths
        ctx.param_is_tensor = tuple(isinstance(p, Tensor) for p in params)
        ctx.params = tuple(p if not isinstance(p, Tensor) else None for p in params)
        ctx.save_for_backward(primary, *[p for p in params if isinstance(p, Tensor)])
        return bwd.marginals

    @staticmethod
    def backward(ctx, grad_map):
        saved = ctx.saved_tensors
        primary = saved[0]
        tensor_params = list(saved[1:])
        params = []
        ti = 0
        for is_t, p in zip(ctx.param_is_tensor, ctx.params):
            if is_t:
                params.append(tensor_params[ti].detach())
                ti += 1
            else:
                params.append(p)
        call = _call(ctx.op, _to_compute(primary.detach()), ctx.lengths, ctx.config, params)
        grad_map = _to_compute(grad_map).contiguous()
        need_params = any(
            ctx.param_is_tensor[k] and ctx.needs_input_grad[4 + k] for k in range(len(params))
        )
        if need_params and ctx.op._kernels.map_vjp is not None:
            g_scores, g_params = ctx.op._kernels.map_vjp(call, grad_map)
        else:
            g_scores = ctx.op._kernels.map_backward(call, None, grad_map)
            g_params = tuple(None for _ in params)
            if need_params:
                g_params = tuple(
                    ctx.op._kernels.param_backward(call, None, name, grad_map)
                    if ctx.param_is_tensor[k] and ctx.needs_input_grad[4 + k] else None
                    for k, name in enumerate(ctx.op.params)
                )
        out = []
        ti = 0
        for k in range(len(params)):
            if not ctx.param_is_tensor[k]:
                out.append(None)
                continue
            p = tensor_params[ti]
            ti += 1
            if ctx.needs_input_grad[4 + k] and g_params[k] is not None:
                out.append(_restore(g_params[k].reshape(p.shape), p))
            else:
                out.append(None)
        g = _restore(g_scores, primary) if ctx.needs_input_grad[1] else None
        return (None, g, None, None, *out)


def _entropy_param_error(op_name: str, param: str) -> NotImplementedError:
    public = {"gap": "gap_score", "gap_open": "gap_open_score", "gap_ext": "gap_extend_score",
              "temp": "temperature"}.get(param, param)
    return NotImplementedError(
        f"{op_name}_entropy is differentiable only w.r.t. its primary pair_scores input in "
        f"this release; the entropy gradient w.r.t. parameter {public} (a second derivative "
        "involving temperature) is not available -- see DERIVATIVE_COVERAGE.md. Detach the "
        "parameter or use finite differences."
    )


class _EntropyFn(torch.autograd.Function):
    """Entropy = dS/dT.  Backward w.r.t. the scores is grad * dmap/dT
    (orihime/operator.py:1063-1163); parameter directions are unsupported."""

    @staticmethod
    def forward(ctx, op, primary, lengths, config, *params):
        call = _call(op, primary, lengths, config, params)
        fwd = op._kernels.forward(call)
        bwd = op._kernels.backward(call, fwd)
        ctx.op = op
        ctx.config = config
        ctx.lengths = lengths
        ctx.param_is_tensor = tuple(isinstance(p, Tensor) for p in params)
        ctx.params = tuple(p if not isinstance(p, Tensor) else None for p in params)
        ctx.save_for_backward(primary, *[p for p in params if isinstance(p, Tensor)])
        return bwd.entropy.clone()

    @staticmethod
    def backward(ctx, grad_entropy):
        for k, name in enumerate(ctx.op.params):
            if ctx.param_is_tensor[k] and ctx.needs_input_grad[4 + k]:
                raise _entropy_param_error(ctx.op.name, name)
        if grad_entropy is None or not ctx.needs_input_grad[1]:
            return (None,) * (4 + len(ctx.op.params))
        saved = ctx.saved_tensors
        primary = saved[0]
        tensor_params = list(saved[1:])
        params = []
        ti = 0
        for is_t, p in zip(ctx.param_is_tensor, ctx.params):
            if is_t:
                params.append(tensor_params[ti].detach())
                ti += 1
            else:
                params.append(p)
        call = _call(ctx.op, _to_compute(primary.detach()), ctx.lengths, ctx.config, params)
        field_t = ctx.op._kernels.param_backward(call, None, "temp", None)
        g = _to_compute(grad_entropy).reshape(-1, 1, 1) * field_t
        return (None, _restore(g, primary), None, None, *([None] * len(params)))


# --------------------------------------------------------------------------
# Operator
# --------------------------------------------------------------------------


class Operator:
    """Autograd engine for one operator family (one tensor input plus
    scalar parameters, the last of which is the temperature)."""

    def __init__(
        self,
        name: str,
        params: Sequence[str],
        *,
        kernels: OperatorKernels,
        tensor_inputs: Sequence[str] = ("pair_scores",),
        tensor_shapes: Sequence[str] = ("B,L1,L2",),
        length_axes: Sequence[int] | None = (-2, -1),
    ) -> None:
        if not isinstance(name, str) or not name:
            raise ValueError("name must be a non-empty string")
        self.name = name
        self._params = tuple(params)
        self._tensor_inputs = tuple(tensor_inputs)
        self._shape = tuple(p.strip() for p in tensor_shapes[0].split(","))
        self._length_axes = None if length_axes is None else tuple(length_axes)
        if not isinstance(kernels, OperatorKernels):
            raise TypeError("kernels must be an OperatorKernels")
        self._kernels = kernels

    @property
    def params(self) -> tuple[str, ...]:
        return self._params

    @property
    def tensor_inputs(self) -> tuple[str, ...]:
        return self._tensor_inputs

    # -- validation ------------------------------------------------------

    def _check_call(self, primary: Tensor, params: tuple[Param, ...]) -> None:
        if not isinstance(primary, Tensor):
            raise TypeError(f"{self._tensor_inputs[0]} must be a tensor")
        if len(params) != len(self._params):
            raise TypeError(
                f"{self.name} expected {len(self._params)} parameters "
                f"({', '.join(self._params)}), got {len(params)}"
            )
        if not all(isinstance(p, (float, int, Tensor)) and not isinstance(p, bool) for p in params):
            raise TypeError("scoring parameters must be real numbers or tensors")

    def _validate_contract(self, primary: Tensor, params: tuple[Param, ...], lengths,
                           check_range: bool = True) -> None:
        name = self._tensor_inputs[0]
        expected = "[" + ", ".join(self._shape) + "]"
        if primary.ndim != len(self._shape):
            raise ValueError(f"{name} must have shape {expected}, got {tuple(primary.shape)}")
        if primary.shape[0] == 0:
            raise ValueError(f"{self.name} does not support an empty leading batch (B=0)")
        if not primary.is_floating_point():
            raise TypeError(f"{name} must have a floating-point dtype")
        if primary.dtype != torch.float32:
            raise TypeError(
                f"{name} must have dtype torch.float32 outside autocast; got {primary.dtype}"
            )
        for pname, p in zip(self._params, params):
            _validate_scalar_parameter_shape(p, pname)
            if not isinstance(p, Tensor):
                continue
            if not p.is_floating_point():
                raise TypeError(f"{pname} tensor must have a floating-point dtype")
            if p.device != primary.device:
                raise ValueError(
                    f"{pname} tensor must be on the same device as {name}, got "
                    f"{p.device} and {primary.device}"
                )
            if p.dtype != torch.float32:
                raise TypeError(
                    f"{pname} tensor must have dtype torch.float32 outside autocast; got {p.dtype}"
                )
        self._validate_lengths(primary, lengths, check_range)

    def _validate_lengths(self, primary: Tensor, lengths: Tensor | None,
                          check_range: bool = True) -> None:
        if lengths is None:
            return
        if self._length_axes is None:
            raise TypeError(f"{self.name} does not accept lengths")
        if not isinstance(lengths, Tensor):
            raise TypeError("lengths must be a tensor or None")
        expected = (primary.shape[0], len(self._length_axes))
        if tuple(lengths.shape) != expected:
            raise ValueError(f"lengths must have shape [B, 2], got {tuple(lengths.shape)}")
        if lengths.dtype != torch.int32:
            raise TypeError(f"lengths must have dtype torch.int32, got {lengths.dtype}")
        if lengths.device != primary.device:
            raise ValueError(
                "lengths must be on the same device as the primary input, got "
                f"{lengths.device} and {primary.device}"
            )
        if not lengths.is_contiguous():
            raise ValueError("lengths must be contiguous")
        if not check_range:
            return
        ok = torch.stack(
            tuple(
                (lengths[:, k] >= 0) & (lengths[:, k] <= primary.shape[ax])
                for k, ax in enumerate(self._length_axes)
            ),
            dim=-1,
        )
        _graph_safe_assert(ok, _LENGTHS_RANGE)

    def _mask_primary(self, primary: Tensor, mask: Tensor | None) -> Tensor:
        if not isinstance(primary, Tensor):
            raise TypeError(f"{self._tensor_inputs[0]} must be a tensor")
        if mask is None:
            return primary
        if not isinstance(mask, Tensor) or mask.dtype != torch.bool:
            got = mask.dtype if isinstance(mask, Tensor) else type(mask).__name__
            raise TypeError(f"mask must be a boolean tensor flagging excluded cells; got {got}")
        if tuple(mask.shape) != tuple(primary.shape):
            raise ValueError(
                f"mask shape {tuple(mask.shape)} must match "
                f"{self._tensor_inputs[0]} shape {tuple(primary.shape)}"
            )
        if mask.device != primary.device:
            raise ValueError(
                f"mask must be on the same device as {self._tensor_inputs[0]}, "
                f"got {mask.device} and {primary.device}"
            )
        return primary.masked_fill(mask, float("-inf"))

    def _config_items(self, config: Mapping[str, Any]) -> tuple[tuple[str, Any], ...]:
        unknown = set(config) - set(self._kernels.config)
        if unknown:
            raise TypeError(f"{self.name} got unexpected configuration: {', '.join(sorted(unknown))}")
        return tuple((k, config.get(k, d)) for k, d in self._kernels.config.items())

    def _prepare(self, primary, params, lengths, dtype, mask, config):
        primary = self._mask_primary(primary, mask)
        params = tuple(params)
        self._check_call(primary, params)
        if dtype is None and torch.is_autocast_enabled(primary.device.type):
            dtype = torch.float32
        if dtype is not None:
            if not isinstance(dtype, torch.dtype):
                raise TypeError("dtype must be a torch.dtype or None")
            if dtype != torch.float32:
                raise ValueError(
                    "orihime computes in float32; dtype= is an FP32-accumulation "
                    "escape hatch - only torch.float32 is supported"
                )
            if primary.is_floating_point():
                primary = primary.to(dtype)
            params = tuple(
                p.to(dtype) if isinstance(p, Tensor) and p.is_floating_point() else p
                for p in params
            )
        temp = params[self._params.index("temp")]
        # device tensors: one native scan does the range, domain and mask
        # checks with a single read-back (the torch checks below do each
        # with reductions of their own)
        fast = self._length_axes in ((-2, -1), (1, 2)) and _checks.scan_applies(primary, temp, lengths)
        self._validate_contract(primary, params, lengths, check_range=not fast)
        if fast:
            primary = _checks.scan(self._tensor_inputs[0], primary, self._params, params, lengths,
                                   _LENGTHS_RANGE)
        else:
            _validate_numerical_domain(self._tensor_inputs[0], primary, self._params, params)
            primary = _normalize_masked(primary, temp)
        return primary, params, self._config_items(config)

    @staticmethod
    def _needs_grad(primary: Tensor, params: tuple[Param, ...]) -> bool:
        return torch.is_grad_enabled() and (
            primary.requires_grad or any(isinstance(p, Tensor) and p.requires_grad for p in params)
        )

    # -- observables -----------------------------------------------------

    def __call__(self, primary, *params, lengths=None, dtype=None, mask=None, **config) -> Tensor:
        """Return the marginal map."""
        primary, params, items = self._prepare(primary, params, lengths, dtype, mask, config)
        if not self._needs_grad(primary, params):
            call = _call(self, primary, lengths, items, params)
            fwd = self._kernels.forward(call)
            return self._kernels.backward(call, fwd).marginals
        return _MapFn.apply(self, primary, lengths, items, *params)

    def value(self, primary, *params, lengths=None, dtype=None, mask=None, **config) -> Tensor:
        """Return the per-pair value."""
        primary, params, items = self._prepare(primary, params, lengths, dtype, mask, config)
        if not self._needs_grad(primary, params):
            call = _call(self, primary, lengths, items, params)
            if self._kernels.value_only is not None:
                return self._kernels.value_only(call)
            return self._kernels.forward(call).value
        return _ValueFn.apply(self, primary, lengths, items, *params)

    def entropy(self, primary, *params, lengths=None, dtype=None, mask=None, **config) -> Tensor:
        """Return the entropy observable (dS/dT)."""
        primary, params, items = self._prepare(primary, params, lengths, dtype, mask, config)
        if not self._needs_grad(primary, params):
            call = _call(self, primary, lengths, items, params)
            fwd = self._kernels.forward(call)
            return self._kernels.backward(call, fwd).entropy
        return _EntropyFn.apply(self, primary, lengths, items, *params)

    def _vjp_one(self, primary, *params, grad_map, wrt, lengths=None, dtype=None, mask=None,
                 **config) -> Tensor:
        """Contract one selected map derivative with ``grad_map``
        (orihime/operator.py:2312-2366)."""
        primary = self._mask_primary(primary, mask)
        valid = (*self._tensor_inputs, *self._params)
        if not isinstance(wrt, str):
            raise TypeError("wrt must be a string")
        if wrt not in valid:
            raise ValueError(f"invalid wrt field {wrt!r}; expected one of {valid!r}")
        if not isinstance(grad_map, Tensor):
            raise TypeError("grad_map (cotangent) must be a tensor")
        if grad_map.shape != primary.shape:
            raise ValueError(
                f"grad_map (cotangent) shape {tuple(grad_map.shape)} must match "
                f"the {self._tensor_inputs[0]} map shape {tuple(primary.shape)}"
            )
        if grad_map.device != primary.device:
            raise ValueError(
                "grad_map (cotangent) must be on the same device as the map, "
                f"got {grad_map.device} and {primary.device}"
            )
        if grad_map.dtype != torch.float32:
            raise TypeError(
                f"grad_map (cotangent) must have dtype torch.float32, got {grad_map.dtype}"
            )
        if not grad_map.is_contiguous():
            raise ValueError("grad_map (cotangent) must be contiguous")
        primary, params, items = self._prepare(primary, params, lengths, dtype, None, config)
        call = _call(self, primary, lengths, items, params)
        if wrt in self._tensor_inputs:
            return self._kernels.map_backward(call, None, grad_map)
        return self._kernels.param_backward(call, None, wrt, grad_map)

    def _sensitivity_one(self, primary, *params, wrt, lengths=None, dtype=None, mask=None,
                         **config) -> Tensor:
        """Raw map sensitivity to one scalar parameter (orihime/operator.py:2368-2399)."""
        primary = self._mask_primary(primary, mask)
        if not isinstance(wrt, str):
            raise TypeError("wrt must be a string")
        if wrt not in self._params:
            raise ValueError(f"invalid sensitivity field {wrt!r}; expected one of {self._params!r}")
        primary, params, items = self._prepare(primary, params, lengths, dtype, None, config)
        call = _call(self, primary, lengths, items, params)
        return self._kernels.param_backward(call, None, wrt, None)


__all__ = ["BackwardPass", "ForwardPass", "KernelCall", "Operator", "OperatorKernels"]
