"""Stateful alignment modules with learnable parameters (orihime/nn.py:51-390):
Smith-Waterman, Saigo-Vert and Needleman-Wunsch, linear and affine gaps."""

from __future__ import annotations

from collections.abc import Callable, Mapping, Sequence
from typing import Any, ClassVar

import torch
from torch import Tensor, nn

from .operator import _validate_scalar_parameter_shape, _validate_temperature
from .nw import nw, nw_affine, nw_affine_entropy, nw_affine_value, nw_entropy, nw_value
from .sv import sv, sv_affine, sv_affine_entropy, sv_affine_value, sv_entropy, sv_value
from .sw import sw, sw_affine, sw_affine_entropy, sw_affine_value, sw_entropy, sw_value

ParamValue = float | Tensor


class _OperatorModule(nn.Module):
    """Thin stateful wrapper around a map/value/entropy function triple.
    Names listed in ``learnable`` become ``nn.Parameter``s; the other model
    parameters are persistent buffers."""

    _param_names: ClassVar[tuple[str, ...]] = ()

    def __init__(self, *, params: Mapping[str, ParamValue], map_function: Callable[..., Tensor],
                 value_function: Callable[..., Tensor], entropy_function: Callable[..., Tensor],
                 learnable: Sequence[str] = (), device=None, dtype=None) -> None:
        super().__init__()
        if set(params) != set(self._param_names):
            missing = set(self._param_names) - set(params)
            unknown = set(params) - set(self._param_names)
            details = []
            if missing:
                details.append("missing " + ", ".join(sorted(missing)))
            if unknown:
                details.append("unknown " + ", ".join(sorted(unknown)))
            raise TypeError("invalid parameter mapping: " + "; ".join(details))
        if isinstance(learnable, (str, bytes)):
            raise TypeError("learnable must be a sequence of parameter names")
        names = tuple(learnable)
        if not all(isinstance(n, str) and n for n in names):
            raise TypeError("learnable must contain non-empty strings")
        if len(set(names)) != len(names):
            raise ValueError("learnable must not contain duplicates")
        unknown_l = set(names) - set(self._param_names)
        if unknown_l:
            raise ValueError("unknown learnable parameter(s): " + ", ".join(sorted(unknown_l)))
        object.__setattr__(self, "_map_function", map_function)
        object.__setattr__(self, "_value_function", value_function)
        object.__setattr__(self, "_entropy_function", entropy_function)
        for name in self._param_names:
            t = self._parameter_tensor(params[name], device=device, dtype=dtype)
            _validate_scalar_parameter_shape(t, name)
            if name in names:
                self.register_parameter(name, nn.Parameter(t))
            else:
                self.register_buffer(name, t, persistent=True)
        self._validate_constraints()

    @staticmethod
    def _parameter_tensor(value: ParamValue, *, device, dtype) -> Tensor:
        if isinstance(value, Tensor):
            if not value.is_floating_point() and dtype is None:
                dtype = torch.get_default_dtype()
            t = value.detach().to(device=device, dtype=dtype).clone()
        elif isinstance(value, (int, float)):
            t = torch.tensor(value, device=device, dtype=dtype or torch.get_default_dtype())
        else:
            raise TypeError("operator parameters must be numbers or tensors")
        if not t.is_floating_point():
            raise TypeError("operator parameters must have a floating-point dtype")
        return t

    def _validate_constraints(self) -> None:
        temperature = getattr(self, "temperature", None)
        if temperature is not None:
            _validate_temperature(temperature)

    def _parameter_kwargs(self) -> dict[str, Tensor]:
        self._validate_constraints()
        return {name: getattr(self, name) for name in self._param_names}

    def forward(self, *inputs: Tensor, **structural: Any) -> Tensor:
        """Return the operator map."""
        return self._map_function(*inputs, **self._parameter_kwargs(), **structural)

    def value(self, *inputs: Tensor, **structural: Any) -> Tensor:
        """Return the operator value."""
        return self._value_function(*inputs, **self._parameter_kwargs(), **structural)

    def entropy(self, *inputs: Tensor, **structural: Any) -> Tensor:
        """Return the operator entropy."""
        return self._entropy_function(*inputs, **self._parameter_kwargs(), **structural)

    def extra_repr(self) -> str:
        learnable = tuple(n for n in self._param_names if n in self._parameters)
        return f"learnable={learnable!r}"


class SmithWaterman(_OperatorModule):
    """Stateful Smith-Waterman layer with per-parameter learnability."""

    _param_names = ("gap_score", "temperature")

    def __init__(self, *, gap_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                 learnable: tuple[str, ...] = (), device=None, dtype=None) -> None:
        super().__init__(params={"gap_score": gap_score, "temperature": temperature},
                         map_function=sw, value_function=sw_value, entropy_function=sw_entropy,
                         learnable=learnable, device=device, dtype=dtype)


class SmithWatermanAffine(_OperatorModule):
    """Affine Smith-Waterman layer with independently learnable parameters."""

    _param_names = ("gap_open_score", "gap_extend_score", "temperature")

    def __init__(self, *, gap_open_score: float | Tensor = 0.0,
                 gap_extend_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                 learnable: tuple[str, ...] = (), device=None, dtype=None) -> None:
        super().__init__(params={"gap_open_score": gap_open_score,
                                 "gap_extend_score": gap_extend_score, "temperature": temperature},
                         map_function=sw_affine, value_function=sw_affine_value,
                         entropy_function=sw_affine_entropy, learnable=learnable, device=device,
                         dtype=dtype)


class _Linear(_OperatorModule):
    _param_names = ("gap_score", "temperature")
    _functions: ClassVar[tuple] = ()

    def __init__(self, *, gap_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                 learnable: tuple[str, ...] = (), device=None, dtype=None) -> None:
        m, v, e = self._functions
        super().__init__(params={"gap_score": gap_score, "temperature": temperature},
                         map_function=m, value_function=v, entropy_function=e,
                         learnable=learnable, device=device, dtype=dtype)


class _Affine(_OperatorModule):
    _param_names = ("gap_open_score", "gap_extend_score", "temperature")
    _functions: ClassVar[tuple] = ()

    def __init__(self, *, gap_open_score: float | Tensor = 0.0,
                 gap_extend_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                 learnable: tuple[str, ...] = (), device=None, dtype=None) -> None:
        m, v, e = self._functions
        super().__init__(params={"gap_open_score": gap_open_score,
                                 "gap_extend_score": gap_extend_score, "temperature": temperature},
                         map_function=m, value_function=v, entropy_function=e,
                         learnable=learnable, device=device, dtype=dtype)


class SaigoVertLinear(_Linear):
    """Canonical linear-gap Saigo-Vert layer."""

    _functions = (staticmethod(sv), staticmethod(sv_value), staticmethod(sv_entropy))


class SaigoVertAffine(_Affine):
    """Canonical affine-gap Saigo-Vert layer."""

    _functions = (staticmethod(sv_affine), staticmethod(sv_affine_value),
                  staticmethod(sv_affine_entropy))


class NeedlemanWunsch(_Linear):
    """Score-native soft Needleman-Wunsch global alignment."""

    _functions = (staticmethod(nw), staticmethod(nw_value), staticmethod(nw_entropy))


class NeedlemanWunschAffine(_Affine):
    """Stateful affine Needleman-Wunsch layer."""

    _functions = (staticmethod(nw_affine), staticmethod(nw_affine_value),
                  staticmethod(nw_affine_entropy))


SVLinear = SaigoVertLinear
SVAffine = SaigoVertAffine
NW = NeedlemanWunsch
NWAffine = NeedlemanWunschAffine

__all__ = ["SmithWaterman", "SmithWatermanAffine", "SaigoVertLinear", "SVLinear",
           "SaigoVertAffine", "SVAffine", "NeedlemanWunsch", "NW", "NeedlemanWunschAffine",
           "NWAffine"]
