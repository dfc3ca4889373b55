"""Finite named operations ``ops.sw`` / ``ops.sw_affine`` with ``forward``,
``backward`` and ``sensitivity`` (orihime/_public_ops.py:79-309)."""

from __future__ import annotations

from collections.abc import Callable, Mapping, Sequence
from inspect import Parameter, Signature, signature
from types import MappingProxyType, ModuleType
from typing import Any

import torch
from torch import Tensor

from .operator import Operator

_FORWARD_FIELDS = ("map", "value", "entropy")


def _selected_fields(output, allowed):
    if output is None:
        return allowed, False
    if isinstance(output, str):
        fields, single = (output,), True
    elif isinstance(output, Sequence):
        fields, single = tuple(output), False
    else:
        raise TypeError("output must be None, a field name, or a sequence of names")
    if not fields:
        raise ValueError("output must select at least one field")
    if not all(isinstance(f, str) for f in fields):
        raise TypeError("output field names must be strings")
    if len(set(fields)) != len(fields):
        raise ValueError("output must not contain duplicate field names")
    bad = tuple(f for f in fields if f not in allowed)
    if bad:
        raise ValueError(f"unknown output field {bad[0]!r}; expected one of {allowed!r}")
    return fields, single


def _result(fields, single, evaluate):
    if single:
        return evaluate(fields[0])
    return {f: evaluate(f) for f in fields}


def _with_kw(base: Signature, *params: Parameter) -> Signature:
    return base.replace(parameters=(*base.parameters.values(), *params))


class PublicOperation:
    """One algorithm's forward, backward, and sensitivity surface."""

    def __init__(self, *, name: str, operator: Operator, map_function: Callable[..., Tensor],
                 value_function: Callable[..., Tensor],
                 entropy_function: Callable[..., Tensor]) -> None:
        self.name = name
        self._operator = operator
        self._functions = MappingProxyType(
            {"map": map_function, "value": value_function, "entropy": entropy_function})
        self._sig = signature(map_function)
        names = tuple(self._sig.parameters)
        n_in, n_p = len(operator.tensor_inputs), len(operator.params)
        self.input_fields = names[:n_in]
        self.parameter_fields = names[n_in:n_in + n_p]
        self.backward_fields = (*self.input_fields, *self.parameter_fields)
        self.sensitivity_fields = self.parameter_fields
        self._internal = MappingProxyType(
            dict(zip(self.backward_fields, (*operator.tensor_inputs, *operator.params))))

        def forward(*args: Any, output=None, **kwargs: Any):
            """Return selected map, value, and entropy outputs."""
            fields, single = _selected_fields(output, _FORWARD_FIELDS)
            return _result(fields, single, lambda f: self._functions[f](*args, **kwargs))

        def backward(*args: Any, grad_map: Tensor, output=None, **kwargs: Any):
            """Contract a map cotangent against selected inputs and parameters."""
            fields, single = _selected_fields(output, self.backward_fields)
            prep = self._prepare(args, kwargs)

            def evaluate(field: str) -> Tensor:
                with torch.no_grad():
                    res = self._operator._vjp_one(
                        prep["primary"], *prep["params"], grad_map=grad_map,
                        wrt=self._internal[field], lengths=prep["lengths"], dtype=prep["dtype"],
                        mask=prep["mask"])
                if field not in self.parameter_fields:
                    return res
                original = prep["values"][field]
                if isinstance(original, Tensor):
                    return res.to(device=original.device, dtype=original.dtype).reshape(original.shape)
                return res.reshape(())

            return _result(fields, single, evaluate)

        def sensitivity(*args: Any, output=None, **kwargs: Any):
            """Return selected map sensitivities to scalar scoring parameters."""
            fields, single = _selected_fields(output, self.sensitivity_fields)
            prep = self._prepare(args, kwargs)

            def evaluate(field: str) -> Tensor:
                with torch.no_grad():
                    return self._operator._sensitivity_one(
                        prep["primary"], *prep["params"], wrt=self._internal[field],
                        lengths=prep["lengths"], dtype=prep["dtype"], mask=prep["mask"])

            return _result(fields, single, evaluate)

        forward.__name__ = "forward"
        forward.__qualname__ = f"ops.{name}.forward"
        forward.__signature__ = _with_kw(
            self._sig, Parameter("output", Parameter.KEYWORD_ONLY, default=None))
        backward.__name__ = "backward"
        backward.__qualname__ = f"ops.{name}.backward"
        backward.__signature__ = _with_kw(
            self._sig, Parameter("grad_map", Parameter.KEYWORD_ONLY),
            Parameter("output", Parameter.KEYWORD_ONLY, default=None))
        sensitivity.__name__ = "sensitivity"
        sensitivity.__qualname__ = f"ops.{name}.sensitivity"
        sensitivity.__signature__ = _with_kw(
            self._sig, Parameter("output", Parameter.KEYWORD_ONLY, default=None))
        self.forward = forward
        self.backward = backward
        self.sensitivity = sensitivity

    def _prepare(self, args, kwargs) -> Mapping[str, Any]:
        bound = self._sig.bind(*args, **kwargs)
        bound.apply_defaults()
        values = bound.arguments
        primary = values[self.input_fields[0]]
        if not isinstance(primary, Tensor):
            raise TypeError(f"{self.input_fields[0]} must be a tensor")
        return {
            "primary": primary,
            "params": tuple(values[f] for f in self.parameter_fields),
            "lengths": values.get("lengths"),
            "dtype": values.get("dtype"),
            "mask": values.get("mask"),
            "values": MappingProxyType(dict(values)),
        }

    def __repr__(self) -> str:
        return f"<orihime_amd.ops.{self.name}>"


def install_public_operations(namespace: ModuleType) -> None:
    """Populate ``orihime_amd.ops`` with the named operations."""
    from importlib import import_module

    swmod = import_module(".sw", __package__)
    k_sw = import_module(".ops.sw", __package__)
    k_aff = import_module(".ops.sw_affine", __package__)
    nwmod = import_module(".nw", __package__)
    k_nwa = import_module(".ops.nw_affine", __package__)
    svmod = import_module(".sv", __package__)
    k_sva = import_module(".ops.sv_affine", __package__)
    k_nwl = import_module(".ops.nw", __package__)
    k_svl = import_module(".ops.sv_linear", __package__)

    ops = {
        "sw": PublicOperation(name="sw", operator=swmod._sw_operator, map_function=swmod.sw,
                              value_function=swmod.sw_value, entropy_function=swmod.sw_entropy),
        "sw_affine": PublicOperation(name="sw_affine", operator=swmod._sw_affine_operator,
                                     map_function=swmod.sw_affine,
                                     value_function=swmod.sw_affine_value,
                                     entropy_function=swmod.sw_affine_entropy),
        "nw_affine": PublicOperation(name="nw_affine", operator=nwmod._nw_affine_operator,
                                     map_function=nwmod.nw_affine,
                                     value_function=nwmod.nw_affine_value,
                                     entropy_function=nwmod.nw_affine_entropy),
        "nw": PublicOperation(name="nw", operator=nwmod._nw_operator, map_function=nwmod.nw,
                              value_function=nwmod.nw_value, entropy_function=nwmod.nw_entropy),
        "sv": PublicOperation(name="sv", operator=svmod._sv_operator, map_function=svmod.sv,
                              value_function=svmod.sv_value, entropy_function=svmod.sv_entropy),
        "sv_affine": PublicOperation(name="sv_affine", operator=svmod._sv_affine_operator,
                                     map_function=svmod.sv_affine,
                                     value_function=svmod.sv_affine_value,
                                     entropy_function=svmod.sv_affine_entropy),
    }
    namespace._kernels = MappingProxyType({"sw": k_sw, "sw_affine": k_aff, "nw": k_nwl,
                                           "nw_affine": k_nwa, "sv_linear": k_svl,
                                           "sv_affine": k_sva})
    for name, op in ops.items():
        setattr(namespace, name, op)
    namespace.__all__ = list(ops)


__all__ = ["PublicOperation"]
