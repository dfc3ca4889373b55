"""Type-stable no-graph VJP machinery (orihime/raw/_base.py)."""

from __future__ import annotations

from typing import Any, Callable, TypeAlias

import torch
from torch import Tensor

from ..operator import _validate_derivative_vector

VJPFields: TypeAlias = tuple[str, ...]


class _RawVJP:
    """Dispatch one operator's VJP fields: ``vjp_one`` returns one tensor,
    ``vjp`` a dictionary.  The cotangent must be contiguous FP32 with the
    map's shape and device; invalid cotangents fail before native dispatch."""

    def __init__(self, *, vjp_fields: VJPFields, vjp_one: Callable[..., Tensor]) -> None:
        if not isinstance(vjp_fields, tuple) or not vjp_fields:
            raise TypeError("vjp_fields must be a non-empty tuple")
        self.vjp_fields = vjp_fields
        self._vjp_one = vjp_one

    def _field(self, wrt: str) -> str:
        if not isinstance(wrt, str):
            raise TypeError("wrt must be a string")
        if wrt not in self.vjp_fields:
            raise ValueError(f"invalid wrt field {wrt!r}; expected one of {self.vjp_fields!r}")
        return wrt

    @staticmethod
    def _cot(primary: Tensor, cotangent: Tensor) -> Tensor:
        return _validate_derivative_vector(primary, cotangent, name="cotangent",
                                           primary_name="map", error_type=None)

    def vjp_one(self, *inputs: Any, wrt: str, cotangent: Tensor, **kwargs: Any) -> Tensor:
        field = self._field(wrt)
        cotangent = self._cot(inputs[0], cotangent)
        with torch.no_grad():
            res = self._vjp_one(*inputs, wrt=field, cotangent=cotangent, **kwargs)
        if not isinstance(res, Tensor):
            raise TypeError("raw vjp_one implementation must return a tensor")
        return res

    def vjp(self, *inputs: Any, wrt: VJPFields, cotangent: Tensor, **kwargs: Any) -> dict[str, Tensor]:
        if not isinstance(wrt, tuple):
            raise TypeError("wrt must be a tuple of field names")
        if len(set(wrt)) != len(wrt):
            raise ValueError("wrt must not contain duplicates")
        fields = tuple(self._field(f) for f in wrt)
        cotangent = self._cot(inputs[0], cotangent)
        return {f: self.vjp_one(*inputs, wrt=f, cotangent=cotangent, **kwargs) for f in fields}


__all__ = ["VJPFields"]
