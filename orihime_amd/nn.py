"""``torch.nn`` layers for every alignment family, with per-parameter
learnability (reference surface: orihime/nn.py:51-390; class names and
constructor arguments are the reference's).

A layer stores its family's scalar parameters (gap scores and temperature)
as tensors: those named in ``learnable`` as ``nn.Parameter``s, the others as
persistent buffers, so both travel with ``state_dict`` and ``.to()``.
``layer(pair_scores, ...)`` returns the map, ``layer.value`` and
``layer.entropy`` the per-pair observables; extra keyword arguments
(``lengths``, ``mask``, ``dtype``) go to the public function unchanged.
"""

from __future__ import annotations

from collections.abc import Sequence
from inspect import Parameter, Signature
from typing import Any, ClassVar

import torch
from torch import Tensor, nn

from . import _checks
from .families import FAMILIES, Family, public_functions

_DEFAULT = {"temperature": 1.0}


def _stored(value: float | Tensor, name: str, device, dtype) -> Tensor:
    """A fresh floating tensor holding ``value`` (a number or a tensor)."""
    if isinstance(value, Tensor):
        t = value.detach().to(device=device, dtype=dtype if dtype is not None else (
            value.dtype if value.is_floating_point() else torch.get_default_dtype())).clone()
    elif isinstance(value, (int, float)) and not isinstance(value, bool):
        t = torch.tensor(float(value), device=device, dtype=dtype or torch.get_default_dtype())
    else:
        raise TypeError("operator parameters must be numbers or tensors")
    if not t.is_floating_point():
        raise TypeError("operator parameters must have a floating-point dtype")
    _checks.scalar_shape(t, name)
    return t


def _learnable_names(learnable: Sequence[str], allowed: tuple[str, ...]) -> frozenset[str]:
    if isinstance(learnable, (str, bytes)) or not isinstance(learnable, Sequence):
        raise TypeError("learnable must be a sequence of parameter names")
    names = tuple(learnable)
    if any(not isinstance(n, str) or not n for n in names):
        raise TypeError("learnable must contain non-empty strings")
    if len(set(names)) < len(names):
        raise ValueError("learnable must not contain duplicates")
    unknown = sorted(set(names) - set(allowed))
    if unknown:
        raise ValueError("unknown learnable parameter(s): " + ", ".join(unknown))
    return frozenset(names)


class AlignmentLayer(nn.Module):
    """A soft-alignment layer of one family (``family`` on the subclass)."""

    family: ClassVar[Family]

    def __init__(self, *, learnable: Sequence[str] = (), device=None, dtype=None,
                 **values: float | Tensor) -> None:
        super().__init__()
        names = self.family.parameters
        for extra in values:
            if extra not in names:
                raise TypeError(
                    f"{type(self).__name__}() got an unexpected keyword argument {extra!r}")
        trained = _learnable_names(learnable, names)
        for n in names:
            t = _stored(values.get(n, _DEFAULT.get(n, 0.0)), n, device, dtype)
            if n in trained:
                self.register_parameter(n, nn.Parameter(t))
            else:
                self.register_buffer(n, t, persistent=True)
        _checks.temperature(self.temperature)

    def _scoring(self) -> dict[str, Tensor]:
        _checks.temperature(self.temperature)
        return {n: getattr(self, n) for n in self.family.parameters}

    def forward(self, pair_scores: Tensor, **options: Any) -> Tensor:
        """Match-probability map, shaped like ``pair_scores``."""
        return public_functions(self.family.name)[0](pair_scores, **self._scoring(), **options)

    def value(self, pair_scores: Tensor, **options: Any) -> Tensor:
        """Soft alignment score per pair, ``[B]``."""
        return public_functions(self.family.name)[1](pair_scores, **self._scoring(), **options)

    def entropy(self, pair_scores: Tensor, **options: Any) -> Tensor:
        """Entropy of the alignment distribution per pair, ``[B]``."""
        return public_functions(self.family.name)[2](pair_scores, **self._scoring(), **options)

    def extra_repr(self) -> str:
        trained = tuple(n for n in self.family.parameters if n in self._parameters)
        return f"learnable={trained!r}"


def _layer_class(fam: Family) -> type[AlignmentLayer]:
    kw = Parameter.KEYWORD_ONLY
    sig = Signature([Parameter("self", Parameter.POSITIONAL_ONLY),
                     *(Parameter(n, kw, default=_DEFAULT.get(n, 0.0), annotation=float | Tensor)
                       for n in fam.parameters),
                     Parameter("learnable", kw, default=(), annotation=tuple[str, ...]),
                     Parameter("device", kw, default=None), Parameter("dtype", kw, default=None)])

    def __init__(self, **kwargs):
        AlignmentLayer.__init__(self, **kwargs)

    __init__.__signature__ = sig
    __init__.__qualname__ = f"{fam.layer}.__init__"
    doc = (f"Soft {fam.title} alignment layer; learnable subset of "
           f"{', '.join(fam.parameters)} chosen with ``learnable``.")
    return type(fam.layer, (AlignmentLayer,), {"family": fam, "__init__": __init__,
                                               "__doc__": doc, "__module__": __name__})


_CLASSES = {fam.layer: _layer_class(fam) for fam in FAMILIES.values()}
SmithWaterman = _CLASSES["SmithWaterman"]
SmithWatermanAffine = _CLASSES["SmithWatermanAffine"]
NeedlemanWunsch = _CLASSES["NeedlemanWunsch"]
NeedlemanWunschAffine = _CLASSES["NeedlemanWunschAffine"]
SaigoVertLinear = _CLASSES["SaigoVertLinear"]
SaigoVertAffine = _CLASSES["SaigoVertAffine"]
# short aliases the reference also exports
NW, NWAffine, SVLinear, SVAffine = NeedlemanWunsch, NeedlemanWunschAffine, SaigoVertLinear, SaigoVertAffine

__all__ = ["AlignmentLayer", "SmithWaterman", "SmithWatermanAffine", "SaigoVertLinear", "SVLinear",
           "SaigoVertAffine", "SVAffine", "NeedlemanWunsch", "NW", "NeedlemanWunschAffine",
           "NWAffine"]
