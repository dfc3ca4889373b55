"""Raw VJPs and kernel bindings for affine Saigo-Vert (orihime/raw/sv_affine.py)."""

from __future__ import annotations

from torch import Tensor

from ..ops.sv_affine import *  # noqa: F401,F403
from ..ops.sv_affine import __all__ as _kernel_all
from ..sv import _sv_affine_operator
from ._base import VJPFields, _RawVJP

vjp_fields: VJPFields = ("gap_open_score", "gap_extend_score", "temperature")
_INTERNAL = {"gap_open_score": "gap_open", "gap_extend_score": "gap_ext", "temperature": "temp"}


def _impl(pair_scores, *, wrt, cotangent, gap_open_score, gap_extend_score, temperature, lengths):
    return _sv_affine_operator._vjp_one(pair_scores, gap_open_score, gap_extend_score,
                                        temperature, grad_map=cotangent, wrt=_INTERNAL[wrt],
                                        lengths=lengths)


_raw = _RawVJP(vjp_fields=vjp_fields, vjp_one=_impl)


def vjp_one(pair_scores: Tensor, *, wrt: str, cotangent: Tensor,
            gap_open_score: float | Tensor = 0.0, gap_extend_score: float | Tensor = 0.0,
            temperature: float | Tensor = 1.0, lengths: Tensor | None = None) -> Tensor:
    """Return one selected affine-gap parameter VJP."""
    return _raw.vjp_one(pair_scores, wrt=wrt, cotangent=cotangent, gap_open_score=gap_open_score,
                        gap_extend_score=gap_extend_score, temperature=temperature,
                        lengths=lengths)


def vjp(pair_scores: Tensor, *, wrt: VJPFields, cotangent: Tensor,
        gap_open_score: float | Tensor = 0.0, gap_extend_score: float | Tensor = 0.0,
        temperature: float | Tensor = 1.0, lengths: Tensor | None = None) -> dict[str, Tensor]:
    """Return selected affine-gap parameter VJPs in a dictionary."""
    return _raw.vjp(pair_scores, wrt=wrt, cotangent=cotangent, gap_open_score=gap_open_score,
                    gap_extend_score=gap_extend_score, temperature=temperature, lengths=lengths)


__all__ = ["vjp_fields", "vjp_one", "vjp", *_kernel_all]
