"""Saigo-Vert (affine gap) public functions (orihime/sv.py:33-214), linear and affine gaps."""

from __future__ import annotations

import torch
from torch import Tensor

from .operator import Operator, _document_public_function
from .ops.sv_linear import SV_KERNELS as _sv_kernels
from .ops.sv_affine import SV_AFFINE_KERNELS as _sv_affine_kernels

_sv_operator = Operator(
    "sv",
    params=("gap", "temp"),
    kernels=_sv_kernels,
    tensor_inputs=("pair_scores",),
    tensor_shapes=("B,L1,L2",),
    length_axes=(-2, -1),
)
_sv_affine_operator = Operator(
    "sv_affine",
    params=("gap_open", "gap_ext", "temp"),
    kernels=_sv_affine_kernels,
    tensor_inputs=("pair_scores",),
    tensor_shapes=("B,L1,L2",),
    length_axes=(-2, -1),
)


@_document_public_function
def sv(pair_scores: Tensor, *, gap_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
       lengths: Tensor | None = None, mask: Tensor | None = None,
       dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native linear-gap Saigo-Vert map (orihime/sv.py:33-123).

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_operator(pair_scores, gap_score, temperature, lengths=lengths, mask=mask,
                        dtype=dtype)


@_document_public_function
def sv_value(pair_scores: Tensor, *, gap_score: float | Tensor = 0.0,
             temperature: float | Tensor = 1.0, lengths: Tensor | None = None,
             mask: Tensor | None = None, dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native linear-gap Saigo-Vert value.

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_operator.value(pair_scores, gap_score, temperature, lengths=lengths, mask=mask,
                              dtype=dtype)


@_document_public_function
def sv_entropy(pair_scores: Tensor, *, gap_score: float | Tensor = 0.0,
               temperature: float | Tensor = 1.0, lengths: Tensor | None = None,
               mask: Tensor | None = None, dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native linear-gap Saigo-Vert entropy.

    Differentiation through the entropy needs the map's temperature
    sensitivity, which has no kernel yet (NotImplementedError).

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_operator.entropy(pair_scores, gap_score, temperature, lengths=lengths, mask=mask,
                                dtype=dtype)


@_document_public_function
def sv_affine(pair_scores: Tensor, *, gap_open_score: float | Tensor = 0.0,
              gap_extend_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
              lengths: Tensor | None = None, mask: Tensor | None = None,
              dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native affine Saigo-Vert map.

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_affine_operator(pair_scores, gap_open_score, gap_extend_score, temperature,
                               lengths=lengths, mask=mask, dtype=dtype)


@_document_public_function
def sv_affine_value(pair_scores: Tensor, *, gap_open_score: float | Tensor = 0.0,
                    gap_extend_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                    lengths: Tensor | None = None, mask: Tensor | None = None,
                    dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native affine Saigo-Vert value.

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_affine_operator.value(pair_scores, gap_open_score, gap_extend_score, temperature,
                                     lengths=lengths, mask=mask, dtype=dtype)


@_document_public_function
def sv_affine_entropy(pair_scores: Tensor, *, gap_open_score: float | Tensor = 0.0,
                      gap_extend_score: float | Tensor = 0.0, temperature: float | Tensor = 1.0,
                      lengths: Tensor | None = None, mask: Tensor | None = None,
                      dtype: torch.dtype | None = None) -> Tensor:
    """Return the score-native affine Saigo-Vert entropy.

    Differentiation through the entropy needs the map's temperature
    sensitivity, which has no kernel yet (NotImplementedError).

    Temperature must be finite and positive; finite scores and scoring
    parameters must satisfy ``abs(value) / temperature <= 80``.
    """
    return _sv_affine_operator.entropy(pair_scores, gap_open_score, gap_extend_score,
                                       temperature, lengths=lengths, mask=mask, dtype=dtype)


# The reference keeps the native family name as aliases (orihime/sv.py:117-122).
sv_linear = sv
sv_linear_value = sv_value
sv_linear_entropy = sv_entropy

__all__ = ["sv", "sv_value", "sv_entropy", "sv_linear", "sv_linear_value", "sv_linear_entropy",
           "sv_affine", "sv_affine_value", "sv_affine_entropy"]
