"""Saigo-Vert public functions, linear and affine gaps (reference surface:
orihime/sv.py:33-214); built from the family records in ``families``.

``sv_linear*`` name the linear-gap functions by their native family name
(the reference exports both spellings, orihime/sv.py:117-122).  Second
order (map derivatives) is computed by the forward-over-reverse tangent
sweeps of the Saigo-Vert kernels."""

from __future__ import annotations

from .families import public_functions

sv, sv_value, sv_entropy = public_functions("sv")
sv_affine, sv_affine_value, sv_affine_entropy = public_functions("sv_affine")
sv_linear, sv_linear_value, sv_linear_entropy = sv, sv_value, sv_entropy

__all__ = ["sv", "sv_value", "sv_entropy", "sv_linear", "sv_linear_value", "sv_linear_entropy",
           "sv_affine", "sv_affine_value", "sv_affine_entropy"]
