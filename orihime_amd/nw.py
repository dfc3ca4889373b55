"""Needleman-Wunsch (global alignment) public functions, linear and affine
gaps (reference surface: orihime/nw.py:33-198); built from the family
records in ``families``.

Second order (map derivatives) is computed by the forward-over-reverse
tangent sweeps of the Needleman-Wunsch kernels."""

from __future__ import annotations

from .families import public_functions

nw, nw_value, nw_entropy = public_functions("nw")
nw_affine, nw_affine_value, nw_affine_entropy = public_functions("nw_affine")

__all__ = ["nw", "nw_value", "nw_entropy", "nw_affine", "nw_affine_value", "nw_affine_entropy"]
