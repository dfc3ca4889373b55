"""Smith-Waterman (local alignment) public functions, linear and affine gaps
(reference surface: orihime/sw.py:14-210); built from the family records in
``families``."""

from __future__ import annotations

from .families import public_functions

sw, sw_value, sw_entropy = public_functions("sw")
sw_affine, sw_affine_value, sw_affine_entropy = public_functions("sw_affine")

__all__ = ["sw", "sw_entropy", "sw_value", "sw_affine", "sw_affine_entropy", "sw_affine_value"]
