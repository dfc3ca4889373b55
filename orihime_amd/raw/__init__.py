"""Low-level tier: no-graph VJPs plus the raw kernel bindings
(orihime/raw/__init__.py), for the Smith-Waterman family."""

from ._base import VJPFields
from . import sw as sw
from . import sw_affine as sw_affine

__all__ = ["VJPFields", "sw", "sw_affine"]
