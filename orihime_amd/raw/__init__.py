"""Low-level tier: no-graph VJPs plus the raw kernel bindings
(orihime/raw/__init__.py), for the Smith-Waterman, Needleman-Wunsch and
Saigo-Vert families."""

from ._base import VJPFields
from . import nw as nw
from . import nw_affine as nw_affine
from . import sv_affine as sv_affine
from . import sv_linear as sv_linear
from . import sw as sw
from . import sw_affine as sw_affine

__all__ = ["VJPFields", "sw", "sw_affine", "nw", "nw_affine", "sv_linear", "sv_affine"]
