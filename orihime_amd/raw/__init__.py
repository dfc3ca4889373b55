"""Low-level tier (reference surface: orihime/raw/__init__.py): per family,
the kernel bindings and graph-free parameter VJPs."""

from . import nw as nw
from . import nw_affine as nw_affine
from . import sv_affine as sv_affine
from . import sv_linear as sv_linear
from . import sw as sw
from . import sw_affine as sw_affine

#: Type of a ``wrt`` selection: a tuple of parameter names.
VJPFields = tuple[str, ...]

__all__ = ["VJPFields", "sw", "sw_affine", "nw", "nw_affine", "sv_linear", "sv_affine"]
