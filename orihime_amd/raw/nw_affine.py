"""Raw tier of affine-gap Needleman-Wunsch (reference surface: orihime/raw/nw_affine.py):
the kernel bindings of ``ops/nw_affine`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.nw_affine import *  # noqa: F401,F403
from ..ops.nw_affine import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("nw_affine")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
