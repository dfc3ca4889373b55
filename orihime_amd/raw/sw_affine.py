"""Raw tier of affine-gap Smith-Waterman (reference surface: orihime/raw/sw_affine.py):
the kernel bindings of ``ops/sw_affine`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.sw_affine import *  # noqa: F401,F403
from ..ops.sw_affine import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("sw_affine")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
