"""Raw tier of affine-gap Saigo-Vert (reference surface: orihime/raw/sv_affine.py):
the kernel bindings of ``ops/sv_affine`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.sv_affine import *  # noqa: F401,F403
from ..ops.sv_affine import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("sv_affine")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
