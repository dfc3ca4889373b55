"""Raw tier of linear-gap Saigo-Vert (reference surface: orihime/raw/sv_linear.py):
the kernel bindings of ``ops/sv_linear`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.sv_linear import *  # noqa: F401,F403
from ..ops.sv_linear import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("sv")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
