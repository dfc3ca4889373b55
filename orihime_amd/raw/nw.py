"""Raw tier of linear-gap Needleman-Wunsch (reference surface: orihime/raw/nw.py):
the kernel bindings of ``ops/nw`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.nw import *  # noqa: F401,F403
from ..ops.nw import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("nw")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
