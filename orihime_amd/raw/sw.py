"""Raw tier of Smith-Waterman (reference surface: orihime/raw/sw.py):
the kernel bindings of ``ops/sw`` and parameter VJPs computed without an
autograd graph."""

from ..families import raw_vjp
from ..ops.sw import *  # noqa: F401,F403
from ..ops.sw import __all__ as _bindings

vjp_fields, vjp_one, vjp = raw_vjp("sw")

__all__ = ["vjp_fields", "vjp_one", "vjp", *_bindings]
