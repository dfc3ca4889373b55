"""orihime_amd — soft Smith-Waterman for AMD Instinct MI355X (gfx950).

Drop-in for the soft Smith-Waterman path of orihime: the same public
functions (``sw``, ``sw_value``, ``sw_entropy``, ``sw_affine``,
``sw_affine_value``, ``sw_affine_entropy``; the global-alignment ``nw`` and
``nw_affine`` families; the Saigo-Vert ``sv`` and ``sv_affine`` families,
each with ``_value`` and ``_entropy``), ``nn`` modules, ``ops``
named operations, ``raw`` low-level tier and ``torch.ops.orihime.*``
primitives, computed by hand-written gfx950 HIP kernels (GPU tensors) or
the library's host implementation (CPU tensors), both behind the C ABI in
include/ohm_sw.h.  Packed ragged batches are in ``ragged``.
"""

__version__ = "0.1.0"

from . import _ops as _ops  # registers torch.ops.orihime.* (loads libohm_sw.so)
from .sw import sw, sw_affine, sw_affine_entropy, sw_affine_value, sw_entropy, sw_value
from .nw import nw, nw_affine, nw_affine_entropy, nw_affine_value, nw_entropy, nw_value
from .sv import (sv, sv_affine, sv_affine_entropy, sv_affine_value, sv_entropy, sv_linear,
                 sv_linear_entropy, sv_linear_value, sv_value)
from . import nn as nn
from . import raw as raw
from . import ops as ops
from . import ragged as ragged
from .families import install_operations as _install_operations

_install_operations(ops)

__all__ = [
    "__version__",
    "sw",
    "sw_value",
    "sw_entropy",
    "sw_affine",
    "sw_affine_value",
    "sw_affine_entropy",
    "nw",
    "nw_value",
    "nw_entropy",
    "nw_affine",
    "nw_affine_value",
    "nw_affine_entropy",
    "sv",
    "sv_value",
    "sv_entropy",
    "sv_linear",
    "sv_linear_value",
    "sv_linear_entropy",
    "sv_affine",
    "sv_affine_value",
    "sv_affine_entropy",
    "nn",
    "ops",
    "raw",
    "ragged",
]
