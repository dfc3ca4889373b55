"""Linear-gap Needleman-Wunsch kernel adapter and low-level bindings
(orihime/ops/nw.py)."""

from .linear_gl import NW_KERNELS, _bindings

forward, forward_t, value_grad_params = _bindings("nw")

__all__ = ["NW_KERNELS", "forward", "forward_t", "value_grad_params"]
