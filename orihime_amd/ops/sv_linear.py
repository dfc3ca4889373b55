"""Linear-gap Saigo-Vert kernel adapter and low-level bindings
(orihime/ops/sv_linear.py)."""

from .linear_gl import SV_KERNELS, _bindings

forward, forward_t, value_grad_params = _bindings("sv_linear")

__all__ = ["SV_KERNELS", "forward", "forward_t", "value_grad_params"]
