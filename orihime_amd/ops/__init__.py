"""Named operations (``ops.sw``, ``ops.sw_affine``) and, under ``_kernels``,
the low-level kernel adapter modules.  Populated by the package init."""
