"""Named primitives of the soft Smith-Waterman path (drop-in boundary).

Registers the reference's dispatcher schemas under the same ``orihime::``
namespace and names (src/sw_affine/registry.cpp:27-42,
src/sw/registry.cpp:30-44) with CPU and CUDA (HIP) kernels that call the C
ABI of libohm_sw.so through ``_lib``; fake (meta) kernels for torch.compile
(orihime/_compile.py:25-134); autograd for the forward primitives
(src/sw_affine/torch_hip.cpp:126-400); and batch-folding vmap rules keyed on
the schema argument names (orihime/_ops.py:237-406).

Module-level wrappers (``sw_affine_forward`` ...) have the call signatures
orihime/_ops.py:434-449 exposes to the kernel adapters.

Two primitives are additions with no reference counterpart:
``*_value_only`` (value without the map, for value calls that need no
gradient) and ``*_marginals_backward_pairs`` (the map VJP with per-pair,
unsummed parameter contractions, which the vmap rule of
``*_marginals_backward`` folds per mapped element).
"""

from __future__ import annotations

from typing import Optional

import torch
from torch import Tensor

from . import _checks, _lib

_LIB = torch.library.Library("orihime", "FRAGMENT")

# --------------------------------------------------------------------------
# Validation at the native boundary (src/sw_affine/torch_hip.cpp:35-99,
# src/sw_affine/torch_cpu.cpp:40-107)
# --------------------------------------------------------------------------


def _check_scores(scores: Tensor, short: bool = False) -> None:
    if scores.dim() != 3:
        raise RuntimeError("scores must be 3D" if short else "scores must be 3D (B, L1, L2)")
    if scores.dtype != torch.float32:
        raise RuntimeError("scores must be float32")
    if not scores.is_contiguous():
        raise RuntimeError("scores must be contiguous")


def _check_vector(scores: Tensor, vec: Tensor, name: str) -> Tensor:
    if vec.shape != scores.shape:
        raise RuntimeError(f"scores and {name} must have same shape")
    if vec.device != scores.device:
        raise RuntimeError(f"{name} must be on same device as scores")
    if vec.dtype != torch.float32:
        raise RuntimeError(f"{name} must have dtype torch.float32")
    if not vec.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    return vec


def _check_lengths(lengths: Optional[Tensor], scores: Tensor) -> Optional[Tensor]:
    if lengths is None:
        return None
    B, L1, L2 = scores.shape
    if lengths.device != scores.device:
        raise RuntimeError("lengths must be on same device as scores")
    if not lengths.is_contiguous():
        raise RuntimeError("lengths must be contiguous")
    if lengths.dim() != 2 or lengths.shape[0] != B or lengths.shape[1] != 2:
        raise RuntimeError("lengths must have shape [B, 2]")
    if lengths.dtype != torch.int32:
        raise RuntimeError("lengths must be int32")
    # The reference bindings copy lengths to the host to range-check them
    # (src/sw_affine/torch_hip.cpp:69-99).  Skipped while a HIP graph is
    # being captured (nothing can be read back; the kernels clamp each length
    # into [0, max]) and when the engine's input scan has just checked these
    # lengths against this shape.
    if (lengths.numel() and not _checks.deferred(lengths)
            and not _checks.lengths_checked(lengths, L1, L2)):
        lo = lengths.amin(dim=0)
        hi = lengths.amax(dim=0)
        bounds = torch.stack([lo, hi]).cpu()
        if bounds.min() < 0 or bounds[1, 0] > L1 or bounds[1, 1] > L2:
            host = lengths.cpu()
            for b in range(B):
                l1, l2 = int(host[b, 0]), int(host[b, 1])
                if not 0 <= l1 <= L1:
                    raise RuntimeError(f"lengths[{b},0] must be between 0 and {L1}, got {l1}")
                if not 0 <= l2 <= L2:
                    raise RuntimeError(f"lengths[{b},1] must be between 0 and {L2}, got {l2}")
    return lengths


# Shape guards of the reference bindings.  The reference keeps (L1+1) x (L2+1)
# DP tables (three planes for the affine models) indexed with int32 and
# refuses, before allocating anything, shapes whose table would overflow that
# index (src/sw_affine/torch_cpu.cpp:39-62, torch_hip.cpp:44-67;
# src/sw/torch_cpu.cpp:38-62, torch_hip.cpp:40-66; src/nw/torch_cpu.cpp:43-61;
# src/nw_affine/torch_cpu.cpp:32-62, torch_hip.cpp:32-80;
# src/sv_affine/torch_cpu.cpp:39-62, torch_hip.cpp:36-66;
# src/sv_linear/torch_cpu.cpp:33-60, torch_hip.cpp:35-62).  These sweeps keep
# no such table, but a drop-in refuses the same shapes with the same messages,
# so a caller sees the same contract.  Zero-storage tensors such as
# (1, 0, 715827882) reach these checks without any allocation.
_I32 = 2**31 - 1


def _guard_shape(family: str, scores: Tensor) -> None:
    B, L1, L2 = (int(x) for x in scores.shape)
    rows, cols = L1 + 1, L2 + 1
    cells = rows * cols
    gpu = scores.device.type != "cpu"
    for n, what in ((B, "batch"), (L1, "L1"), (L2, "L2")):
        if n > _I32:
            raise RuntimeError(f"scores {what} dimension is too large for {family}")
    if family in ("sw_affine", "sv_affine"):
        if gpu:
            api = "CUDA" if family == "sw_affine" else "HIP"
            if L1 > _I32 - L2:
                raise RuntimeError(
                    f"{family} dimensions too large: L1 + L2 exceeds {api} int index range")
            if 3 * cells > _I32:
                raise RuntimeError(f"{family} DP table too large: 3 * (L1 + 1) * (L2 + 1) "
                                   f"exceeds {api} int index range")
        elif 3 * cells > _I32:
            raise RuntimeError(
                f"{family} CPU DP table is too large: 3 * (L1 + 1) * (L2 + 1) = {3 * cells}")
    elif family == "nw_affine":
        if gpu:
            if L1 > _I32 - L2:
                raise RuntimeError(f"scores dimensions L1 + L2 must fit in int32, got {L1} + {L2}")
            if 3 * cells > _I32:
                raise RuntimeError(
                    f"nw_affine alpha workspace size must fit in int32, got {3 * cells}")
        elif 3 * cells > _I32:
            raise RuntimeError(f"nw_affine CPU DP workspace is too large: 3 * (L1 + 1) * "
                               f"(L2 + 1) = {3 * cells} exceeds {_I32}")
    elif family == "sw":
        if cells > _I32:
            if gpu:
                raise RuntimeError(f"SW CUDA DP table is too large for 32-bit kernel indexing: "
                                   f"(L1 + 1) * (L2 + 1) = {cells} exceeds {_I32}")
            raise RuntimeError(f"SW CPU DP table size exceeds supported int32 range: "
                               f"(L1 + 1) * (L2 + 1) must be <= {_I32}, got {rows} * {cols}")
    elif family == "sv_linear":
        if cells > _I32 // 3:
            if gpu:
                raise RuntimeError(f"SV linear HIP DP table is too large for 32-bit kernel "
                                   f"indexing: 3 * (L1 + 1) * (L2 + 1) = {3 * cells} exceeds {_I32}")
            raise RuntimeError(f"SV linear CPU DP table size exceeds supported int32 range: "
                               f"3 * (L1 + 1) * (L2 + 1) must be <= {_I32}, got {rows} * {cols}")
    elif family == "nw" and not gpu and cells > _I32:
        raise RuntimeError(f"NW alpha workspace size is too large: ({L1} + 1) * ({L2} + 1) "
                           f"exceeds {_I32}")


def _family_of(schema_name: str) -> str:
    name = schema_name[5:] if schema_name.startswith("soft_") else schema_name
    for fam in ("sw_affine", "nw_affine", "sv_affine", "sv_linear"):
        if name.startswith(fam):
            return fam
    return "sw" if name.startswith("sw") else "nw"


def _guarded(schema_name: str, impl):
    family = _family_of(schema_name)

    def run(scores, *args):
        if isinstance(scores, Tensor) and scores.dim() == 3:
            _guard_shape(family, scores)
        return impl(scores, *args)

    return run


def _scalar(t: Tensor, name: str) -> float:
    if t.numel() != 1:
        raise RuntimeError(f"{name} must be a scalar tensor")
    return _checks.host_scalar(t, name)


def _vec(scores: Tensor, n: int) -> Tensor:
    return torch.empty(n, dtype=torch.float32, device=scores.device)


# --------------------------------------------------------------------------
# Affine gap implementations
# --------------------------------------------------------------------------


def _aff_forward(scores, go, ge, temp, lengths, want_map=True, want_grads=False):
    _check_scores(scores)
    ln = _check_lengths(lengths, scores)
    B = scores.shape[0]
    value = _vec(scores, B)
    mp = torch.empty_like(scores) if want_map else None
    g = [_vec(scores, B) for _ in range(3)] if want_grads else [None] * 3
    _lib.affine_forward(scores, ln, None, go, ge, temp, value=value, map=mp,
                        grad_open=g[0], grad_ext=g[1], grad_temp=g[2])
    return value, mp, g


def _aff_tangent(scores, go, ge, temp, lengths, direction=None, dgo=0.0, dge=0.0, dT=0.0,
                 contract=False):
    ln = _check_lengths(lengths, scores)
    B = scores.shape[0]
    mt = torch.empty_like(scores)
    c = torch.empty(B, 3, dtype=torch.float32, device=scores.device) if contract else None
    _lib.affine_tangent(scores, ln, None, go, ge, temp, direction=direction, d_gap_open=dgo,
                        d_gap_ext=dge, d_temp=dT, map_tangent=mt, contract=c)
    return mt, c


def sw_affine_forward_impl(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
    value, mp, _ = _aff_forward(scores, gap_open, gap_ext, temp, lengths)
    return [value, mp]


# The tensor-parameter forwards also request the parameter gradients (the
# same sweeps produce them) and attach them to the returned value tensor, so
# the autograd backward reuses them instead of running the sweeps again for
# value_grad_params (orihime/ops/sw_affine.py:133-146 recomputes).  Where the
# attribute does not survive (fake or traced tensors) the backward recomputes.
_PARAM_GRADS = "_ohm_param_grads"


def sw_affine_forward_t_impl(scores, gap_open, gap_ext, temp, lengths):
    go = _scalar(gap_open, "gap_open")
    ge = _scalar(gap_ext, "gap_ext")
    T = _scalar(temp, "temperature")
    value, mp, g = _aff_forward(scores, go, ge, T, lengths, want_grads=True)
    setattr(value, _PARAM_GRADS, tuple(g))
    return [value, mp]


def sw_affine_value_grad_params_impl(scores, gap_open: float, gap_ext: float, temp: float,
                                     lengths=None):
    _, _, g = _aff_forward(scores, gap_open, gap_ext, temp, lengths, want_map=False,
                           want_grads=True)
    return g[0], g[1], g[2]


def sw_affine_value_only_impl(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
    value, _, _ = _aff_forward(scores, gap_open, gap_ext, temp, lengths, want_map=False)
    return value


def sw_affine_marginals_hvp_impl(scores, v, gap_open: float, gap_ext: float, temp: float,
                                 lengths=None):
    _check_scores(scores, short=True)
    _check_vector(scores, v, "tangent")
    mt, _ = _aff_tangent(scores, gap_open, gap_ext, temp, lengths, direction=v)
    return mt


def _aff_sens(param: int):
    def impl(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
        _check_scores(scores, short=True)
        mt, _ = _aff_tangent(scores, gap_open, gap_ext, temp, lengths,
                             dgo=float(param == 0), dge=float(param == 1), dT=float(param == 2))
        return mt

    return impl


def sw_affine_marginals_backward_pairs_impl(scores, grad_marginals, gap_open: float,
                                            gap_ext: float, temp: float, lengths=None):
    _check_scores(scores)
    _check_vector(scores, grad_marginals, "grad_marginals")
    return _aff_tangent(scores, gap_open, gap_ext, temp, lengths, direction=grad_marginals,
                        contract=True)


def sw_affine_marginals_backward_impl(scores, grad_marginals, gap_open: float, gap_ext: float,
                                      temp: float, lengths=None):
    gs, c = sw_affine_marginals_backward_pairs_impl(scores, grad_marginals, gap_open, gap_ext,
                                                    temp, lengths)
    tot = c.sum(dim=0)
    return gs, tot[0:1].clone(), tot[1:2].clone(), tot[2:3].clone()


# --------------------------------------------------------------------------
# Linear gap implementations
# --------------------------------------------------------------------------


def _lin_forward(scores, gap, temp, lengths, want_map=True, want_grads=False):
    _check_scores(scores)
    ln = _check_lengths(lengths, scores)
    B = scores.shape[0]
    value = _vec(scores, B)
    mp = torch.empty_like(scores) if want_map else None
    g = [_vec(scores, B) for _ in range(2)] if want_grads else [None] * 2
    _lib.linear_forward(scores, ln, None, gap, temp, value=value, map=mp, grad_gap=g[0],
                        grad_temp=g[1])
    return value, mp, g


def _lin_tangent(scores, gap, temp, lengths, direction=None, dg=0.0, dT=0.0, contract=False):
    ln = _check_lengths(lengths, scores)
    B = scores.shape[0]
    mt = torch.empty_like(scores)
    c = torch.empty(B, 2, dtype=torch.float32, device=scores.device) if contract else None
    _lib.linear_tangent(scores, ln, None, gap, temp, direction=direction, d_gap=dg, d_temp=dT,
                        map_tangent=mt, contract=c)
    return mt, c


def sw_forward_impl(scores, gap: float, temp: float, lengths=None):
    value, mp, _ = _lin_forward(scores, gap, temp, lengths)
    return [value, mp]


def sw_forward_t_impl(scores, gap, temp, lengths):
    value, mp, g = _lin_forward(scores, _scalar(gap, "gap"), _scalar(temp, "temperature"), lengths,
                                want_grads=True)
    setattr(value, _PARAM_GRADS, tuple(g))
    return [value, mp]


def sw_value_grad_params_impl(scores, gap: float, temp: float, lengths=None):
    _, _, g = _lin_forward(scores, gap, temp, lengths, want_map=False, want_grads=True)
    return g[0], g[1]


def sw_value_only_impl(scores, gap: float, temp: float, lengths=None):
    value, _, _ = _lin_forward(scores, gap, temp, lengths, want_map=False)
    return value


def sw_marginals_hvp_impl(scores, v, gap: float, temp: float, lengths=None):
    _check_scores(scores, short=True)
    _check_vector(scores, v, "tangent")
    mt, _ = _lin_tangent(scores, gap, temp, lengths, direction=v)
    return mt


def _lin_sens(param: int):
    def impl(scores, gap: float, temp: float, lengths=None):
        _check_scores(scores, short=True)
        mt, _ = _lin_tangent(scores, gap, temp, lengths, dg=float(param == 0), dT=float(param == 1))
        return mt

    return impl


def sw_marginals_backward_pairs_impl(scores, grad_marginals, gap: float, temp: float, lengths=None):
    _check_scores(scores)
    _check_vector(scores, grad_marginals, "grad_marginals")
    return _lin_tangent(scores, gap, temp, lengths, direction=grad_marginals, contract=True)


def sw_marginals_backward_impl(scores, grad_marginals, gap: float, temp: float, lengths=None):
    gs, c = sw_marginals_backward_pairs_impl(scores, grad_marginals, gap, temp, lengths)
    tot = c.sum(dim=0)
    return gs, tot[0:1].clone(), tot[1:2].clone()


# --------------------------------------------------------------------------
# Legacy primitives (src/sw_affine/registry.cpp:27-32, src/sw/registry.cpp:30-35)
# --------------------------------------------------------------------------


def soft_sw_affine_with_grads_impl(scores, gap_open: float, gap_ext: float, temperature: float,
                                   lengths=None):
    value, mp, g = _aff_forward(scores, gap_open, gap_ext, temperature, lengths, want_grads=True)
    return value, mp, g[0], g[1], g[2]


def soft_sw_affine_param_jacobian_impl(scores, param_type: int, gap_open: float, gap_ext: float,
                                       temperature: float, lengths=None):
    if not 0 <= param_type <= 2:
        raise RuntimeError("param_type must be 0, 1, or 2")
    return _aff_sens(int(param_type))(scores, gap_open, gap_ext, temperature, lengths)


def soft_sw_with_grads_impl(scores, gap: float, temperature: float, lengths=None):
    value, mp, g = _lin_forward(scores, gap, temperature, lengths, want_grads=True)
    return value, mp, g[0], g[1]


def soft_sw_param_jacobian_impl(scores, param_type: int, gap: float, temperature: float,
                                lengths=None):
    if not 0 <= param_type <= 1:
        raise RuntimeError("param_type must be 0 or 1")
    return _lin_sens(int(param_type))(scores, gap, temperature, lengths)


# --------------------------------------------------------------------------
# Needleman-Wunsch and Saigo-Vert, affine gap (src/nw_affine/registry.cpp:33-46,
# src/sv_affine/registry.cpp:29-44): forward + reverse sweeps for value, map
# and parameter gradients; the second-order schemas run the families'
# forward-over-reverse tangent sweeps.
# --------------------------------------------------------------------------


def _second_order(tangent, linear: bool):
    """Second-order schema implementations on a forward-over-reverse entry
    point: HVP (score direction), parameter sensitivities (unit parameter
    direction) and the map VJP (score direction with contractions).  For a
    linear-gap family the gap direction moves both affine gap scores and its
    contraction is the sum of theirs."""

    def run(scores, gap_args, temp, lengths, direction=None, dgo=0.0, dge=0.0, dT=0.0,
            contract=False):
        _check_scores(scores, short=True)
        ln = _check_lengths(lengths, scores)
        go, ge = gap_args
        mt = torch.empty_like(scores)
        c = torch.empty(scores.shape[0], 3, dtype=torch.float32,
                        device=scores.device) if contract else None
        tangent(scores, ln, None, go, ge, temp, direction=direction, d_gap_open=dgo,
                d_gap_ext=dge, d_temp=dT, map_tangent=mt, contract=c)
        return mt, c

    if linear:
        def hvp(scores, v, gap: float, temp: float, lengths=None):
            return run(scores, (gap, gap), temp, lengths, _check_vector(scores, v, "v"))[0]

        def sens(param: int):
            def impl(scores, gap: float, temp: float, lengths=None):
                d = (1.0, 1.0, 0.0) if param == 0 else (0.0, 0.0, 1.0)
                return run(scores, (gap, gap), temp, lengths, None, *d)[0]
            return impl

        def backward(scores, grad_marginals, gap: float, temp: float, lengths=None):
            mt, c = run(scores, (gap, gap), temp, lengths,
                        _check_vector(scores, grad_marginals, "grad_marginals"), contract=True)
            return mt, (c[:, 0] + c[:, 1]).sum().reshape(1), c[:, 2].sum().reshape(1)

        def jacobian(scores, param_type: int, gap: float, temperature: float, lengths=None):
            if param_type not in (0, 1):
                raise RuntimeError("param_type must be 0 (gap) or 1 (temperature)")
            return sens(param_type)(scores, gap, temperature, lengths)
    else:
        def hvp(scores, v, go: float, ge: float, temp: float, lengths=None):
            return run(scores, (go, ge), temp, lengths, _check_vector(scores, v, "v"))[0]

        def sens(param: int):
            def impl(scores, go: float, ge: float, temp: float, lengths=None):
                d = [0.0, 0.0, 0.0]
                d[param] = 1.0
                return run(scores, (go, ge), temp, lengths, None, *d)[0]
            return impl

        def backward(scores, grad_marginals, go: float, ge: float, temp: float, lengths=None):
            mt, c = run(scores, (go, ge), temp, lengths,
                        _check_vector(scores, grad_marginals, "grad_marginals"), contract=True)
            return mt, c[:, 0].sum().reshape(1), c[:, 1].sum().reshape(1), c[:, 2].sum().reshape(1)

        def jacobian(scores, param_type: int, go: float, ge: float, temperature: float,
                     lengths=None):
            if param_type not in (0, 1, 2):
                raise RuntimeError("param_type must be 0 (gap_open), 1 (gap_ext) or 2 (temperature)")
            return sens(param_type)(scores, go, ge, temperature, lengths)

    return hvp, sens, backward, jacobian


def _three_state_family(family: str, call, value_first: bool):
    """Implementations of one family's first-order schemas.  ``call`` is the
    ``_lib`` entry point; ``value_first``: ``*_value_grad_params`` returns the
    value before the three gradients (NW, src/nw_affine/torch_cpu.cpp:596-610)
    or the gradients only (SV, src/sv_affine/torch_cpu.cpp:615-624)."""

    def run(scores, go, ge, temp, lengths, want_map=True, want_grads=False):
        _check_scores(scores)
        ln = _check_lengths(lengths, scores)
        B = scores.shape[0]
        value = _vec(scores, B)
        # the gradients come out of the reverse sweep that writes the map
        mp = torch.empty_like(scores) if (want_map or want_grads) else None
        g = [_vec(scores, B) for _ in range(3)] if want_grads else [None] * 3
        call(scores, ln, None, go, ge, temp, value=value, map=mp, grad_open=g[0],
             grad_ext=g[1], grad_temp=g[2])
        return value, mp, g

    def forward(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
        value, mp, _ = run(scores, gap_open, gap_ext, temp, lengths)
        return [value, mp]

    def forward_t(scores, gap_open, gap_ext, temp, lengths):
        go = _scalar(gap_open, "gap_open")
        ge = _scalar(gap_ext, "gap_ext")
        T = _scalar(temp, "temperature")
        value, mp, g = run(scores, go, ge, T, lengths, want_grads=True)
        setattr(value, _PARAM_GRADS, tuple(g))
        return [value, mp]

    def value_grad_params(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
        value, _, g = run(scores, gap_open, gap_ext, temp, lengths, want_map=False,
                          want_grads=True)
        return (value, g[0], g[1], g[2]) if value_first else (g[0], g[1], g[2])

    def with_grads(scores, gap_open: float, gap_ext: float, temperature: float, lengths=None):
        value, mp, g = run(scores, gap_open, gap_ext, temperature, lengths, want_grads=True)
        return value, mp, g[0], g[1], g[2]

    def value_only(scores, gap_open: float, gap_ext: float, temp: float, lengths=None):
        value, _, _ = run(scores, gap_open, gap_ext, temp, lengths, want_map=False)
        return value

    return forward, forward_t, value_grad_params, with_grads, value_only


def _three_state_schemas(family: str, call, value_first: bool, tangent) -> dict:
    fwd, fwd_t, vgp, wg, vo = _three_state_family(family, call, value_first)
    hvp, sens, mb, jac = _second_order(tangent, linear=False)
    so = {"marginals_backward": mb, "marginals_hvp": hvp, "marginals_grad_gap_open": sens(0),
          "marginals_grad_gap_ext": sens(1), "marginals_grad_temp": sens(2), "hvp": hvp,
          "param_jacobian": jac, "backward_full": mb}.__getitem__
    f = family
    S = "(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths)"
    L = "(Tensor scores, float gap_open, float gap_ext, float temperature, Tensor? lengths)"
    grads = "(Tensor, Tensor, Tensor, Tensor)" if value_first else "(Tensor, Tensor, Tensor)"
    return {
        f"{f}_affine_forward": (f"{S} -> Tensor[]", fwd),
        f"{f}_affine_forward_t": ("(Tensor scores, Tensor gap_open, Tensor gap_ext, Tensor temp, Tensor lengths) -> Tensor[]", fwd_t),
        f"{f}_affine_value_grad_params": (f"{S} -> {grads}", vgp),
        f"{f}_affine_marginals_backward": ("(Tensor scores, Tensor grad_marginals, float gap_open, float gap_ext, float temp, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor)", so("marginals_backward")),
        f"{f}_affine_marginals_hvp": ("(Tensor scores, Tensor v, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", so("marginals_hvp")),
        f"{f}_affine_marginals_grad_gap_open": (f"{S} -> Tensor", so("marginals_grad_gap_open")),
        f"{f}_affine_marginals_grad_gap_ext": (f"{S} -> Tensor", so("marginals_grad_gap_ext")),
        f"{f}_affine_marginals_grad_temp": (f"{S} -> Tensor", so("marginals_grad_temp")),
        f"{f}_affine_value_only": (f"{S} -> Tensor", vo),
        f"soft_{f}_affine": ("(Tensor scores, Tensor gap_open, Tensor gap_ext, Tensor temperature, Tensor lengths) -> Tensor[]", fwd_t),
        f"soft_{f}_affine_float": (f"{L} -> Tensor[]", fwd),
        f"soft_{f}_affine_with_grads": (f"{L} -> (Tensor, Tensor, Tensor, Tensor, Tensor)", wg),
        f"soft_{f}_affine_hvp": ("(Tensor scores, Tensor tangent, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> Tensor", so("hvp")),
        f"soft_{f}_affine_param_jacobian": ("(Tensor scores, int param_type, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> Tensor", so("param_jacobian")),
        f"soft_{f}_affine_backward_full": ("(Tensor scores, Tensor grad_alignment, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor)", so("backward_full")),
    }


def _linear_family_schemas(ns: str, soft: str, call, model: str, tangent) -> dict:
    """First-order schemas of a linear-gap family (src/nw/registry.cpp:35-49,
    src/sv_linear/registry.cpp): the affine sweeps with one gap score in both
    roles, dS/dgap = dS/dgap_open + dS/dgap_ext (summed in the kernel)."""

    def run(scores, gap, temp, lengths, want_map=True, want_grads=False):
        _check_scores(scores)
        ln = _check_lengths(lengths, scores)
        B = scores.shape[0]
        value = _vec(scores, B)
        mp = torch.empty_like(scores) if (want_map or want_grads) else None
        g = [_vec(scores, B) for _ in range(2)] if want_grads else [None] * 2
        call(scores, ln, None, gap, temp, value=value, map=mp, grad_gap=g[0], grad_temp=g[1])
        return value, mp, g

    def forward(scores, gap: float, temp: float, lengths=None):
        value, mp, _ = run(scores, gap, temp, lengths)
        return [value, mp]

    def forward_t(scores, gap, temp, lengths):
        value, mp, g = run(scores, _scalar(gap, "gap"), _scalar(temp, "temperature"), lengths,
                           want_grads=True)
        setattr(value, _PARAM_GRADS, tuple(g))
        return [value, mp]

    def value_grad_params(scores, gap: float, temp: float, lengths=None):
        _, _, g = run(scores, gap, temp, lengths, want_map=False, want_grads=True)
        return g[0], g[1]

    def with_grads(scores, gap: float, temperature: float, lengths=None):
        value, mp, g = run(scores, gap, temperature, lengths, want_grads=True)
        return value, mp, g[0], g[1]

    def value_only(scores, gap: float, temp: float, lengths=None):
        value, _, _ = run(scores, gap, temp, lengths, want_map=False)
        return value

    hvp, sens, mb, jac = _second_order(tangent, linear=True)
    second_order = {"marginals_backward": mb, "marginals_hvp": hvp,
                    "marginals_grad_gap": sens(0), "marginals_grad_temp": sens(1),
                    "hvp": hvp, "param_jacobian": jac, "backward_full": mb}.__getitem__

    S = "(Tensor scores, float gap, float temp, Tensor? lengths)"
    L = "(Tensor scores, float gap, float temperature, Tensor? lengths)"
    return {
        f"{ns}_forward": (f"{S} -> Tensor[]", forward),
        f"{ns}_forward_t": ("(Tensor scores, Tensor gap, Tensor temp, Tensor lengths) -> Tensor[]", forward_t),
        f"{ns}_value_grad_params": (f"{S} -> (Tensor, Tensor)", value_grad_params),
        f"{ns}_marginals_backward": ("(Tensor scores, Tensor grad_marginals, float gap, float temp, Tensor? lengths) -> (Tensor, Tensor, Tensor)", second_order("marginals_backward")),
        f"{ns}_marginals_hvp": ("(Tensor scores, Tensor v, float gap, float temp, Tensor? lengths) -> Tensor", second_order("marginals_hvp")),
        f"{ns}_marginals_grad_gap": (f"{S} -> Tensor", second_order("marginals_grad_gap")),
        f"{ns}_marginals_grad_temp": (f"{S} -> Tensor", second_order("marginals_grad_temp")),
        f"{ns}_value_only": (f"{S} -> Tensor", value_only),
        soft: ("(Tensor scores, Tensor gap, Tensor temperature, Tensor lengths) -> Tensor[]", forward_t),
        f"{soft}_float": (f"{L} -> Tensor[]", forward),
        f"{soft}_with_grads": (f"{L} -> (Tensor, Tensor, Tensor, Tensor)", with_grads),
        f"{soft}_hvp": ("(Tensor scores, Tensor tangent, float gap, float temperature, Tensor? lengths) -> Tensor", second_order("hvp")),
        f"{soft}_param_jacobian": ("(Tensor scores, int param_type, float gap, float temperature, Tensor? lengths) -> Tensor", second_order("param_jacobian")),
        f"{soft}_backward_full": ("(Tensor scores, Tensor grad_alignment, float gap, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor)", second_order("backward_full")),
    }


# --------------------------------------------------------------------------
# Registration
# --------------------------------------------------------------------------

_SCHEMAS = {
    # affine — namespaced API
    "sw_affine_forward": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor[]", sw_affine_forward_impl),
    "sw_affine_forward_t": ("(Tensor scores, Tensor gap_open, Tensor gap_ext, Tensor temp, Tensor lengths) -> Tensor[]", sw_affine_forward_t_impl),
    "sw_affine_value_grad_params": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> (Tensor, Tensor, Tensor)", sw_affine_value_grad_params_impl),
    "sw_affine_marginals_backward": ("(Tensor scores, Tensor grad_marginals, float gap_open, float gap_ext, float temp, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor)", sw_affine_marginals_backward_impl),
    "sw_affine_marginals_hvp": ("(Tensor scores, Tensor v, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", sw_affine_marginals_hvp_impl),
    "sw_affine_marginals_grad_gap_open": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", _aff_sens(0)),
    "sw_affine_marginals_grad_gap_ext": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", _aff_sens(1)),
    "sw_affine_marginals_grad_temp": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", _aff_sens(2)),
    "sw_affine_value_only": ("(Tensor scores, float gap_open, float gap_ext, float temp, Tensor? lengths) -> Tensor", sw_affine_value_only_impl),
    "sw_affine_marginals_backward_pairs": ("(Tensor scores, Tensor grad_marginals, float gap_open, float gap_ext, float temp, Tensor? lengths) -> (Tensor, Tensor)", sw_affine_marginals_backward_pairs_impl),
    # affine — legacy
    "soft_sw_affine": ("(Tensor scores, Tensor gap_open, Tensor gap_ext, Tensor temperature, Tensor lengths) -> Tensor[]", sw_affine_forward_t_impl),
    "soft_sw_affine_float": ("(Tensor scores, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> Tensor[]", sw_affine_forward_impl),
    "soft_sw_affine_with_grads": ("(Tensor scores, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor, Tensor)", soft_sw_affine_with_grads_impl),
    "soft_sw_affine_hvp": ("(Tensor scores, Tensor tangent, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> Tensor", sw_affine_marginals_hvp_impl),
    "soft_sw_affine_param_jacobian": ("(Tensor scores, int param_type, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> Tensor", soft_sw_affine_param_jacobian_impl),
    "soft_sw_affine_backward_full": ("(Tensor scores, Tensor grad_alignment, float gap_open, float gap_ext, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor)", sw_affine_marginals_backward_impl),
    # linear — namespaced API
    "sw_forward": ("(Tensor scores, float gap, float temp, Tensor? lengths) -> Tensor[]", sw_forward_impl),
    "sw_forward_t": ("(Tensor scores, Tensor gap, Tensor temp, Tensor lengths) -> Tensor[]", sw_forward_t_impl),
    "sw_value_grad_params": ("(Tensor scores, float gap, float temp, Tensor? lengths) -> (Tensor, Tensor)", sw_value_grad_params_impl),
    "sw_marginals_backward": ("(Tensor scores, Tensor grad_marginals, float gap, float temp, Tensor? lengths) -> (Tensor, Tensor, Tensor)", sw_marginals_backward_impl),
    "sw_marginals_hvp": ("(Tensor scores, Tensor v, float gap, float temp, Tensor? lengths) -> Tensor", sw_marginals_hvp_impl),
    "sw_marginals_grad_gap": ("(Tensor scores, float gap, float temp, Tensor? lengths) -> Tensor", _lin_sens(0)),
    "sw_marginals_grad_temp": ("(Tensor scores, float gap, float temp, Tensor? lengths) -> Tensor", _lin_sens(1)),
    "sw_value_only": ("(Tensor scores, float gap, float temp, Tensor? lengths) -> Tensor", sw_value_only_impl),
    "sw_marginals_backward_pairs": ("(Tensor scores, Tensor grad_marginals, float gap, float temp, Tensor? lengths) -> (Tensor, Tensor)", sw_marginals_backward_pairs_impl),
    # linear — legacy
    "soft_sw": ("(Tensor scores, Tensor gap, Tensor temperature, Tensor lengths) -> Tensor[]", sw_forward_t_impl),
    "soft_sw_float": ("(Tensor scores, float gap, float temperature, Tensor? lengths) -> Tensor[]", sw_forward_impl),
    "soft_sw_with_grads": ("(Tensor scores, float gap, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor, Tensor)", soft_sw_with_grads_impl),
    "soft_sw_hvp": ("(Tensor scores, Tensor tangent, float gap, float temperature, Tensor? lengths) -> Tensor", sw_marginals_hvp_impl),
    "soft_sw_param_jacobian": ("(Tensor scores, int param_type, float gap, float temperature, Tensor? lengths) -> Tensor", soft_sw_param_jacobian_impl),
    "soft_sw_backward_full": ("(Tensor scores, Tensor grad_alignment, float gap, float temperature, Tensor? lengths) -> (Tensor, Tensor, Tensor)", sw_marginals_backward_impl),
}


_SCHEMAS.update(_three_state_schemas("nw", _lib.nw_affine_forward, True, _lib.nw_affine_tangent))
_SCHEMAS.update(_three_state_schemas("sv", _lib.sv_affine_forward, False, _lib.sv_affine_tangent))
_SCHEMAS.update(_linear_family_schemas("nw", "soft_nw", _lib.nw_forward, "nw",
                                       _lib.nw_affine_tangent))
_SCHEMAS.update(_linear_family_schemas("sv_linear", "soft_sv_linear", _lib.sv_forward, "sv",
                                       _lib.sv_affine_tangent))


def _fake_for(name: str, schema: str):
    """Shape-only kernel: outputs follow the schema's return types."""
    ret = schema.split("->")[1].strip()

    def like(scores):
        return torch.empty_like(scores)

    def vec(scores):
        return scores.new_empty((scores.shape[0],))

    def one(scores):
        return scores.new_empty((1,))

    def fake(scores, *args):
        if ret == "Tensor[]":
            return [vec(scores), like(scores)]
        if name.endswith("value_grad_params"):
            n = 4 if name.startswith("nw_affine") else (3 if "affine" in name else 2)
            return tuple(vec(scores) for _ in range(n))
        if name.endswith("_with_grads"):
            n = 3 if "affine" in name else 2
            return (vec(scores), like(scores), *(vec(scores) for _ in range(n)))
        if name.endswith("marginals_backward_pairs"):
            n = 3 if "affine" in name else 2
            return like(scores), scores.new_empty((scores.shape[0], n))
        if name.endswith(("marginals_backward", "backward_full")):
            n = 3 if "affine" in name else 2
            return (like(scores), *(one(scores) for _ in range(n)))
        if name.endswith("value_only"):
            return vec(scores)
        return like(scores)

    return fake


for _name, (_schema, _impl) in _SCHEMAS.items():
    _LIB.define(_name + _schema)
    _LIB.impl(_name, _guarded(_name, _impl), "CPU")
    _LIB.impl(_name, _guarded(_name, _impl), "CUDA")
    torch.library.register_fake(f"orihime::{_name}", _fake_for(_name, _schema), lib=_LIB)

orihime = torch.ops.orihime

# --------------------------------------------------------------------------
# Autograd of the forward primitives: d(value) = map, d(map) = the map VJP,
# and for tensor parameters the contracted parameter derivatives
# (src/sw_affine/torch_hip.cpp:210-400).
# --------------------------------------------------------------------------


def _forward_setup(ctx, inputs, output):
    # an unused output's gradient arrives as None (not zeros), so a backward
    # through the value alone runs no map VJP
    ctx.set_materialize_grads(False)
    scores = inputs[0]
    ctx.save_for_backward(scores, output[1])
    ctx.inputs = inputs
    ctx.param_grads = getattr(output[0], _PARAM_GRADS, None)


def _make_forward_backward(affine: bool, tensor_params: bool, family: str | None = None):
    nparam = 3 if affine else 2
    family = family or ("sw_affine" if affine else "sw")

    def backward(ctx, grads):
        scores, mp = ctx.saved_tensors
        inputs = ctx.inputs
        params = inputs[1:1 + nparam]
        lengths = inputs[1 + nparam]
        fparams = tuple(_scalar(p, "param") if isinstance(p, Tensor) else float(p) for p in params)
        gv = grads[0] if grads is not None else None
        gm = grads[1] if grads is not None and len(grads) > 1 else None
        g_scores = torch.zeros_like(scores)
        g_params = [torch.zeros(1, dtype=scores.dtype, device=scores.device) for _ in range(nparam)]
        if gv is not None:
            g_scores = g_scores + gv.reshape(-1, 1, 1) * mp
            if tensor_params:
                pg = ctx.param_grads
                if pg is None:
                    pg = getattr(orihime, f"{family}_value_grad_params")(scores.detach(), *fparams, lengths)
                    if family == "nw_affine":
                        pg = pg[1:]  # (value, grads...)
                for k in range(nparam):
                    g_params[k] = g_params[k] + (gv * pg[k]).sum().reshape(1)
        if gm is not None:
            mb = getattr(orihime, f"{family}_marginals_backward")(
                scores.detach(), gm.contiguous().to(torch.float32), *fparams, lengths)
            g_scores = g_scores + mb[0]
            for k in range(nparam):
                g_params[k] = g_params[k] + mb[1 + k]
        if tensor_params:
            outs = [g.reshape(p.shape).to(p.dtype) for g, p in zip(g_params, params)]
        else:
            outs = [None] * nparam
        return (g_scores, *outs, None)

    return backward


for _name, _aff, _tp in (
    ("sw_affine_forward", True, False), ("sw_affine_forward_t", True, True),
    ("soft_sw_affine_float", True, False), ("soft_sw_affine", True, True),
    ("sw_forward", False, False), ("sw_forward_t", False, True),
    ("soft_sw_float", False, False), ("soft_sw", False, True),
):
    torch.library.register_autograd(
        f"orihime::{_name}", _make_forward_backward(_aff, _tp), setup_context=_forward_setup,
        lib=_LIB,
    )
for _ns, _soft in (("nw", "soft_nw"), ("sv_linear", "soft_sv_linear")):
    for _name, _tp in ((f"{_ns}_forward", False), (f"{_ns}_forward_t", True),
                       (f"{_soft}_float", False), (_soft, True)):
        torch.library.register_autograd(
            f"orihime::{_name}", _make_forward_backward(False, _tp, _ns),
            setup_context=_forward_setup, lib=_LIB,
        )
for _fam in ("nw", "sv"):
    for _name, _tp in ((f"{_fam}_affine_forward", False), (f"{_fam}_affine_forward_t", True),
                       (f"soft_{_fam}_affine_float", False), (f"soft_{_fam}_affine", True)):
        torch.library.register_autograd(
            f"orihime::{_name}", _make_forward_backward(True, _tp, f"{_fam}_affine"),
            setup_context=_forward_setup, lib=_LIB,
        )

# --------------------------------------------------------------------------
# vmap: fold the mapped dimension into the native batch (orihime/_ops.py:237-406)
# --------------------------------------------------------------------------

_VMAP_BATCH_ARGUMENTS = frozenset({"scores", "grad_marginals", "v", "lengths"})


def _make_batch_folding_vmap_rule(op_name: str, pairs_op: str | None = None):
    raw = getattr(orihime, op_name).default
    names = tuple(a.name for a in raw._schema.arguments)
    batch_idx = tuple(i for i, n in enumerate(names) if n in _VMAP_BATCH_ARGUMENTS)
    pairs_raw = getattr(orihime, pairs_op).default if pairs_op else None

    def rule(info, in_dims, *args):
        for i, d in enumerate(in_dims):
            if d is not None and i not in batch_idx:
                raise RuntimeError(
                    f"orihime::{op_name} does not support mapping scalar argument "
                    f"{names[i]!r}; keep scalar parameters unbatched"
                )
        native = None
        folded = list(args)
        for i in batch_idx:
            a = args[i]
            if not isinstance(a, Tensor):
                continue
            d = in_dims[i]
            x = a.unsqueeze(0).expand(info.batch_size, *a.shape) if d is None else a.movedim(d, 0)
            if x.ndim < 2:
                raise RuntimeError(f"orihime::{op_name} batch argument {names[i]!r} has no native batch dimension")
            if native is None:
                native = x.shape[1]
            elif x.shape[1] != native:
                raise RuntimeError(
                    f"orihime::{op_name} received inconsistent native batch sizes {native} and {x.shape[1]}"
                )
            folded[i] = x.reshape(info.batch_size * native, *x.shape[2:]).contiguous()
        if native is None:
            raise RuntimeError(f"orihime::{op_name} requires a mapped batch-aligned tensor")

        def restore(t):
            return t.reshape(info.batch_size, native, *t.shape[1:])

        if pairs_raw is not None:
            gs, contract = pairs_raw(*folded)
            sums = restore(contract).sum(dim=1)  # [V, nparams]
            outs = (restore(gs), *(sums[:, k:k + 1] for k in range(sums.shape[1])))
            return outs, tuple(0 for _ in outs)
        result = raw(*folded)
        if isinstance(result, (list, tuple)):
            outs = [restore(t) for t in result]
            dims = [0] * len(outs)
            return (outs, dims) if isinstance(result, list) else (tuple(outs), tuple(dims))
        return restore(result), 0

    return rule


for _family, _fields in (("sw", ("gap", "temp")), ("sw_affine", ("gap_open", "gap_ext", "temp"))):
    for _op in (f"{_family}_forward", f"{_family}_forward_t", f"{_family}_value_grad_params",
                f"{_family}_marginals_hvp", f"{_family}_value_only",
                *(f"{_family}_marginals_grad_{f}" for f in _fields)):
        torch.library.register_vmap(f"orihime::{_op}", _make_batch_folding_vmap_rule(_op), lib=_LIB)
    torch.library.register_vmap(
        f"orihime::{_family}_marginals_backward",
        _make_batch_folding_vmap_rule(f"{_family}_marginals_backward", f"{_family}_marginals_backward_pairs"),
        lib=_LIB,
    )

# --------------------------------------------------------------------------
# Wrappers used by the kernel adapters (orihime/_ops.py:434-449)
# --------------------------------------------------------------------------

sw_forward = orihime.sw_forward
sw_forward_t = orihime.sw_forward_t
sw_value_grad_params = orihime.sw_value_grad_params
sw_marginals_backward = orihime.sw_marginals_backward
sw_marginals_hvp = orihime.sw_marginals_hvp
sw_marginals_grad_gap = orihime.sw_marginals_grad_gap
sw_marginals_grad_temp = orihime.sw_marginals_grad_temp
sw_value_only = orihime.sw_value_only

sw_affine_forward = orihime.sw_affine_forward
sw_affine_forward_t = orihime.sw_affine_forward_t
sw_affine_value_grad_params = orihime.sw_affine_value_grad_params
sw_affine_marginals_backward = orihime.sw_affine_marginals_backward
sw_affine_marginals_hvp = orihime.sw_affine_marginals_hvp
sw_affine_marginals_grad_gap_open = orihime.sw_affine_marginals_grad_gap_open
sw_affine_marginals_grad_gap_ext = orihime.sw_affine_marginals_grad_gap_ext
sw_affine_marginals_grad_temp = orihime.sw_affine_marginals_grad_temp
sw_affine_value_only = orihime.sw_affine_value_only

for _fam, _fields in (("nw_affine", ("gap_open", "gap_ext", "temp")),
                      ("sv_affine", ("gap_open", "gap_ext", "temp")),
                      ("nw", ("gap", "temp")), ("sv_linear", ("gap", "temp"))):
    for _op in (f"{_fam}_forward", f"{_fam}_forward_t", f"{_fam}_value_grad_params",
                f"{_fam}_value_only", f"{_fam}_marginals_hvp",
                *(f"{_fam}_marginals_grad_{f}" for f in _fields)):
        torch.library.register_vmap(f"orihime::{_op}", _make_batch_folding_vmap_rule(_op), lib=_LIB)

nw_affine_forward = orihime.nw_affine_forward
nw_affine_forward_t = orihime.nw_affine_forward_t
nw_affine_value_grad_params = orihime.nw_affine_value_grad_params
nw_affine_marginals_backward = orihime.nw_affine_marginals_backward
nw_affine_marginals_hvp = orihime.nw_affine_marginals_hvp
nw_affine_marginals_grad_gap_open = orihime.nw_affine_marginals_grad_gap_open
nw_affine_marginals_grad_gap_ext = orihime.nw_affine_marginals_grad_gap_ext
nw_affine_marginals_grad_temp = orihime.nw_affine_marginals_grad_temp
nw_affine_value_only = orihime.nw_affine_value_only

sv_affine_forward = orihime.sv_affine_forward
sv_affine_forward_t = orihime.sv_affine_forward_t
sv_affine_value_grad_params = orihime.sv_affine_value_grad_params
sv_affine_marginals_backward = orihime.sv_affine_marginals_backward
sv_affine_marginals_hvp = orihime.sv_affine_marginals_hvp
sv_affine_marginals_grad_gap_open = orihime.sv_affine_marginals_grad_gap_open
sv_affine_marginals_grad_gap_ext = orihime.sv_affine_marginals_grad_gap_ext
sv_affine_marginals_grad_temp = orihime.sv_affine_marginals_grad_temp
sv_affine_value_only = orihime.sv_affine_value_only

nw_forward = orihime.nw_forward
nw_forward_t = orihime.nw_forward_t
nw_value_grad_params = orihime.nw_value_grad_params
nw_marginals_backward = orihime.nw_marginals_backward
nw_marginals_hvp = orihime.nw_marginals_hvp
nw_marginals_grad_gap = orihime.nw_marginals_grad_gap
nw_marginals_grad_temp = orihime.nw_marginals_grad_temp
nw_value_only = orihime.nw_value_only

sv_linear_forward = orihime.sv_linear_forward
sv_linear_forward_t = orihime.sv_linear_forward_t
sv_linear_value_grad_params = orihime.sv_linear_value_grad_params
sv_linear_marginals_backward = orihime.sv_linear_marginals_backward
sv_linear_marginals_hvp = orihime.sv_linear_marginals_hvp
sv_linear_marginals_grad_gap = orihime.sv_linear_marginals_grad_gap
sv_linear_marginals_grad_temp = orihime.sv_linear_marginals_grad_temp
sv_linear_value_only = orihime.sv_linear_value_only


def _has_cuda_dispatch() -> bool:
    try:
        return torch._C._dispatch_has_kernel_for_dispatch_key("orihime::sw_forward", "CUDA")
    except (AttributeError, RuntimeError):
        return False
