"""Packed ragged batches: many pairs of different shapes in one launch.

A packed batch stores pair b's L1_b x L2_b scores row-major at element
``offsets[b]`` of one flat FP32 buffer (row stride L2_b).  One persistent
launch processes every pair with no padding and no mask, which the padded
``[B, L1max, L2max]`` + ``mask=`` form cannot do (there the gap states still
run through the padding).  Maps and map-shaped derivatives use the same
packed layout.

    flat, lengths, offsets = pack([s0, s1, ...])        # s_b: [L1_b, L2_b]
    value = sw_affine_value(flat, lengths, offsets, gap_open_score=-8.0, ...)
    value.sum().backward()                              # flat.grad: packed map

Semantics per pair are those of ``orihime_amd.sw_affine`` on that pair alone.

Every call checks, on the device, that each pair's block lies inside ``flat``
(``0 <= offsets[b]`` and ``offsets[b] + L1_b * L2_b <= flat.numel()``) and
that the lengths lie in [0, max]; that costs one read-back.  While a HIP graph
is being captured nothing is read back: pass ``L1max``/``L2max`` and Python
numbers for the parameters, and the checks are left to the caller (the
kernels still clamp lengths into [0, max]).
"""

from __future__ import annotations

from collections.abc import Sequence

import torch
from torch import Tensor

from . import _checks, _lib
from .operator import _validate_temperature


def pack(pairs: Sequence[Tensor]) -> tuple[Tensor, Tensor, Tensor]:
    """Concatenate [L1_b, L2_b] score matrices into (flat, lengths, offsets)."""
    if not pairs:
        raise ValueError("pack needs at least one pair")
    dev = pairs[0].device
    for p in pairs:
        if p.ndim != 2:
            raise ValueError("every pair must be a [L1, L2] matrix")
        if p.device != dev:
            raise ValueError("every pair must be on the same device")
    sizes = [p.numel() for p in pairs]
    offs = [0]
    for n in sizes[:-1]:
        offs.append(offs[-1] + n)
    flat = torch.cat([p.reshape(-1).to(torch.float32) for p in pairs]) if sum(sizes) else \
        torch.zeros(0, dtype=torch.float32, device=dev)
    lengths = torch.tensor([[p.shape[0], p.shape[1]] for p in pairs], dtype=torch.int32,
                           device=dev)
    offsets = torch.tensor(offs, dtype=torch.int64, device=dev)
    return flat.contiguous(), lengths, offsets


def unpack(flat: Tensor, lengths: Tensor, offsets: Tensor) -> list[Tensor]:
    """Split a packed map back into [L1_b, L2_b] matrices."""
    ln = lengths.cpu().tolist()
    of = offsets.cpu().tolist()
    return [flat[o:o + a * b].reshape(a, b) for (a, b), o in zip(ln, of)]


def _check(flat: Tensor, lengths: Tensor, offsets: Tensor, L1max, L2max):
    if flat.ndim != 1 or flat.dtype != torch.float32 or not flat.is_contiguous():
        raise ValueError("packed scores must be a contiguous 1-D float32 tensor")
    if lengths.dtype != torch.int32 or lengths.ndim != 2 or lengths.shape[1] != 2:
        raise ValueError("lengths must be int32 [B, 2]")
    if offsets.dtype != torch.int64 or offsets.ndim != 1 or offsets.shape[0] != lengths.shape[0]:
        raise ValueError("offsets must be int64 [B]")
    if lengths.device != flat.device or offsets.device != flat.device:
        raise ValueError("scores, lengths and offsets must be on one device")
    lengths = lengths.contiguous()
    offsets = offsets.contiguous()
    if L1max is None or L2max is None:
        if _checks.deferred(lengths):
            raise ValueError("pass L1max and L2max while capturing a HIP graph (otherwise they "
                             "are read from lengths on the host)")
        mx = lengths.amax(dim=0).cpu() if lengths.shape[0] else torch.zeros(2, dtype=torch.int32)
        L1max = int(mx[0]) if L1max is None else int(L1max)
        L2max = int(mx[1]) if L2max is None else int(L2max)
    L1max, L2max = int(L1max), int(L2max)
    if L1max < 0 or L2max < 0:
        raise ValueError("L1max and L2max must be non-negative")
    if lengths.shape[0] and not _checks.deferred(lengths):
        l1, l2 = lengths[:, 0], lengths[:, 1]
        _checks.require((l1 >= 0) & (l1 <= L1max) & (l2 >= 0) & (l2 <= L2max),
                        f"lengths must lie in [0, L1max={L1max}] x [0, L2max={L2max}]")
        end = offsets + l1.long() * l2.long()
        _checks.require((offsets >= 0) & (end <= flat.numel()),
                        "every pair's block offsets[b] + L1_b * L2_b must lie within the "
                        f"packed scores ({flat.numel()} elements)")
    return lengths, offsets, L1max, L2max


def _params(*vals):
    return tuple(_checks.host_scalar(v, "scoring parameter") if isinstance(v, Tensor) else float(v)
                 for v in vals)


def sw_affine_forward(flat: Tensor, lengths: Tensor, offsets: Tensor, *, gap_open_score=0.0,
                      gap_extend_score=0.0, temperature=1.0, need_map: bool = True,
                      need_grads: bool = True, L1max=None, L2max=None, workspace=None):
    """(value[B], packed map or None, grad_open, grad_ext, grad_temp or None)."""
    _validate_temperature(temperature)
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    go, ge, T = _params(gap_open_score, gap_extend_score, temperature)
    B = lengths.shape[0]
    value = torch.empty(B, dtype=torch.float32, device=flat.device)
    mp = torch.empty_like(flat) if need_map else None
    g = [torch.empty(B, dtype=torch.float32, device=flat.device) for _ in range(3)] \
        if need_grads else [None] * 3
    _lib.affine_forward(flat, lengths, offsets, go, ge, T, value=value, map=mp, grad_open=g[0],
                        grad_ext=g[1], grad_temp=g[2], L1max=L1, L2max=L2, workspace=workspace)
    return value, mp, g[0], g[1], g[2]


def sw_affine_tangent(flat: Tensor, lengths: Tensor, offsets: Tensor, *, gap_open_score=0.0,
                      gap_extend_score=0.0, temperature=1.0, direction: Tensor | None = None,
                      d_gap_open: float = 0.0, d_gap_extend: float = 0.0, d_temperature: float = 0.0,
                      contract: bool = False, L1max=None, L2max=None, workspace=None):
    """(packed map tangent, per-pair contractions [B, 3] or None)."""
    _validate_temperature(temperature)
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    go, ge, T = _params(gap_open_score, gap_extend_score, temperature)
    mt = torch.empty_like(flat)
    c = torch.empty(lengths.shape[0], 3, dtype=torch.float32, device=flat.device) if contract \
        else None
    _lib.affine_tangent(flat, lengths, offsets, go, ge, T, direction=direction,
                        d_gap_open=d_gap_open, d_gap_ext=d_gap_extend, d_temp=d_temperature,
                        map_tangent=mt, contract=c, L1max=L1, L2max=L2, workspace=workspace)
    return mt, c


def sw_forward(flat: Tensor, lengths: Tensor, offsets: Tensor, *, gap_score=0.0, temperature=1.0,
               need_map: bool = True, need_grads: bool = True, L1max=None, L2max=None,
               workspace=None):
    """(value[B], packed map or None, grad_gap, grad_temp or None)."""
    _validate_temperature(temperature)
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    g_, T = _params(gap_score, temperature)
    B = lengths.shape[0]
    value = torch.empty(B, dtype=torch.float32, device=flat.device)
    mp = torch.empty_like(flat) if need_map else None
    g = [torch.empty(B, dtype=torch.float32, device=flat.device) for _ in range(2)] \
        if need_grads else [None] * 2
    _lib.linear_forward(flat, lengths, offsets, g_, T, value=value, map=mp, grad_gap=g[0],
                        grad_temp=g[1], L1max=L1, L2max=L2, workspace=workspace)
    return value, mp, g[0], g[1]


def sw_tangent(flat: Tensor, lengths: Tensor, offsets: Tensor, *, gap_score=0.0, temperature=1.0,
               direction: Tensor | None = None, d_gap: float = 0.0, d_temperature: float = 0.0,
               contract: bool = False, L1max=None, L2max=None, workspace=None):
    """(packed map tangent, per-pair contractions [B, 2] or None)."""
    _validate_temperature(temperature)
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    g_, T = _params(gap_score, temperature)
    mt = torch.empty_like(flat)
    c = torch.empty(lengths.shape[0], 2, dtype=torch.float32, device=flat.device) if contract \
        else None
    _lib.linear_tangent(flat, lengths, offsets, g_, T, direction=direction, d_gap=d_gap,
                        d_temp=d_temperature, map_tangent=mt, contract=c, L1max=L1, L2max=L2,
                        workspace=workspace)
    return mt, c


class _PackedValue(torch.autograd.Function):
    @staticmethod
    def forward(ctx, flat, lengths, offsets, affine, params, L1max, L2max):
        if affine:
            value, mp, *_ = sw_affine_forward(flat, lengths, offsets, gap_open_score=params[0],
                                              gap_extend_score=params[1], temperature=params[2],
                                              need_grads=False, L1max=L1max, L2max=L2max)
        else:
            value, mp, *_ = sw_forward(flat, lengths, offsets, gap_score=params[0],
                                       temperature=params[1], need_grads=False, L1max=L1max,
                                       L2max=L2max)
        ctx.save_for_backward(mp, lengths, offsets)
        return value

    @staticmethod
    def backward(ctx, grad_value):
        mp, lengths, offsets = ctx.saved_tensors
        per_pair = torch.repeat_interleave(grad_value, (lengths[:, 0] * lengths[:, 1]).long(),
                                           output_size=mp.numel())
        return mp * per_pair, None, None, None, None, None, None


class _PackedMap(torch.autograd.Function):
    @staticmethod
    def forward(ctx, flat, lengths, offsets, affine, params, L1max, L2max):
        if affine:
            _, mp, *_ = sw_affine_forward(flat, lengths, offsets, gap_open_score=params[0],
                                          gap_extend_score=params[1], temperature=params[2],
                                          need_grads=False, L1max=L1max, L2max=L2max)
        else:
            _, mp, *_ = sw_forward(flat, lengths, offsets, gap_score=params[0],
                                   temperature=params[1], need_grads=False, L1max=L1max,
                                   L2max=L2max)
        ctx.save_for_backward(flat, lengths, offsets)
        ctx.affine, ctx.params, ctx.L = affine, params, (L1max, L2max)
        return mp

    @staticmethod
    def backward(ctx, grad_map):
        flat, lengths, offsets = ctx.saved_tensors
        p = ctx.params
        g = grad_map.contiguous().to(torch.float32)
        if ctx.affine:
            mt, _ = sw_affine_tangent(flat, lengths, offsets, gap_open_score=p[0],
                                      gap_extend_score=p[1], temperature=p[2], direction=g,
                                      L1max=ctx.L[0], L2max=ctx.L[1])
        else:
            mt, _ = sw_tangent(flat, lengths, offsets, gap_score=p[0], temperature=p[1],
                               direction=g, L1max=ctx.L[0], L2max=ctx.L[1])
        return mt, None, None, None, None, None, None


def sw_affine_value(flat, lengths, offsets, *, gap_open_score=0.0, gap_extend_score=0.0,
                    temperature=1.0, L1max=None, L2max=None) -> Tensor:
    """Per-pair affine soft-SW value [B] of a packed batch (differentiable in the scores)."""
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    return _PackedValue.apply(flat, lengths, offsets, True,
                              _params(gap_open_score, gap_extend_score, temperature), L1, L2)


def sw_affine(flat, lengths, offsets, *, gap_open_score=0.0, gap_extend_score=0.0,
              temperature=1.0, L1max=None, L2max=None) -> Tensor:
    """Packed affine soft-SW map (differentiable in the scores)."""
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    return _PackedMap.apply(flat, lengths, offsets, True,
                            _params(gap_open_score, gap_extend_score, temperature), L1, L2)


def sw_value(flat, lengths, offsets, *, gap_score=0.0, temperature=1.0, L1max=None,
             L2max=None) -> Tensor:
    """Per-pair linear soft-SW value [B] of a packed batch."""
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    return _PackedValue.apply(flat, lengths, offsets, False, _params(gap_score, temperature), L1,
                              L2)


def sw(flat, lengths, offsets, *, gap_score=0.0, temperature=1.0, L1max=None,
       L2max=None) -> Tensor:
    """Packed linear soft-SW map."""
    lengths, offsets, L1, L2 = _check(flat, lengths, offsets, L1max, L2max)
    return _PackedMap.apply(flat, lengths, offsets, False, _params(gap_score, temperature), L1, L2)


__all__ = [
    "pack",
    "unpack",
    "sw_affine",
    "sw_affine_value",
    "sw_affine_forward",
    "sw_affine_tangent",
    "sw",
    "sw_value",
    "sw_forward",
    "sw_tangent",
]
