"""Launch-shape and tile autotuner for the GPU sweeps.

The kernels' results do not depend on the launch shape (waves per workgroup,
persistent workgroups per CU, per-pair workgroups or wave-scheduled team
kernels) nor on the tile of the affine sweeps' adjoint (the checkpoint
stride: anti-diagonal steps per recomputed block, 4, 8 or 16) — only their
speed does, and the best choice depends on the batch: many mid-size pairs
favour 4 waves per pair, few long pairs the team or grid kernels; a shorter
stride stores more checkpoints and holds fewer registers per block.
``autotune`` times the candidate shapes on the caller's own batch (the best
of five forward+backward calls each, after a warm-up), then the strides under the
fastest shape (affine sweeps only: the linear sweeps keep full tables),
applies the winner through ``ohm_set_launch_config`` and
``ohm_set_checkpoint_stride`` and caches it per batch signature (pair count
bucket, maximum lengths, layout), so a training loop pays the search once.

    cfg = orihime_amd.tune.autotune(scores, 2.0, -0.5, 1.0, lengths=lens)

``reset()`` returns to the automatic choice.  Workspace sizes depend on the
stride: size workspaces after tuning (the library's own callers query the
size on every call).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from . import _lib

# candidate shapes: (waves, groups per CU, teams) with teams 1 = team kernels,
# 2 = per-pair workgroup kernels, 3 = grid kernels (batches of at most
# GRID_MAX_PAIRS pairs), 4 = per-pair workgroups running each pair's forward
# and reverse sweep in one launch
CANDIDATES = (
    (4, 0, 2),
    (4, 8, 2),
    (2, 0, 2),
    (8, 0, 2),
    (0, 0, 1),
    (0, 0, 3),
    (4, 0, 4),
    (4, 8, 4),
    (8, 0, 4),
)
GRID_MAX_PAIRS = 64
# checkpoint strides of the affine sweeps (steps per recomputed block)
STRIDES = (8, 4, 16)


@dataclass(frozen=True)
class LaunchConfig:
    waves: int
    groups_per_cu: int
    teams: int
    ms: float
    ckpt: int = 0  # checkpoint stride (0: automatic)


_cache: dict[tuple, LaunchConfig] = {}


def _signature(scores, lengths, offsets, affine):
    B = int(lengths.shape[0]) if lengths is not None else int(scores.shape[0])
    if offsets is not None:
        L1max = int(lengths[:, 0].max())
        L2max = int(lengths[:, 1].max())
    else:
        L1max, L2max = int(scores.shape[-2]), int(scores.shape[-1])
    bucket = 1 << max(0, (B - 1).bit_length())
    return (affine, offsets is not None, bucket, L1max, L2max, scores.device.index)


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def autotune(scores, gap_open, gap_ext=None, temperature=1.0, *, lengths=None, offsets=None,
             reps: int = 5, candidates=CANDIDATES, strides=STRIDES,
             apply: bool = True) -> LaunchConfig:
    """Choose the fastest launch shape for this batch (affine when ``gap_ext``
    is given, linear otherwise) and apply it.  ``scores`` is a GPU tensor in
    the padded ([B, L1, L2]) or packed layout (``offsets``, see
    orihime_amd.ragged)."""
    if not scores.is_cuda:
        raise ValueError("autotune needs GPU tensors")
    affine = gap_ext is not None
    key = _signature(scores, lengths, offsets, affine)
    if key in _cache:
        cfg = _cache[key]
    else:
        B = int(lengths.shape[0]) if lengths is not None else int(scores.shape[0])
        L1max, L2max = key[3], key[4]
        total = -1 if offsets is None else int(scores.numel())
        strides = strides if affine else ()
        # the finest stride needs the largest workspace
        _lib.set_checkpoint_stride(min(strides) if strides else 0)
        try:
            nbytes = _lib.workspace_bytes(affine, _lib.OP_FORWARD, B, L1max, L2max, total)
        finally:
            _lib.set_checkpoint_stride()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=scores.device)
        value = torch.empty(B, device=scores.device)
        mp = torch.empty_like(scores)
        g0, g1, g2 = (torch.empty(B, device=scores.device) for _ in range(3))

        def call():
            if affine:
                _lib.affine_forward(scores, lengths, offsets, gap_open, gap_ext, temperature,
                                    value=value, map=mp, grad_open=g0, grad_ext=g1, grad_temp=g2,
                                    L1max=L1max, L2max=L2max, workspace=ws)
            else:
                _lib.linear_forward(scores, lengths, offsets, gap_open, temperature, value=value,
                                    map=mp, grad_gap=g0, grad_temp=g1, L1max=L1max, L2max=L2max,
                                    workspace=ws)

        results = []
        try:
            for w, g, t in candidates:
                if t == 3 and B > GRID_MAX_PAIRS:
                    continue
                _lib.set_launch_config(w, g, t)
                results.append(LaunchConfig(w, g, t, _time(call, reps)))
            best = min(results, key=lambda c: c.ms)
            _lib.set_launch_config(best.waves, best.groups_per_cu, best.teams)
            default = _lib.checkpoint_stride()  # what the shapes above ran with
            for ck in strides:
                if ck == default:
                    continue
                _lib.set_checkpoint_stride(ck)
                results.append(LaunchConfig(best.waves, best.groups_per_cu, best.teams,
                                            _time(call, reps), ck))
        finally:
            _lib.set_launch_config()
            _lib.set_checkpoint_stride()
        cfg = min(results, key=lambda c: c.ms)
        _cache[key] = cfg
    if apply:
        _lib.set_launch_config(cfg.waves, cfg.groups_per_cu, cfg.teams)
        _lib.set_checkpoint_stride(cfg.ckpt)
    return cfg


def reset() -> None:
    """Back to the automatic launch shape and stride (the cache is kept)."""
    _lib.set_launch_config()
    _lib.set_checkpoint_stride()
