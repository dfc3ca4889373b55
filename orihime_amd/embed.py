"""Embedding-scored soft alignment, end to end.

Pair scores are inner products of per-position embeddings,
``s_b = Q_b K_b^T`` (BF16 inputs, FP32 accumulation), fed into the ragged
soft Smith-Waterman affine sweep, and differentiated back to the embeddings
through the map: with ``P_b = dS_b / ds_b`` the marginal map,

    dQ_b = g_b P_b K_b,    dK_b = g_b P_b^T Q_b

Everything runs through the C ABI (include/ohm_embed.h, include/ohm_sw.h).
On the GPU a ragged MFMA kernel writes each pair's M_b x N_b score block
into the sweeps' packed layout, the packed persistent sweeps run on it, and
two ragged MFMA kernels contract the packed map with K and Q.

``fused=True`` takes the other GPU path: the sweeps generate the scores
themselves, each wave forming its strip's 16 x 16 score blocks with BF16
MFMA into LDS as it sweeps (ohm_embed_sw_affine_forward), so the scores
never go through memory.  It gives the same scores, values and maps bit for
bit; it applies for D = 32, 64 and 128 within the exponent-tracked gap range
(``fused_applies``), and on 512 pairs of U[128, 768]^2 at D = 128 it is the
slower of the two (sweeps 2.27 ms against 0.24 ms of score kernel + 1.29 ms
of sweeps on scores in memory): those sweeps are bound by their instruction
latency, not by the score traffic, and the score blocks on each wave's
critical path, at two waves per SIMD for the registers they need, cost more
than the traffic saves.  CPU tensors run the library's host implementation
of the same entry points.

    value = soft_align_embeddings(Q, K, lengths, gap_open=-4.0, gap_ext=-0.5,
                                  temperature=0.25)
    value.sum().backward()      # Q.grad, K.grad

Q: [B, M, D], K: [B, N, D] bfloat16 (D = 32, 64, 128 or 256), same device;
lengths: [B, 2] int32 (M_b, N_b).  Rows past a pair's lengths are ignored and
receive zero gradient.
"""

from __future__ import annotations

import math

import torch

from . import _lib, _ops  # noqa: F401  (registers torch.ops.orihime)
from . import _checks


def _check(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor) -> None:
    if Q.dim() != 3 or K.dim() != 3 or Q.shape[0] != K.shape[0] or Q.shape[2] != K.shape[2]:
        raise ValueError("Q must be [B, M, D] and K [B, N, D] with matching B and D")
    if Q.dtype != torch.bfloat16 or K.dtype != torch.bfloat16:
        raise TypeError(f"Q and K must be torch.bfloat16, got {Q.dtype} and {K.dtype}")
    if Q.device != K.device:
        raise ValueError("Q and K must be on the same device")
    if Q.shape[2] not in (32, 64, 128, 256):
        raise ValueError(f"embedding dimension must be 32, 64, 128 or 256, got {Q.shape[2]}")
    if tuple(lengths.shape) != (Q.shape[0], 2):
        raise ValueError("lengths must be [B, 2] (M_b, N_b)")


def _lengths(lengths: torch.Tensor, Q: torch.Tensor, K: torch.Tensor) -> torch.Tensor:
    ln = lengths.to(device=Q.device, dtype=torch.int32).contiguous()
    if not (Q.is_cuda and torch.cuda.is_current_stream_capturing()):
        _checks.require((ln[:, 0] >= 0) & (ln[:, 0] <= Q.shape[1]) & (ln[:, 1] >= 0) & (ln[:, 1] <= K.shape[1]),
                "lengths entries must lie within the padded input shape")
    return ln


def pair_scores(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor
                ) -> tuple[torch.Tensor, torch.Tensor]:
    """Packed pair scores (capacity B * M * N floats; pair b's block at
    ``offsets[b]``, row stride N_b) and the int64 offsets."""
    _check(Q, K, lengths)
    Q, K = Q.contiguous(), K.contiguous()
    ln = _lengths(lengths, Q, K)
    B, M, _ = Q.shape
    N = K.shape[1]
    offsets = torch.empty(B, dtype=torch.int64, device=Q.device)
    scores = torch.empty(B * M * N, dtype=torch.float32, device=Q.device)
    _lib.embed_scores(Q, K, ln, offsets, scores)
    return scores, offsets


def embedding_grads(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor, offsets: torch.Tensor,
                    post: torch.Tensor, grad_value: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """``dQ_b = g_b P_b K_b`` and ``dK_b = g_b P_b^T Q_b`` (BF16) from the
    packed map ``post``."""
    Q, K = Q.contiguous(), K.contiguous()
    dQ, dK = torch.empty_like(Q), torch.empty_like(K)
    g = grad_value.to(device=Q.device, dtype=torch.float32).contiguous()
    _lib.embed_grads(Q, K, lengths, offsets, post, g, dQ, dK)
    return dQ, dK


# |gap| log2(e) / T bound of the exponent-tracked sweeps (the host check of
# use_scaled in the library)
_SCALED_GAP_LIMIT = 120.0


def fused_applies(Q: torch.Tensor, gap_open: float, gap_ext: float, temp: float) -> bool:
    """Whether the sweeps can generate the scores themselves for this call."""
    k = 1.0 / (math.log(2.0) * float(temp))
    return (Q.shape[2] in (32, 64, 128) and abs(gap_open * k) <= _SCALED_GAP_LIMIT
            and abs(gap_ext * k) <= _SCALED_GAP_LIMIT)


def _sweep(scores, lengths, offsets, M, N, gap_open, gap_ext, temp):
    B = lengths.shape[0]
    value = torch.empty(B, dtype=torch.float32, device=scores.device)
    post = torch.empty_like(scores)
    _lib.affine_forward(scores, lengths, offsets, gap_open, gap_ext, temp, value=value, map=post,
                        L1max=M, L2max=N)
    return value, post


def _fused_sweep(Q, K, lengths, gap_open, gap_ext, temp):
    B, M, _ = Q.shape
    N = K.shape[1]
    dev = Q.device
    offsets = torch.empty(B, dtype=torch.int64, device=dev)
    scores = torch.empty(B * M * N, dtype=torch.float32, device=dev)  # flagged pairs only
    value = torch.empty(B, dtype=torch.float32, device=dev)
    post = torch.empty(B * M * N, dtype=torch.float32, device=dev)
    _lib.embed_sw_affine(Q, K, lengths, gap_open, gap_ext, temp, offsets=offsets, scores=scores,
                         value=value, map=post)
    return value, post, offsets


class _SoftAlignEmbeddings(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, K, lengths, gap_open, gap_ext, temp, fused):
        Q, K = Q.contiguous(), K.contiguous()
        ln = _lengths(lengths, Q, K)
        if fused:
            value, post, offsets = _fused_sweep(Q, K, ln, gap_open, gap_ext, temp)
        else:
            scores, offsets = pair_scores(Q, K, ln)
            value, post = _sweep(scores, ln, offsets, Q.shape[1], K.shape[1], gap_open, gap_ext,
                                 temp)
        ctx.save_for_backward(Q, K, ln, offsets, post)
        return value

    @staticmethod
    def backward(ctx, grad_value):
        Q, K, ln, offsets, post = ctx.saved_tensors
        dQ, dK = embedding_grads(Q, K, ln, offsets, post, grad_value)
        return dQ, dK, None, None, None, None, None


def soft_align_embeddings(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor,
                          gap_open: float = -4.0, gap_ext: float = -0.5,
                          temperature: float = 0.25, fused: bool = False) -> torch.Tensor:
    """Soft Smith-Waterman (affine) value of every pair, scored by ``Q K^T``;
    differentiable in Q and K.  ``fused``: generate the scores inside the
    sweeps instead of a score kernel (GPU tensors; see ``fused_applies``)."""
    _check(Q, K, lengths)
    _checks.temperature(temperature)
    go, ge, T = float(gap_open), float(gap_ext), float(temperature)
    if fused and not (Q.is_cuda and fused_applies(Q, go, ge, T)):
        raise ValueError("the fused sweeps take GPU tensors, D = 32, 64 or 128 and "
                         "|gap| log2(e) / T <= 120")
    return _SoftAlignEmbeddings.apply(Q, K, lengths, go, ge, T, bool(fused))
