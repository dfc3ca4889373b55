"""Embedding-scored soft alignment, end to end.

Pair scores are inner products of per-position embeddings,
``s_b = Q_b K_b^T`` (BF16 inputs, FP32 accumulation and output: one batched
MFMA GEMM through hipBLASLt), fed into the ragged soft Smith-Waterman affine
sweep (padded storage with per-pair lengths: cells past a pair's lengths are
never computed), and differentiated back to the embeddings through the map:
with ``P_b = dS_b / ds_b`` the marginal map,

    dQ_b = g_b (P_b K_b),    dK_b = g_b (Q_b^T P_b)^T

(FP32 GEMMs on the FP32 map).  Rows and columns past a pair's lengths have a
zero map and so receive zero gradient.

    value = soft_align_embeddings(Q, K, lengths, gap_open=-4.0, gap_ext=-0.5,
                                  temperature=0.25)
    value.sum().backward()      # Q.grad, K.grad

Q: [B, M, D], K: [B, N, D] (bfloat16 or float32, same device), lengths:
[B, 2] int32 (M_b, N_b).
"""

from __future__ import annotations

import torch

from . import _ops  # noqa: F401  (registers torch.ops.orihime)


def pair_scores(Q: torch.Tensor, K: torch.Tensor) -> torch.Tensor:
    """``Q K^T`` per pair in FP32: BF16 products with FP32 accumulation."""
    Kt = K.transpose(1, 2).contiguous()
    if Q.is_cuda and Q.dtype == torch.bfloat16:
        return torch.bmm(Q, Kt, out_dtype=torch.float32)
    return torch.bmm(Q.float(), Kt.float())


class _SoftAlignEmbeddings(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Q, K, lengths, gap_open, gap_ext, temperature):
        scores = pair_scores(Q, K)
        value, post = torch.ops.orihime.sw_affine_forward(scores, float(gap_open), float(gap_ext),
                                                          float(temperature), lengths)
        ctx.save_for_backward(Q, K, post)
        return value

    @staticmethod
    def backward(ctx, grad_value):
        # FP32 GEMMs on the FP32 map; the per-pair factor g is applied to the
        # small [B, *, D] products, and dK is formed as (Q^T P)^T, so the
        # [B, M, N] map is only read by the two GEMMs (no scaled or transposed
        # copy of it)
        Q, K, post = ctx.saved_tensors
        gv = grad_value.to(torch.float32)[:, None, None]
        dQ = torch.bmm(post, K.float()) * gv
        dKt = torch.bmm(Q.float().transpose(1, 2).contiguous(), post)
        dK = dKt.transpose(1, 2) * gv
        return dQ.to(Q.dtype), dK.to(K.dtype), None, None, None, None


def soft_align_embeddings(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor,
                          gap_open: float = -4.0, gap_ext: float = -0.5,
                          temperature: float = 0.25) -> torch.Tensor:
    """Soft Smith-Waterman (affine) value of every pair, scored by ``Q K^T``;
    differentiable in Q and K."""
    if Q.dim() != 3 or K.dim() != 3 or Q.shape[0] != K.shape[0] or Q.shape[2] != K.shape[2]:
        raise ValueError("Q must be [B, M, D] and K [B, N, D] with matching B and D")
    if lengths.shape != (Q.shape[0], 2):
        raise ValueError("lengths must be [B, 2] (M_b, N_b)")
    return _SoftAlignEmbeddings.apply(Q, K, lengths.to(device=Q.device, dtype=torch.int32),
                                      gap_open, gap_ext, temperature)
