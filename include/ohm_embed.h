/*
 * ohm_embed.h — C ABI of the embedding-scored alignment path: pair scores
 * formed from per-position embeddings, s_b = Q_b K_b^T, fed into the ragged
 * soft Smith-Waterman sweeps (ohm_sw.h), and the embeddings' gradients
 * through the marginal map.  Built into orihime_amd/libohm_sw.so.
 *
 * The reference has no fused path: its callers form Q K^T themselves and
 * pass the scores to orihime.sw_affine (orihime/sw.py:116-141).  These entry
 * points replace that caller-side GEMM and the two map GEMMs of its backward
 * with ragged BF16 MFMA kernels that touch only each pair's own box.
 *
 * Layout:
 *   Q        BF16 [B, Lq, D] row-major (row stride D), pair b's rows 0..M_b-1 used
 *   K        BF16 [B, Lk, D] likewise, rows 0..N_b-1 used
 *   lengths  int32 [B, 2] = (M_b, N_b), 0 <= M_b <= Lq, 0 <= N_b <= Lk
 *   offsets  int64 [B], written by ohm_embed_scores: the exclusive prefix sum
 *            of M_b * N_b, i.e. the packed layout of ohm_sw.h
 *   scores   FP32, packed: pair b's M_b x N_b block at offsets[b] (row
 *            stride N_b); capacity B * Lq * Lk floats bounds every batch
 *   map      FP32 packed like scores (the marginal map of the sweep)
 * D is 32, 64, 128 or 256; Q, K, dQ, dK must be 16-byte aligned.
 *
 * GPU contract as in ohm_sw.h: enqueue-only on `stream`, no allocation, no
 * host synchronisation, bitwise repeatable (fixed-order reductions, no
 * atomics).  The *_cpu twins take host pointers and compute in FP32 from
 * the same BF16 values.
 */
#ifndef OHM_EMBED_H
#define OHM_EMBED_H

#include <stddef.h>
#include <stdint.h>

#include "ohm_sw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Packed pair scores s_b[i, j] = sum_d Q[b, i, d] K[b, j, d] (BF16 products,
 * FP32 accumulation on the matrix cores) and the packed offsets.
 * Replaces: the caller-side Q K^T GEMM before sw_affine (orihime/sw.py:116-141). */
ohm_status ohm_embed_scores(const void* Q, const void* K, const int32_t* lengths, int32_t B,
                            int32_t Lq, int32_t Lk, int32_t D, int64_t* offsets, float* scores,
                            void* stream);
ohm_status ohm_embed_scores_cpu(const void* Q, const void* K, const int32_t* lengths, int32_t B,
                                int32_t Lq, int32_t Lk, int32_t D, int64_t* offsets,
                                float* scores);

/* The packed offsets alone (the exclusive prefix sum of M_b * N_b), as
 * ohm_embed_scores writes them. */
ohm_status ohm_embed_offsets(const int32_t* lengths, int32_t B, int32_t Lq, int32_t Lk,
                             int64_t* offsets, void* stream);

/* Soft Smith-Waterman (affine) of every pair scored by Q_b K_b^T, with the
 * scores generated inside the sweeps: each wave forms its strip's 16 x 16
 * score blocks on the matrix cores (the BF16 products and FP32 reduction
 * order of ohm_embed_scores, so each score is bitwise the one it writes)
 * into LDS, so the scores never go through memory.  Outputs as
 * ohm_sw_affine_forward on the packed layout (value, map, the three
 * parameter gradients; map and gradients nullable) and the packed offsets.
 * `scores` (capacity B * Lq * Lk floats) is written only for the pairs that
 * leave the exponent-tracked range, which the log-domain sweeps then redo
 * from it.  D is 32, 64 or 128; the gap factors must be in the
 * exponent-tracked range (|gap| log2(e) / T <= 120; otherwise: OHM_EINVAL,
 * use ohm_embed_scores and ohm_sw_affine_forward).  Workspace:
 * ohm_sw_affine_workspace(OHM_OP_FORWARD, B, Lq, Lk, B * Lq * Lk).
 * Bit for bit the result of ohm_embed_scores + ohm_sw_affine_forward; on
 * short pairs (512 pairs of U[128, 768]^2, D = 128) those two are the
 * faster route (1.53 against 2.27 ms): the sweeps there are latency-bound
 * and the score blocks sit on each wave's critical path.
 * Replaces: the caller-side Q K^T GEMM and the sweep over its result
 * (orihime/sw.py:116-141 with orihime/ops/sw_affine.py). */
ohm_status ohm_embed_sw_affine_forward(const void* Q, const void* K, const int32_t* lengths,
                                       int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                                       float gap_open, float gap_ext, float temp,
                                       int64_t* offsets, float* scores, float* value, float* map,
                                       float* grad_open, float* grad_ext, float* grad_temp,
                                       void* workspace, size_t workspace_bytes, void* stream);
/* Host twin: ohm_embed_scores_cpu, then ohm_sw_affine_forward_cpu (scores
 * written for every pair). */
ohm_status ohm_embed_sw_affine_forward_cpu(const void* Q, const void* K, const int32_t* lengths,
                                           int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                                           float gap_open, float gap_ext, float temp,
                                           int64_t* offsets, float* scores, float* value,
                                           float* map, float* grad_open, float* grad_ext,
                                           float* grad_temp, int32_t threads);

/* Workspace bytes of ohm_embed_grads (transposed copies of Q and K). */
size_t ohm_embed_grads_workspace(int32_t B, int32_t Lq, int32_t Lk, int32_t D);

/* Gradients of sum_b g_b S_b through the marginal map P_b = dS_b/ds_b:
 *   dQ[b, i, :] = g_b sum_j P_b[i, j] K[b, j, :]    (i < M_b; rows >= M_b are 0)
 *   dK[b, j, :] = g_b sum_i P_b[i, j] Q[b, i, :]    (j < N_b; rows >= N_b are 0)
 * grad_value: FP32 [B] (g_b).  dQ [B, Lq, D], dK [B, Lk, D] BF16 (the inputs'
 * dtype; FP32 accumulation, P split into two BF16 terms so the map's FP32
 * precision reaches the products).
 * Replaces: the two GEMMs of the caller's backward through the map. */
ohm_status ohm_embed_grads(const void* Q, const void* K, const int32_t* lengths,
                           const int64_t* offsets, int32_t B, int32_t Lq, int32_t Lk, int32_t D,
                           const float* map, const float* grad_value, void* dQ, void* dK,
                           void* workspace, size_t workspace_bytes, void* stream);
ohm_status ohm_embed_grads_cpu(const void* Q, const void* K, const int32_t* lengths,
                               const int64_t* offsets, int32_t B, int32_t Lq, int32_t Lk,
                               int32_t D, const float* map, const float* grad_value, void* dQ,
                               void* dK);

#ifdef __cplusplus
}
#endif

#endif /* OHM_EMBED_H */
