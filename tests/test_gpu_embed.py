"""The ragged BF16 MFMA kernels of the embedding-scored path
(ohm_embed_scores, ohm_embed_grads) against float64 products of the same
BF16 values, on shapes that exercise tile edges: lengths of 0 and 1, lengths
that are not multiples of the 64-row tiles or the 32-wide reduction runs,
rows that start at any 4-byte offset of the packed map, and every supported
embedding dimension."""

import numpy as np
import pytest
import torch

from orihime_amd import _lib

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

CASES = [
    # B, Lq, Lk, D, lengths
    (3, 70, 90, 128, [[70, 90], [1, 1], [33, 65]]),
    (4, 64, 64, 32, [[64, 64], [0, 10], [10, 0], [63, 1]]),
    (2, 129, 200, 64, [[129, 200], [100, 17]]),
    (2, 40, 50, 256, [[40, 50], [39, 49]]),
    (5, 768, 768, 128, [[768, 768], [128, 700], [513, 129], [767, 1], [300, 333]]),
]


def _inputs(B, Lq, Lk, D, seed):
    g = torch.Generator().manual_seed(seed)
    Q = (torch.randn(B, Lq, D, generator=g) / D ** 0.5).bfloat16()
    K = (torch.randn(B, Lk, D, generator=g) / D ** 0.5).bfloat16()
    return Q, K


def _offsets(ln):
    sizes = [int(m) * int(n) for m, n in ln]
    return np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64), sum(sizes)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"B{c[0]}_{c[1]}x{c[2]}_D{c[3]}")
def test_scores_packed_match_float64(case):
    B, Lq, Lk, D, lens = case
    print(f"[gpu case] embed scores {case[:4]} lengths {lens}", flush=True)
    Q, K = _inputs(B, Lq, Lk, D, seed=B * 7 + D)
    ln = torch.tensor(lens, dtype=torch.int32)
    off = torch.empty(B, dtype=torch.int64, device=DEV)
    S = torch.full((B * Lq * Lk,), float("nan"), device=DEV)
    _lib.embed_scores(Q.to(DEV), K.to(DEV), ln.to(DEV), off, S)
    torch.cuda.synchronize()
    want_off, total = _offsets(lens)
    assert np.array_equal(off.cpu().numpy(), want_off)
    Sh = S.cpu().numpy()
    assert np.isnan(Sh[total:]).all()  # nothing written past the packed blocks
    Q64, K64 = Q.double().numpy(), K.double().numpy()
    for b, (m, n) in enumerate(lens):
        got = Sh[want_off[b]:want_off[b] + m * n].reshape(m, n)
        ref = Q64[b, :m] @ K64[b, :n].T
        bound = np.abs(Q64[b, :m]) @ np.abs(K64[b, :n]).T
        assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-7), b


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"B{c[0]}_{c[1]}x{c[2]}_D{c[3]}")
def test_grads_match_float64_and_zero_past_lengths(case):
    B, Lq, Lk, D, lens = case
    print(f"[gpu case] embed grads {case[:4]} lengths {lens}", flush=True)
    Q, K = _inputs(B, Lq, Lk, D, seed=B * 11 + D)
    want_off, total = _offsets(lens)
    g = torch.Generator().manual_seed(D + Lq)
    P = torch.rand(max(total, 1), generator=g)  # any FP32 map in [0, 1)
    gv = torch.randn(B, generator=g)
    ln = torch.tensor(lens, dtype=torch.int32)
    dQ = torch.full((B, Lq, D), float("nan"), dtype=torch.bfloat16, device=DEV)
    dK = torch.full((B, Lk, D), float("nan"), dtype=torch.bfloat16, device=DEV)
    _lib.embed_grads(Q.to(DEV), K.to(DEV), ln.to(DEV), torch.from_numpy(want_off).to(DEV),
                     P.to(DEV), gv.to(DEV), dQ, dK)
    torch.cuda.synchronize()
    gq, gk = dQ.float().cpu().numpy(), dK.float().cpu().numpy()
    Q64, K64, P64, g64 = Q.double().numpy(), K.double().numpy(), P.double().numpy(), gv.double().numpy()
    for b, (m, n) in enumerate(lens):
        Pb = P64[want_off[b]:want_off[b] + m * n].reshape(m, n)
        rq = g64[b] * (Pb @ K64[b, :n])
        rk = g64[b] * (Pb.T @ Q64[b, :m])
        bq = abs(g64[b]) * (Pb @ np.abs(K64[b, :n]))
        bk = abs(g64[b]) * (Pb.T @ np.abs(Q64[b, :m]))
        # BF16 outputs (8 significant bits: rounding within 2^-8 of the
        # entry), plus the products' FP32 accumulation and the map's two-term
        # BF16 split
        assert np.all(np.abs(gq[b, :m] - rq) <= 2.0 ** -8 * np.abs(rq) + 2e-5 * bq + 1e-30), b
        assert np.all(np.abs(gk[b, :n] - rk) <= 2.0 ** -8 * np.abs(rk) + 2e-5 * bk + 1e-30), b
        assert not np.any(gq[b, m:]) and not np.any(gk[b, n:]), b  # exact zeros past lengths


def test_grads_and_scores_are_bitwise_repeatable():
    B, Lq, Lk, D = 6, 300, 257, 128
    Q, K = _inputs(B, Lq, Lk, D, seed=3)
    ln = torch.tensor([[300, 257], [1, 200], [299, 3], [150, 150], [0, 5], [64, 64]],
                      dtype=torch.int32).to(DEV)
    Qd, Kd = Q.to(DEV), K.to(DEV)
    outs = []
    for _ in range(2):
        off = torch.empty(B, dtype=torch.int64, device=DEV)
        S = torch.zeros(B * Lq * Lk, device=DEV)
        _lib.embed_scores(Qd, Kd, ln, off, S)
        gv = torch.ones(B, device=DEV)
        dQ, dK = torch.empty_like(Qd), torch.empty_like(Kd)
        _lib.embed_grads(Qd, Kd, ln, off, S.sigmoid(), gv, dQ, dK)
        outs.append((S, dQ, dK))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


# --------------------------------------------------------------------------
# Scores generated inside the sweeps (ohm_embed_sw_affine_forward)
# --------------------------------------------------------------------------

from orihime_amd import embed  # noqa: E402

FUSED_CASES = [
    # B, Lq, Lk, D, lengths
    (5, 300, 257, 128, [[300, 257], [1, 200], [299, 3], [150, 150], [64, 64]]),
    (4, 70, 90, 64, [[70, 90], [1, 1], [33, 65], [0, 7]]),
    (3, 200, 130, 32, [[200, 130], [129, 1], [65, 128]]),
]


def _unfused(Qd, Kd, ln, B, Lq, Lk, params):
    """Score kernel, then the per-pair fused forward + reverse sweep kernel
    with four waves: the launch the embedding sweep kernel mirrors."""
    off = torch.empty(B, dtype=torch.int64, device=DEV)
    S = torch.zeros(B * Lq * Lk, device=DEV)
    _lib.embed_scores(Qd, Kd, ln, off, S)
    out = [torch.empty(B, device=DEV), torch.full_like(S, float("nan"))] + \
        [torch.empty(B, device=DEV) for _ in range(3)]
    try:
        _lib.set_launch_config(4, 0, 4)
        _lib.affine_forward(S, ln, off, *params, value=out[0], map=out[1], grad_open=out[2],
                            grad_ext=out[3], grad_temp=out[4], L1max=Lq, L2max=Lk)
    finally:
        _lib.set_launch_config()
    return off, S, out


def _fused(Qd, Kd, ln, B, Lq, Lk, params):
    off = torch.empty(B, dtype=torch.int64, device=DEV)
    S = torch.full((B * Lq * Lk,), float("nan"), device=DEV)
    out = [torch.empty(B, device=DEV), torch.full_like(S, float("nan"))] + \
        [torch.empty(B, device=DEV) for _ in range(3)]
    _lib.embed_sw_affine(Qd, Kd, ln, *params, offsets=off, scores=S, value=out[0], map=out[1],
                         grad_open=out[2], grad_ext=out[3], grad_temp=out[4])
    return off, S, out


@pytest.mark.parametrize("case", FUSED_CASES, ids=lambda c: f"B{c[0]}_{c[1]}x{c[2]}_D{c[3]}")
def test_fused_sweeps_equal_score_kernel_plus_sweeps_bitwise(case):
    """The sweeps that generate their scores on the matrix cores give the
    bits of the score kernel followed by the same sweep launch, and never
    write the score buffer for pairs they keep."""
    B, Lq, Lk, D, lens = case
    print(f"[gpu case] fused embedding sweeps {case[:4]} lengths {lens}", flush=True)
    Q, K = _inputs(B, Lq, Lk, D, seed=B * 13 + D)
    Qd, Kd = Q.to(DEV), K.to(DEV)
    ln = torch.tensor(lens, dtype=torch.int32).to(DEV)
    params = (-4.0, -0.5, 0.25)
    off_u, _, ref = _unfused(Qd, Kd, ln, B, Lq, Lk, params)
    off_f, S_f, got = _fused(Qd, Kd, ln, B, Lq, Lk, params)
    torch.cuda.synchronize()
    assert torch.equal(off_u, off_f)
    total = sum(int(m) * int(n) for m, n in lens)
    for k, (a, b) in enumerate(zip(ref, got)):
        if k == 1:
            a, b = a[:total], b[:total]
        assert torch.equal(a, b), k
    assert torch.isnan(S_f).all()  # no pair was handed to the log-domain sweeps


def test_fused_flagged_pair_scores_written_for_the_log_domain_redo():
    """A pair whose scores jump from -80 T to +80 T between adjacent columns
    leaves the exponent-tracked range: the fused kernel writes its scores to
    the packed buffer and the log-domain sweeps redo it, as in the score
    kernel + sweep path; the other pair stays in the fused kernel."""
    B, Lq, Lk, D = 2, 40, 50, 32
    print("[gpu case] fused embedding sweeps, flagged pair", flush=True)
    Q, K = _inputs(B, Lq, Lk, D, seed=77)
    Q[0] = 0.0
    K[0] = 0.0
    Q[0, :, 0] = 4.0
    K[0, :, 0] = -5.0
    K[0, 25, 0] = 5.0  # column 25: +20 = 80 T, every other column -20 = -80 T
    Qd, Kd = Q.to(DEV), K.to(DEV)
    lens = [[40, 50], [40, 50]]
    ln = torch.tensor(lens, dtype=torch.int32).to(DEV)
    params = (-4.0, -0.5, 0.25)
    off_u, S_u, ref = _unfused(Qd, Kd, ln, B, Lq, Lk, params)
    off_f, S_f, got = _fused(Qd, Kd, ln, B, Lq, Lk, params)
    torch.cuda.synchronize()
    flagged = torch.isfinite(S_f[:2000])
    assert flagged.all() and torch.isnan(S_f[2000:4000]).all()  # pair 0 written, pair 1 not
    assert torch.equal(S_f[:2000], S_u[:2000])
    for k, (a, b) in enumerate(zip(ref, got)):
        if k == 1:
            a, b = a[:4000], b[:4000]
        assert torch.equal(a, b), k
    s64 = (Q.double() @ K.double().transpose(1, 2)).numpy()
    from oracle import oracle as O
    f64 = O.sw_affine_forward(s64, *params, np.array(lens, np.int32), prec="f64")
    assert np.max(np.abs(got[0].cpu().numpy() - f64[0]) / np.maximum(1e-3, np.abs(f64[0]))) < 1e-5


def test_soft_align_embeddings_fused_and_unfused_agree():
    """Value and embedding gradients through autograd, fused against the
    score kernel path under the same sweep launch."""
    B, Lq, Lk, D = 4, 160, 140, 128
    print("[gpu case] soft_align_embeddings fused vs unfused", flush=True)
    Q, K = _inputs(B, Lq, Lk, D, seed=5)
    ln = torch.tensor([[160, 140], [17, 100], [159, 1], [80, 140]], dtype=torch.int32).to(DEV)
    outs = []
    for fused in (True, False):
        q = Q.to(DEV).requires_grad_(True)
        k = K.to(DEV).requires_grad_(True)
        try:
            _lib.set_launch_config(4, 0, 4)
            v = embed.soft_align_embeddings(q, k, ln, fused=fused)
            v.sum().backward()
        finally:
            _lib.set_launch_config()
        outs.append((v.detach(), q.grad, k.grad))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
