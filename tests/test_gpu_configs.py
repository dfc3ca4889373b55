"""The benchmark workloads at their real sizes, checked against the oracle.

* The ragged batch of the benchmark (2048 pairs, L1, L2 ~ U[64, 1024],
  torch seed 2; go=-8, ge=-0.5, T=0.5), packed: every value against the
  FP32 reference kernels and the FP64 oracle; maps and parameter gradients of
  64 pairs spanning the size range.
* The same batch padded to [2048, 1024, 1024] with the cells beyond each
  pair's lengths masked: every value against the reference kernels on the
  same sentinel-filled input (the mask leaves gap paths into the padding, so
  the padded launch computes a different function from the packed one).
* The fixed batch (256 x 512 x 512, seed 1): HVP and temperature sensitivity
  at size, sampled pairs against the oracle.
* The first pair of the long batch (16 x 4096 x 4096, seed 3): the HVP along
  its N(0,1) direction against the oracle.
* The embedding-scored batch at 64 pairs (M, N ~ U[128, 768], seed 4): value
  and dQ, dK against the FP64 oracle on the same BF16 values.

Tolerances are the reference suite's (values 1e-5 relative, maps 1e-4
absolute, tests/test_soft_sw_affine.py:29-30); second-order outputs follow
helpers.close_or_nearer_truth.
"""

import numpy as np
import pytest
import torch

import bench
from helpers import abs_err, close_or_nearer_truth, rel_err
from oracle import oracle as O
from orihime_amd import ragged
from orihime_amd.operator import _normalize_masked

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
IMPL = "ref" if O.ref_available() else "oracle"


def case(msg):
    print(f"[gpu case] {msg}", flush=True)


def h(t):
    return t.detach().cpu().numpy()


def _spread(sizes, n):
    """Indices of n pairs at evenly spaced ranks of cell count."""
    order = np.argsort(np.asarray(sizes), kind="stable")
    return sorted(int(order[k]) for k in np.linspace(0, len(order) - 1, n).round().astype(int))


@pytest.fixture(scope="module")
def c3_batch():
    pairs = [p.numpy() for p in bench.ragged_pairs(torch)]
    return pairs


def test_c3_packed_every_value_and_spread_maps(c3_batch):
    pairs = c3_batch
    go, ge, T = -8.0, -0.5, 0.5
    case(f"C3 packed fwd+bwd, {len(pairs)} pairs, {sum(p.size for p in pairs)} cells")
    flat, ln, off = ragged.pack([torch.from_numpy(p) for p in pairs])
    v, m, a, b, c = ragged.sw_affine_forward(flat.to(DEV), ln.to(DEV), off.to(DEV),
                                             gap_open_score=go, gap_extend_score=ge, temperature=T)
    torch.cuda.synchronize()
    vh = h(v)
    ref_v = np.asarray(O.over_pairs(O.sw_affine_value, pairs, go, ge, T, impl=IMPL, chunk=64)[0])
    err = rel_err(vh, ref_v, 1e-3)
    f64_v = np.asarray(O.over_pairs(O.sw_affine_value, pairs, go, ge, T, prec="f64", chunk=16)[0])
    print(f"  all {len(pairs)} values: rel err vs FP32 reference {err:.3e}, "
          f"vs FP64 oracle {rel_err(vh, f64_v, 1e-3):.3e}", flush=True)
    assert err < 1e-5
    assert rel_err(vh, f64_v, 1e-3) < 1e-5
    # maps and parameter gradients of 64 pairs spread over the size range,
    # against the FP64 oracle (at a million cells the FP32 reference's own map
    # carries ~1e-4 of rounding; its distance is reported beside)
    idx = _spread([p.size for p in pairs], 64)
    sub = [pairs[i] for i in idx]
    tv, tm, ta, tb, tc = O.over_pairs(O.sw_affine_forward, sub, go, ge, T, prec="f64", chunk=1)
    rv, rm, ra, rb, rc = O.over_pairs(O.sw_affine_forward, sub, go, ge, T, impl=IMPL, chunk=8)
    maps = ragged.unpack(m, ln.to(DEV), off.to(DEV))
    worst = worst_ref = 0.0
    for k, i in enumerate(idx):
        e = abs_err(h(maps[i]), tm[k])
        worst, worst_ref = max(worst, e), max(worst_ref, abs_err(rm[k], tm[k]))
        assert e < 1e-4, (i, pairs[i].shape, e, abs_err(rm[k], tm[k]))
    for got, want in ((a, ta), (b, tb)):
        assert rel_err(h(got)[idx], np.asarray(want), 1.0) < 1e-4
    # the temperature gradient (S - E[score]) / T cancels: its FP32 noise grows
    # with |S| / T, so it passes within 1e-4 or twice the FP32 reference's error
    gt, t64, r32 = h(c)[idx].astype(np.float64), np.asarray(tc), np.asarray(rc, np.float64)
    lim = np.maximum(1e-4 * np.maximum(1.0, np.abs(t64)), 2.0 * np.abs(r32 - t64))
    assert np.all(np.abs(gt - t64) <= lim), float(np.max(np.abs(gt - t64) - lim))
    print(f"  64 spread pairs ({min(pairs[i].size for i in idx)}..{max(pairs[i].size for i in idx)} "
          f"cells): map max abs err vs FP64 {worst:.3e} (FP32 reference vs FP64: {worst_ref:.3e})",
          flush=True)


def test_c3_batch_linear_gap_values_and_spread_maps(c3_batch, monkeypatch):
    """The ragged batch through the linear-gap sweeps (the reference
    harness's linear parameters, gap -0.75, T 0.73): every value and the maps
    and gradients of 16 pairs spread over the size range against the FP64
    oracle, for the exponent-tracked sweeps and (reported beside) the
    log-domain sweeps."""
    pairs = c3_batch
    gp, T = -0.75, 0.73
    case(f"C3 batch linear gap fwd+bwd, {len(pairs)} pairs")
    flat, ln, off = ragged.pack([torch.from_numpy(p) for p in pairs])
    flat, ln, off = flat.to(DEV), ln.to(DEV), off.to(DEV)
    runs = {}
    for env in ("1", "0"):
        monkeypatch.setenv("OHM_LINEAR_SCALED", env)
        runs[env] = ragged.sw_forward(flat, ln, off, gap_score=gp, temperature=T)
        torch.cuda.synchronize()
    v, m, a, c = runs["1"]
    f64_v = np.asarray(O.over_pairs(O.sw_forward, pairs, gp, T, prec="f64", chunk=16)[0])
    err, err_log = rel_err(h(v), f64_v, 1e-3), rel_err(h(runs["0"][0]), f64_v, 1e-3)
    print(f"  all values: rel err vs FP64 {err:.3e} (log-domain sweeps {err_log:.3e})", flush=True)
    assert err < 1e-5
    idx = _spread([p.size for p in pairs], 16)
    sub = [pairs[i] for i in idx]
    tv, tm, ta, tc = O.over_pairs(O.sw_forward, sub, gp, T, prec="f64", chunk=1)
    maps, maps_log = ragged.unpack(m, ln, off), ragged.unpack(runs["0"][1], ln, off)
    worst = worst_log = 0.0
    for k, i in enumerate(idx):
        e, el = abs_err(h(maps[i]), tm[k]), abs_err(h(maps_log[i]), tm[k])
        worst, worst_log = max(worst, e), max(worst_log, el)
    print(f"  16 spread pairs: map max abs err vs FP64 {worst:.3e} (log-domain sweeps "
          f"{worst_log:.3e})", flush=True)
    assert worst < 1e-4
    assert rel_err(h(a)[idx], np.asarray(ta), 1.0) < 1e-4


def test_c3_padded_mask_every_value_on_its_own_input(c3_batch):
    pairs = c3_batch
    go, ge, T = -8.0, -0.5, 0.5
    B, L = len(pairs), 1024
    case(f"C3 padded+mask fwd+bwd, [{B}, {L}, {L}]")
    pad = torch.full((B, L, L), float("-inf"), device=DEV)
    for i, p in enumerate(pairs):
        pad[i, :p.shape[0], :p.shape[1]] = torch.from_numpy(p).to(DEV)
    sent = _normalize_masked(pad, T).contiguous()
    del pad
    v, mp, *_ = torch.ops.orihime.soft_sw_affine_with_grads(sent, go, ge, T, None)
    vh = h(v)
    ref = np.empty(B, np.float32)
    step = 64
    for lo in range(0, B, step):
        ref[lo:lo + step] = O.sw_affine_value(h(sent[lo:lo + step]), go, ge, T, impl=IMPL)
    err = rel_err(vh, ref, 1e-3)
    print(f"  all {B} padded values: rel err vs FP32 reference on the sentinel input {err:.3e}",
          flush=True)
    assert err < 1e-5
    # whole maps (masked cells included) of a few pairs against the FP64 oracle
    idx = [0, 1000, B - 1]
    t64 = O.over_pairs(O.sw_affine_forward, [h(sent[i]) for i in idx], go, ge, T, prec="f64",
                       chunk=1)[1]
    for k, i in enumerate(idx):
        assert abs_err(h(mp[i]), t64[k]) < 1e-4, (i, pairs[i].shape)


def test_c2_hvp_and_temperature_sensitivity_at_size():
    go, ge, T = -11.0, -1.0, 1.0
    x = bench.fixed_batch(torch)
    gen = torch.Generator().manual_seed(11)
    V = torch.randn(x.shape, generator=gen)
    case("C2 [256, 512, 512] HVP and d(map)/dT")
    xd = x.to(DEV)
    hv = torch.ops.orihime.sw_affine_marginals_hvp(xd, V.to(DEV), go, ge, T, None)
    sT = torch.ops.orihime.sw_affine_marginals_grad_temp(xd, go, ge, T, None)
    idx = [0, 131, 255]
    sub, Vs = x[idx].numpy(), V[idx].numpy()
    r = O.sw_affine_hvp(sub, Vs, go, ge, T)
    truth = lambda: O.sw_affine_hvp(sub.astype(np.float64), Vs, go, ge, T, prec="f64")
    assert close_or_nearer_truth(h(hv[idx]), r, truth, 1e-4 * max(1.0, float(np.abs(r).max())))
    rT = O.sw_affine_param_jacobian(sub, 2, go, ge, T)
    truthT = lambda: O.sw_affine_param_jacobian(sub.astype(np.float64), 2, go, ge, T, prec="f64")
    assert close_or_nearer_truth(h(sT[idx]), rT, truthT, 1e-4 * max(1.0, float(np.abs(rT).max())))


def test_c4_first_pair_hvp_at_4096():
    go, ge, T = -11.0, -1.0, 1.0
    gen = torch.Generator().manual_seed(3)
    x = torch.randn(16, 4096, 4096, generator=gen)
    V = torch.randn(16, 4096, 4096, generator=gen)
    s, v = x[:1].clone(), V[:1].clone()
    del x, V
    case("C4 first pair 4096 x 4096 HVP")
    hv = h(torch.ops.orihime.sw_affine_marginals_hvp(s.to(DEV), v.to(DEV), go, ge, T, None))
    r = O.sw_affine_hvp(s.numpy(), v.numpy(), go, ge, T, impl=IMPL)
    truth = lambda: O.sw_affine_hvp(s.numpy().astype(np.float64), v.numpy(), go, ge, T,
                                    prec="f64")
    e = abs_err(hv, r)
    print(f"  HVP max abs err vs FP32 reference {e:.3e} (largest entry {np.abs(r).max():.3e})",
          flush=True)
    assert close_or_nearer_truth(hv, r, truth, 1e-4 * max(1.0, float(np.abs(r).max())))


def test_c5_embedding_scored_64_pairs_vs_fp64():
    from orihime_amd.embed import soft_align_embeddings
    go, ge, T = -4.0, -0.5, 0.25
    Q, K, ln = bench.embed_inputs(torch, B=64)
    case(f"C5 embedding-scored, 64 pairs, lengths {int(ln.min())}..{int(ln.max())}")
    Qd, Kd = Q.to(DEV).requires_grad_(), K.to(DEV).requires_grad_()
    v = soft_align_embeddings(Qd, Kd, ln.to(DEV), go, ge, T)
    v.sum().backward()
    Q64, K64 = Q.double().numpy(), K.double().numpy()
    lnh = ln.numpy()
    S = [Q64[b, :lnh[b, 0]] @ K64[b, :lnh[b, 1]].T for b in range(len(lnh))]
    val, mp, *_ = O.over_pairs(O.sw_affine_forward, S, go, ge, T, prec="f64", chunk=4)
    assert rel_err(h(v), np.asarray(val), 1e-3) < 1e-5
    dQ, dK = h(Qd.grad.float()), h(Kd.grad.float())
    eq = ek = 0.0
    scale = 0.0
    for b in range(len(lnh)):
        m, n = lnh[b]
        tq, tk = mp[b] @ K64[b, :n], mp[b].T @ Q64[b, :m]
        scale = max(scale, float(np.abs(tq).max()), float(np.abs(tk).max()))
        eq = max(eq, abs_err(dQ[b, :m], tq))
        ek = max(ek, abs_err(dK[b, :n], tk))
        assert float(np.abs(dQ[b, m:]).max(initial=0.0)) == 0.0
        assert float(np.abs(dK[b, n:]).max(initial=0.0)) == 0.0
    print(f"  value rel err {rel_err(h(v), np.asarray(val), 1e-3):.3e}; dQ {eq:.3e}, dK {ek:.3e} "
          f"abs ({max(eq, ek) / scale:.2e} of the largest entry)", flush=True)
    # gradients come back in BF16 (the inputs' dtype): 2^-8 relative rounding
    assert max(eq, ek) < 4e-3 * max(1.0, scale)
