"""Needleman-Wunsch affine (global alignment): oracle pins, host path, API,
and the HIP kernels against the oracle.

The oracle restatement (oracle/sw_oracle.c, nwa_*) is pinned bit for bit
against the reference's own CPU kernels (the golden fixtures were written by
src/nw_affine/kernels_cpu.cpp built into oracle/_ref), against brute-force
enumeration of every global alignment, and against finite differences of its
value.  The library (host and GPU) is then checked against the oracle:
values to 1e-5 relative (of max(|S|, 1): global values cross zero), maps to
1e-4 absolute, parameter gradients (sums over the whole path) to 1e-4
relative of max(|g|, 1).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import abs_err, brute_nw_affine, golden_files, load
from oracle import oracle as O

import orihime_amd as ohm
from orihime_amd import _lib

NW = golden_files("nw_affine")
IDS = [p.split("/")[-1] for p in NW]


def _params(d):
    return float(d["gap_open"]), float(d["gap_ext"]), float(d["temp"])


def _grad_err(got, ref):
    """Error relative to max(|ref|, 1): values of global alignments cross
    zero, parameter gradients are path-length sums."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)))


def _temp_ok(got, ref, value, T):
    """dS/dT = (S - E[path score]) / T cancels two values of size |S|: its
    FP32 noise grows with |S| / T (the FP32 reference carries the same: 1e-4
    of max(|g|, 1) at T = 0.1 on 64 x 64 pairs), hence the bar
    1e-4 + 8 (|S| + 1) 2^-23 / T relative to max(|g|, 1)."""
    tol = 1e-4 + 8.0 * (np.abs(np.asarray(value, np.float64)) + 1.0) * 2.0 ** -23 / T
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return bool(np.all(np.abs(got - ref) <= tol * np.maximum(1.0, np.abs(ref))))


def _temp_close(got, ref, truth, value, T):
    """Within the cancellation bar of the FP32 reference's dS/dT, or at most
    twice as far from the FP64 result (truth(), computed when needed) as the
    FP32 reference is itself (it reads 4e-4 off FP64 at T = 0.1 on 16 x 16
    pairs)."""
    if _temp_ok(got, ref, value, T):
        return True
    t = np.asarray(truth(), np.float64)
    got = np.asarray(got, np.float64)
    bar = np.maximum(1e-4 * np.maximum(1.0, np.abs(t)), 2.0 * np.abs(np.asarray(ref, np.float64) - t))
    return bool(np.all(np.abs(got - t) <= bar))


def _temp_vs_fp64(got, f32, f64, value, T):
    """dS/dT against the FP64 oracle: within the cancellation bar, or the
    batch's worst error at most twice the FP32 reference's worst error on the
    same inputs (its seeds 2^(M - S) carry ulp(|S| / T) of rounding, so at
    T = 0.1 it reads up to 5e-3 of max(|g|, 1) off FP64)."""
    if _temp_ok(got, f64, value, T):
        return True
    got, f32, f64 = (np.asarray(x, np.float64) for x in (got, f32, f64))
    return float(np.max(np.abs(got - f64))) <= 2.0 * float(np.max(np.abs(f32 - f64)))


def _library(scores, lengths, go, ge, T, device="cpu", offsets=None, L1max=None, L2max=None):
    s = torch.as_tensor(scores, dtype=torch.float32).to(device).contiguous()
    ln = torch.as_tensor(lengths, dtype=torch.int32).to(device).contiguous()
    B = ln.shape[0]
    value = torch.empty(B, dtype=torch.float32, device=device)
    mp = torch.empty_like(s)
    g = [torch.empty(B, dtype=torch.float32, device=device) for _ in range(3)]
    off = None if offsets is None else torch.as_tensor(offsets, dtype=torch.int64).to(device)
    _lib.nw_affine_forward(s, ln, off, go, ge, T, value=value, map=mp, grad_open=g[0],
                           grad_ext=g[1], grad_temp=g[2], L1max=L1max, L2max=L2max)
    return [t.cpu().numpy() for t in (value, mp, *g)]


# --------------------------------------------------------------------------
# Oracle pins
# --------------------------------------------------------------------------


@pytest.mark.parametrize("path", NW, ids=IDS)
def test_oracle_matches_reference_golden_bitwise(path):
    d = load(path)
    go, ge, T = _params(d)
    out = O.nw_affine_forward(d["scores"], go, ge, T, d["lengths"])
    for got, key in zip(out, ("value", "map", "grad_open", "grad_ext", "grad_temp")):
        assert np.array_equal(got, d[key]), key


@pytest.mark.skipif(not O.nw_ref_available(), reason="reference kernels not built here")
def test_oracle_matches_reference_build_on_fresh_inputs():
    rng = np.random.default_rng(5)
    for B, L1, L2, go, ge, T in [(3, 9, 13, -1.5, -0.4, 0.3), (2, 70, 65, -4.0, -1.0, 1.5)]:
        s = (rng.standard_normal((B, L1, L2)) * 1.5).astype(np.float32)
        ln = np.array([[L1, L2], [L1 - 5, L2 - 2], [1, L2]][:B], np.int32)
        a = O.nw_affine_forward(s, go, ge, T, ln)
        r = O.nw_affine_forward(s, go, ge, T, ln, impl="ref")
        for x, y in zip(a, r):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("shape,go,ge,T", [((2, 2), -2.0, -0.5, 1.0), ((2, 3), -1.0, -0.3, 0.5),
                                           ((3, 3), -3.0, -1.0, 2.0), ((1, 4), -2.0, -0.5, 1.0),
                                           ((4, 1), -2.0, -0.5, 0.7), ((3, 4), 0.5, 0.2, 1.0)])
def test_oracle_value_is_the_log_sum_over_all_global_alignments(shape, go, ge, T):
    rng = np.random.default_rng(sum(shape))
    s = rng.standard_normal(shape)
    v = O.nw_affine_forward(s[None], go, ge, T, prec="f64")[0][0]
    assert abs(v - brute_nw_affine(s, go, ge, T)) < 1e-10


def test_oracle_empty_sides_are_one_gap_run():
    s = np.zeros((3, 4, 5), np.float64)
    ln = np.array([[0, 5], [4, 0], [0, 0]], np.int32)
    v, mp, gop, gex, gT = O.nw_affine_forward(s, -2.0, -0.5, 1.0, ln, prec="f64")
    assert np.allclose(v, [-2.0 - 4 * 0.5, -2.0 - 3 * 0.5, 0.0])
    assert np.allclose(gop, [1, 1, 0]) and np.allclose(gex, [4, 3, 0]) and np.allclose(gT, 0)
    assert not mp.any()


def test_oracle_derivatives_match_finite_differences():
    rng = np.random.default_rng(11)
    s = rng.standard_normal((1, 6, 8))
    go, ge, T = -1.5, -0.4, 0.8
    v, mp, gop, gex, gT = O.nw_affine_forward(s, go, ge, T, prec="f64")
    h = 1e-6

    def val(ss=s, a=go, b=ge, t=T):
        return O.nw_affine_forward(ss, a, b, t, prec="f64")[0][0]

    fd = np.zeros_like(s)
    for i in range(6):
        for j in range(8):
            sp, sm = s.copy(), s.copy()
            sp[0, i, j] += h
            sm[0, i, j] -= h
            fd[0, i, j] = (val(sp) - val(sm)) / (2 * h)
    assert abs_err(mp, fd) < 1e-7
    assert abs(gop[0] - (val(a=go + h) - val(a=go - h)) / (2 * h)) < 1e-7
    assert abs(gex[0] - (val(b=ge + h) - val(b=ge - h)) / (2 * h)) < 1e-7
    assert abs(gT[0] - (val(t=T + h) - val(t=T - h)) / (2 * h)) < 1e-7


# --------------------------------------------------------------------------
# Library on the host (CPU tensors through the C ABI)
# --------------------------------------------------------------------------


def test_library_exports_the_nw_entry_points():
    for name in ("ohm_nw_affine_workspace", "ohm_nw_affine_forward", "ohm_nw_affine_forward_cpu"):
        assert hasattr(_lib._lib, name)
    assert _lib._ws_nw_affine(1, 4, 100, 100, -1) > _lib._ws_nw_affine(0, 4, 100, 100, -1) > 0
    assert _lib._ws_nw_affine(2, 4, 100, 100, -1) > _lib._ws_nw_affine(1, 4, 100, 100, -1)
    assert _lib._ws_nw_affine(3, 4, 100, 100, -1) == 0


@pytest.mark.parametrize("path", NW, ids=IDS)
def test_host_path_matches_golden(path):
    d = load(path)
    go, ge, T = _params(d)
    v, mp, gop, gex, gT = _library(d["scores"], d["lengths"], go, ge, T)
    assert _grad_err(v, d["value"]) < 1e-5
    assert abs_err(mp, d["map"]) < 1e-4
    for got, key in ((gop, "grad_open"), (gex, "grad_ext")):
        assert _grad_err(got, d[key]) < 1e-4, key
    truth = lambda: O.nw_affine_forward(d["scores"], go, ge, T, d["lengths"], prec="f64")[4]  # noqa: E731
    assert _temp_close(gT, d["grad_temp"], truth, d["value"], T)


def test_public_functions_follow_the_oracle():
    d = load([p for p in NW if "ragged" in p][0])
    go, ge, T = _params(d)
    x = torch.from_numpy(d["scores"])
    ln = torch.from_numpy(d["lengths"])
    kw = dict(gap_open_score=go, gap_extend_score=ge, temperature=T, lengths=ln)
    assert abs_err(ohm.nw_affine(x, **kw).numpy(), d["map"]) < 1e-4
    assert _grad_err(ohm.nw_affine_value(x, **kw).numpy(), d["value"]) < 1e-5
    assert _temp_ok(ohm.nw_affine_entropy(x, **kw).numpy(), d["grad_temp"], d["value"], T)
    out = ohm.ops.nw_affine.forward(x, **kw)
    assert set(out) == {"map", "value", "entropy"}
    assert abs_err(out["map"].numpy(), d["map"]) < 1e-4
    assert _grad_err(out["value"].numpy(), d["value"]) < 1e-5


def test_value_gradients_reach_scores_and_tensor_parameters():
    d = load([p for p in NW if "b4_8x10" in p][0])
    go, ge, T = _params(d)
    x = torch.from_numpy(d["scores"]).clone().requires_grad_()
    tgo = torch.tensor(go, requires_grad=True)
    tge = torch.tensor(ge, requires_grad=True)
    tT = torch.tensor(T, requires_grad=True)
    v = ohm.nw_affine_value(x, gap_open_score=tgo, gap_extend_score=tge, temperature=tT)
    (v * torch.arange(1.0, 5.0)).sum().backward()
    w = np.arange(1.0, 5.0)
    assert abs_err(x.grad.numpy(), d["map"] * w[:, None, None]) < 1e-4
    assert abs(tgo.grad.item() - (w * d["grad_open"]).sum()) < 1e-3 * max(1, abs(tgo.grad.item()))
    assert abs(tge.grad.item() - (w * d["grad_ext"]).sum()) < 1e-3 * max(1, abs(tge.grad.item()))
    assert abs(tT.grad.item() - (w * d["grad_temp"]).sum()) < 1e-3 * max(1, abs(tT.grad.item()))


def test_prefix_masks_equal_lengths():
    d = load([p for p in NW if "ragged" in p][0])
    go, ge, T = _params(d)
    x = torch.from_numpy(d["scores"])
    ln = torch.from_numpy(d["lengths"])
    B, L1, L2 = x.shape
    mask1 = torch.arange(L1)[None, :] < ln[:, 0:1]
    mask2 = torch.arange(L2)[None, :] < ln[:, 1:2]
    v = ohm.nw_affine_value(x, gap_open_score=go, gap_extend_score=ge, temperature=T, lengths=ln)
    from orihime_amd.families import engine; _nw_affine_operator = engine("nw_affine")
    vm = _nw_affine_operator.value(x, go, ge, T, mask1=mask1, mask2=mask2)
    assert torch.equal(v, vm)
    vg = ohm.ops._kernels["nw_affine"].value_grad_params(x, go, ge, T, ln)
    assert torch.equal(vg[0], v)


def test_invalid_lengths_raise():
    x = torch.randn(1, 4, 5)
    with pytest.raises((RuntimeError, ValueError)):
        ohm.nw_affine(x, gap_open_score=-2.0, gap_extend_score=-0.5,
                      lengths=torch.tensor([[4, 6]], dtype=torch.int32))
    with pytest.raises((RuntimeError, ValueError)):
        ohm.nw_affine(x, gap_open_score=-2.0, gap_extend_score=-0.5,
                      lengths=torch.tensor([[-1, 5]], dtype=torch.int32))


def _tangent(s, ln, go, ge, T, device="cpu", direction=None, d=(0.0, 0.0, 0.0), contract=False):
    st = torch.as_tensor(s, dtype=torch.float32).to(device).contiguous()
    lnd = torch.as_tensor(ln, dtype=torch.int32).to(device).contiguous()
    mt = torch.empty_like(st)
    c = torch.empty(st.shape[0], 3, device=device) if contract else None
    dv = None if direction is None else torch.as_tensor(direction, dtype=torch.float32).to(device)
    _lib.nw_affine_tangent(st, lnd, None, go, ge, T, direction=dv, d_gap_open=d[0],
                           d_gap_ext=d[1], d_temp=d[2], map_tangent=mt, contract=c)
    return mt.cpu().numpy(), (None if c is None else c.cpu().numpy())


def _fd_second_order(s, ln, go, ge, T, V, G, h=1e-5):
    """Central differences of the FP64 oracle's map: H.V, dP/dgo, dP/dge,
    dP/dT, and the parameter parts of the map VJP against G."""
    def mp(ss=s, a=go, b=ge, t=T):
        return O.nw_affine_forward(ss, a, b, t, ln, prec="f64")[1]
    hv = (mp(s + h * V) - mp(s - h * V)) / (2 * h)
    sens = [(mp(a=go + h) - mp(a=go - h)) / (2 * h), (mp(b=ge + h) - mp(b=ge - h)) / (2 * h),
            (mp(t=T + h) - mp(t=T - h)) / (2 * h)]
    con = np.stack([(x * G).sum(axis=(1, 2)) for x in sens], axis=1)
    return hv, sens, con


SECOND = [((3, 7, 9), [[7, 9], [4, 6], [1, 5]], -1.3, -0.4, 0.7),
          ((2, 5, 4), [[5, 4], [0, 3]], 0.5, 0.2, 1.0)]


@pytest.mark.parametrize("shape,ln,go,ge,T", SECOND)
def test_host_second_order_matches_finite_differences(shape, ln, go, ge, T):
    rng = np.random.default_rng(shape[1])
    s, V, G = (rng.standard_normal(shape) for _ in range(3))
    ln = np.array(ln, np.int32)
    hv, sens, con = _fd_second_order(s.astype(np.float32).astype(np.float64), ln, go, ge, T, V, G)
    s32 = s.astype(np.float32)
    assert abs_err(_tangent(s32, ln, go, ge, T, direction=V)[0], hv) < 1e-5
    for p in range(3):
        d = [0.0, 0.0, 0.0]
        d[p] = 1.0
        assert abs_err(_tangent(s32, ln, go, ge, T, d=d)[0], sens[p]) < 1e-5
    mt, c = _tangent(s32, ln, go, ge, T, direction=G, contract=True)
    assert _grad_err(c, con) < 1e-4


def test_map_derivatives_flow_through_autograd():
    rng = np.random.default_rng(2)
    s = rng.standard_normal((2, 6, 7)).astype(np.float32)
    G = rng.standard_normal((2, 6, 7)).astype(np.float32)
    x = torch.from_numpy(s).requires_grad_()
    go = torch.tensor(-1.5, requires_grad=True)
    mp = ohm.nw_affine(x, gap_open_score=go, gap_extend_score=-0.4, temperature=0.8)
    (mp * torch.from_numpy(G)).sum().backward()
    hv, sens, con = _fd_second_order(s.astype(np.float64), np.array([[6, 7]] * 2, np.int32),
                                     -1.5, -0.4, 0.8, G, G)
    assert abs_err(x.grad.numpy(), hv) < 1e-5
    assert abs(go.grad.item() - con[:, 0].sum()) < 1e-4 * max(1.0, abs(go.grad.item()))
    x2 = torch.from_numpy(s).requires_grad_()
    ohm.nw(x2, gap_score=-0.7, temperature=0.8).sum().backward()
    assert torch.isfinite(x2.grad).all()


# --------------------------------------------------------------------------
# GPU kernels against the oracle
# --------------------------------------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("path", NW, ids=IDS)
def test_gpu_matches_golden(path):
    d = load(path)
    go, ge, T = _params(d)
    print(f"nw gpu golden {path.split('/')[-1]}", flush=True)
    v, mp, gop, gex, gT = _library(d["scores"], d["lengths"], go, ge, T, device="cuda")
    assert _grad_err(v, d["value"]) < 1e-5
    assert abs_err(mp, d["map"]) < 1e-4
    for got, key in ((gop, "grad_open"), (gex, "grad_ext")):
        assert _grad_err(got, d[key]) < 1e-4, key
    truth = lambda: O.nw_affine_forward(d["scores"], go, ge, T, d["lengths"], prec="f64")[4]  # noqa: E731
    assert _temp_close(gT, d["grad_temp"], truth, d["value"], T)


@pytest.mark.gpu
@pytest.mark.parametrize("B,L1,L2,go,ge,T", [(3, 300, 257, -2.0, -0.5, 0.5),
                                             (2, 130, 400, -4.0, -1.0, 1.0),
                                             (16, 64, 64, -1.0, -0.3, 0.1),
                                             (4, 65, 3, -2.0, -0.5, 2.0)])
def test_gpu_matches_fp64_oracle_multi_strip(B, L1, L2, go, ge, T):
    rng = np.random.default_rng(B * 1000 + L1)
    s = rng.standard_normal((B, L1, L2)).astype(np.float32)
    ln = np.tile(np.array([L1, L2], np.int32), (B, 1))
    ln[0] = [max(L1 - 37, 1), max(L2 - 11, 1)]
    print(f"nw gpu vs oracle B={B} {L1}x{L2}", flush=True)
    v, mp, gop, gex, gT = _library(s, ln, go, ge, T, device="cuda")
    r = O.nw_affine_forward(s, go, ge, T, ln, prec="f64")
    assert _grad_err(v, r[0]) < 1e-5
    assert abs_err(mp, r[1]) < 1e-4
    for got, ref in zip((gop, gex), r[2:4]):
        assert _grad_err(got, ref) < 1e-4
    r32 = O.nw_affine_forward(s, go, ge, T, ln)
    assert _temp_vs_fp64(gT, r32[4], r[4], r[0], T)


@pytest.mark.gpu
def test_gpu_packed_layout_equals_padded():
    rng = np.random.default_rng(3)
    ln = np.array([[130, 70], [5, 200], [64, 64], [0, 9], [77, 1]], np.int32)
    L1m, L2m = int(ln[:, 0].max()), int(ln[:, 1].max())
    s = rng.standard_normal((len(ln), L1m, L2m)).astype(np.float32)
    packed = np.concatenate([s[b, :ln[b, 0], :ln[b, 1]].ravel() for b in range(len(ln))])
    sizes = ln[:, 0].astype(np.int64) * ln[:, 1]
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    print("nw gpu packed vs padded", flush=True)
    a = _library(s, ln, -2.0, -0.5, 0.7, device="cuda")
    b = _library(packed, ln, -2.0, -0.5, 0.7, device="cuda", offsets=offsets, L1max=L1m,
                 L2max=L2m)
    for x, y in zip((a[0], *a[2:]), (b[0], *b[2:])):
        assert np.array_equal(x, y)
    for k in range(len(ln)):
        m = b[1][offsets[k]:offsets[k] + sizes[k]].reshape(ln[k, 0], ln[k, 1])
        assert np.array_equal(m, a[1][k, :ln[k, 0], :ln[k, 1]])


@pytest.mark.gpu
def test_gpu_long_pairs_match_the_host_path():
    rng = np.random.default_rng(9)
    s = rng.standard_normal((2, 1100, 1000)).astype(np.float32)
    ln = np.array([[1100, 1000], [700, 990]], np.int32)
    print("nw gpu vs host 2x1100x1000", flush=True)
    g = _library(s, ln, -2.0, -0.5, 0.5, device="cuda")
    h = _library(s, ln, -2.0, -0.5, 0.5)
    assert _grad_err(g[0], h[0]) < 1e-5
    assert abs_err(g[1], h[1]) < 1e-4
    for x, y in zip(g[2:4], h[2:4]):
        assert _grad_err(x, y) < 1e-4
    assert _temp_ok(g[4], h[4], h[0], 0.5)


@pytest.mark.gpu
def test_gpu_public_functions_run_on_the_hip_kernels():
    d = load([p for p in NW if "ragged" in p][0])
    go, ge, T = _params(d)
    x = torch.from_numpy(d["scores"]).cuda().requires_grad_()
    ln = torch.from_numpy(d["lengths"]).cuda()
    v = ohm.nw_affine_value(x, gap_open_score=go, gap_extend_score=ge, temperature=T, lengths=ln)
    v.sum().backward()
    assert abs_err(x.grad.cpu().numpy(), d["map"]) < 1e-4
    assert _grad_err(v.detach().cpu().numpy(), d["value"]) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("B,L1,L2,go,ge,T", [(3, 300, 257, -2.0, -0.5, 0.5),
                                             (5, 70, 130, -1.0, -0.3, 1.0)])
def test_gpu_second_order_on_multi_strip_pairs(B, L1, L2, go, ge, T):
    """Against central differences of the FP64 oracle.  Second-order outputs
    carry FP32 rounding that accumulates along the reverse recursion (largest
    at the cells it reaches last, near (1, 1)); on the first case the FP32
    reference's own HVP reads 1.3e-4 off FP64, the host twin 0.9e-4, this path
    1.5e-4 (3.9e-4 on the ragged pair: 0.9e-4 of the largest entry), hence the
    bar of 2e-4 of the largest entry."""
    rng = np.random.default_rng(L1 + L2)
    s, V, G = (rng.standard_normal((B, L1, L2)).astype(np.float32) for _ in range(3))
    ln = np.tile(np.array([L1, L2], np.int32), (B, 1))
    ln[0] = [L1 - 41, L2 - 3]
    ln[1] = [0, L2 - 7]
    print(f"nw gpu tangent B={B} {L1}x{L2}", flush=True)
    hv, sens, con = _fd_second_order(s.astype(np.float64), ln, go, ge, T, V, G, h=1e-4)
    for kw, truth in ((dict(direction=V), hv), (dict(d=(1.0, 0.0, 0.0)), sens[0]),
                      (dict(d=(0.0, 1.0, 0.0)), sens[1]), (dict(d=(0.0, 0.0, 1.0)), sens[2])):
        g = _tangent(s, ln, go, ge, T, device="cuda", **kw)[0]
        assert abs_err(g, truth) < 2e-4 * max(1.0, float(np.abs(truth).max(initial=0.0)))
    gm, gc = _tangent(s, ln, go, ge, T, device="cuda", direction=G, contract=True)
    assert _grad_err(gc, con) < 2e-3


@pytest.mark.gpu
@pytest.mark.parametrize("shape,ln,go,ge,T", SECOND)
def test_gpu_second_order_matches_finite_differences(shape, ln, go, ge, T):
    rng = np.random.default_rng(shape[1])
    s, V, G = (rng.standard_normal(shape) for _ in range(3))
    ln = np.array(ln, np.int32)
    s32 = s.astype(np.float32)
    hv, sens, con = _fd_second_order(s32.astype(np.float64), ln, go, ge, T, V, G)
    print(f"nw gpu tangent vs fd {shape}", flush=True)
    assert abs_err(_tangent(s32, ln, go, ge, T, device="cuda", direction=V)[0], hv) < 1e-5
    for p in range(3):
        d = [0.0, 0.0, 0.0]
        d[p] = 1.0
        assert abs_err(_tangent(s32, ln, go, ge, T, device="cuda", d=d)[0], sens[p]) < 1e-5
    _, c = _tangent(s32, ln, go, ge, T, device="cuda", direction=G, contract=True)
    assert _grad_err(c, con) < 1e-4


@pytest.mark.gpu
def test_gpu_nw_map_backward_runs_on_the_hip_kernels():
    rng = np.random.default_rng(8)
    s = rng.standard_normal((4, 90, 80)).astype(np.float32)
    G = rng.standard_normal((4, 90, 80)).astype(np.float32)
    x = torch.from_numpy(s).cuda().requires_grad_()
    mp = ohm.nw_affine(x, gap_open_score=-1.0, gap_extend_score=-0.3, temperature=0.6)
    (mp * torch.from_numpy(G).cuda()).sum().backward()
    h = _tangent(s, np.array([[90, 80]] * 4, np.int32), -1.0, -0.3, 0.6, direction=G)[0]
    assert abs_err(x.grad.cpu().numpy(), h) < 1e-4


SECOND_GOLDEN = [p for p in NW if "hvp" in load(p)]


def _second_order_close(got, ref, truth, tol):
    """Within tol of the reference's FP32 output, or at most twice as far
    from the FP64 central-difference value as that output is."""
    if abs_err(got, ref) < tol:
        return True
    t = truth()
    return abs_err(got, t) <= max(tol, 2.0 * abs_err(ref, t))


def _check_second_order_golden(d, device):
    go, ge, T = _params(d)
    s, ln, V = d["scores"], d["lengths"], d["V"]
    f64 = lambda: _fd_second_order(s.astype(np.float64), ln, go, ge, T, V, V)  # noqa: E731
    tol = lambda r: 1e-4 * max(1.0, float(np.abs(r).max(initial=0.0)))  # noqa: E731
    got = _tangent(s, ln, go, ge, T, device=device, direction=V)[0]
    assert _second_order_close(got, d["hvp"], lambda: f64()[0], tol(d["hvp"]))
    for p, n in enumerate(("gap_open", "gap_ext", "temp")):
        dd = [0.0, 0.0, 0.0]
        dd[p] = 1.0
        got = _tangent(s, ln, go, ge, T, device=device, d=dd)[0]
        ref = d[f"sens_{n}"]
        assert _second_order_close(got, ref, lambda p=p: f64()[1][p], tol(ref)), n


@pytest.mark.parametrize("path", SECOND_GOLDEN, ids=lambda p: p.split("/")[-1])
def test_host_second_order_matches_reference_golden(path):
    _check_second_order_golden(load(path), "cpu")


@pytest.mark.gpu
@pytest.mark.parametrize("path", SECOND_GOLDEN, ids=lambda p: p.split("/")[-1])
def test_gpu_second_order_matches_reference_golden(path):
    print(f"nw gpu second order golden {path.split('/')[-1]}", flush=True)
    _check_second_order_golden(load(path), "cuda")
