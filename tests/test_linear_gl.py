"""Linear-gap Needleman-Wunsch (``nw``) and Saigo-Vert (``sv``): oracle pins,
host path, API, and the HIP kernels.

The library runs both on the affine sweeps with one gap score in both roles
(each move sequence has exactly one state sequence, so the models are the
same function) and reports dS/dgap = dS/dgap_open + dS/dgap_ext.  The oracle
restates the reference's own linear kernels (src/nw/kernels_cpu.cpp,
src/sv_linear/kernels_cpu.cpp) independently and is pinned bit for bit
against their builds (golden fixtures), against enumeration of every path,
and against the affine oracle at gap_open = gap_ext in FP64 (which pins the
reduction itself).  Bars as the affine families: values 1e-5 of
max(|S|, 1), maps 1e-4 absolute, gap gradients 1e-4 of max(|g|, 1), dS/dT
with the low-temperature cancellation bar.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import abs_err, brute_nw_linear, brute_sv_affine, golden_files, load
from oracle import oracle as O

import orihime_amd as ohm
from orihime_amd import _lib

CASES = [(f, p) for f in ("nw", "sv") for p in golden_files(f"{f}_linear")]
IDS = [p.split("/")[-1] for _, p in CASES]


def _err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)))


def _temp_close(got, ref, truth, value, T):
    """Within 1e-4 + 8 (|S| + 1) 2^-23 / T of the FP32 reference's dS/dT, or at
    most twice as far from FP64 as that reference is."""
    tol = 1e-4 + 8.0 * (np.abs(np.asarray(value, np.float64)) + 1.0) * 2.0 ** -23 / T
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if np.all(np.abs(got - ref) <= tol * np.maximum(1.0, np.abs(ref))):
        return True
    t = np.asarray(truth(), np.float64)
    return float(np.max(np.abs(got - t))) <= max(1e-4, 2.0 * float(np.max(np.abs(ref - t))))


def _library(fam, scores, lengths, gap, T, device="cpu"):
    s = torch.as_tensor(scores, dtype=torch.float32).to(device).contiguous()
    ln = torch.as_tensor(lengths, dtype=torch.int32).to(device).contiguous()
    B = ln.shape[0]
    out = [torch.empty(B, device=device), torch.empty_like(s), torch.empty(B, device=device),
           torch.empty(B, device=device)]
    getattr(_lib, f"{fam}_forward")(s, ln, None, gap, T, value=out[0], map=out[1],
                                    grad_gap=out[2], grad_temp=out[3])
    return [t.cpu().numpy() for t in out]


def _check(fam, d, got):
    gap, T = float(d["gap"]), float(d["temp"])
    v, mp, gg, gT = got
    assert _err(v, d["value"]) < 1e-5
    assert abs_err(mp, d["map"]) < 1e-4
    assert _err(gg, d["grad_gap"]) < 1e-4
    truth = lambda: getattr(O, f"{fam}_forward")(d["scores"], gap, T, d["lengths"],  # noqa: E731
                                                 prec="f64")[3]
    assert _temp_close(gT, d["grad_temp"], truth, d["value"], T)


@pytest.mark.parametrize("fam,path", CASES, ids=IDS)
def test_oracle_matches_reference_golden_bitwise(fam, path):
    d = load(path)
    out = getattr(O, f"{fam}_forward")(d["scores"], float(d["gap"]), float(d["temp"]),
                                       d["lengths"])
    for got, key in zip(out, ("value", "map", "grad_gap", "grad_temp")):
        assert np.array_equal(got, d[key]), key


@pytest.mark.parametrize("shape,gap,T", [((2, 2), -1.0, 1.0), ((2, 3), -0.5, 0.5),
                                         ((3, 3), -2.0, 2.0), ((1, 4), -1.0, 1.0),
                                         ((3, 4), 0.3, 1.0)])
def test_oracle_value_is_the_log_sum_over_all_paths(shape, gap, T):
    rng = np.random.default_rng(sum(shape) + 7)
    s = rng.standard_normal(shape)
    nw = O.nw_forward(s[None], gap, T, prec="f64")[0][0]
    sv = O.sv_forward(s[None], gap, T, prec="f64")[0][0]
    assert abs(nw - brute_nw_linear(s, gap, T)) < 1e-10
    assert abs(sv - brute_sv_affine(s, gap, gap, T)) < 1e-10


@pytest.mark.parametrize("fam", ["nw", "sv"])
def test_linear_model_is_the_affine_model_with_one_gap_score(fam):
    rng = np.random.default_rng(3)
    s = rng.standard_normal((3, 23, 19))
    ln = np.array([[23, 19], [9, 14], [0, 5]], np.int32)
    lin = getattr(O, f"{fam}_forward")(s, -0.9, 0.6, ln, prec="f64")
    aff = getattr(O, f"{fam}_affine_forward")(s, -0.9, -0.9, 0.6, ln, prec="f64")
    assert np.allclose(lin[0], aff[0], rtol=0, atol=1e-12)
    assert np.allclose(lin[1], aff[1], rtol=0, atol=1e-12)
    assert np.allclose(lin[2], aff[2] + aff[3], rtol=0, atol=1e-11)
    assert np.allclose(lin[3], aff[4], rtol=0, atol=1e-11)


@pytest.mark.parametrize("fam,path", CASES, ids=IDS)
def test_host_path_matches_golden(fam, path):
    d = load(path)
    _check(fam, d, _library(fam, d["scores"], d["lengths"], float(d["gap"]), float(d["temp"])))


@pytest.mark.parametrize("fam", ["nw", "sv"])
def test_public_functions_and_gradients(fam):
    d = load([p for f, p in CASES if f == fam and "ragged" in p][0])
    gap, T = float(d["gap"]), float(d["temp"])
    x = torch.from_numpy(d["scores"]).clone().requires_grad_()
    ln = torch.from_numpy(d["lengths"])
    g = torch.tensor(gap, requires_grad=True)
    mp = getattr(ohm, fam)(x.detach(), gap_score=gap, temperature=T, lengths=ln)
    assert abs_err(mp.numpy(), d["map"]) < 1e-4
    v = getattr(ohm, f"{fam}_value")(x, gap_score=g, temperature=T, lengths=ln)
    assert _err(v.detach().numpy(), d["value"]) < 1e-5
    v.sum().backward()
    assert abs_err(x.grad.numpy(), d["map"]) < 1e-4
    assert abs(g.grad.item() - float(d["grad_gap"].sum())) < 1e-3 * max(1.0, abs(g.grad.item()))
    out = getattr(ohm.ops, fam).forward(x.detach(), gap_score=gap, temperature=T, lengths=ln)
    assert set(out) == {"map", "value", "entropy"}
    # the map is differentiable too (forward-over-reverse pass with both gap
    # directions set): its VJP in the gap equals <G, dP/dgap> by central
    # differences of the oracle
    x2 = torch.from_numpy(d["scores"]).clone().requires_grad_()
    g2 = torch.tensor(gap, requires_grad=True)
    G = torch.from_numpy(np.random.default_rng(1).standard_normal(d["scores"].shape).astype(np.float32))
    (getattr(ohm, fam)(x2, gap_score=g2, temperature=T, lengths=ln) * G).sum().backward()
    fw = getattr(O, f"{fam}_forward")
    h = 1e-5
    fd = ((fw(d["scores"].astype(np.float64), gap + h, T, d["lengths"], prec="f64")[1] -
           fw(d["scores"].astype(np.float64), gap - h, T, d["lengths"], prec="f64")[1]) /
          (2 * h) * G.numpy()).sum()
    assert abs(g2.grad.item() - fd) < 1e-3 * max(1.0, abs(fd))


@pytest.mark.gpu
@pytest.mark.parametrize("fam,path", CASES, ids=IDS)
def test_gpu_matches_golden(fam, path):
    d = load(path)
    print(f"linear {fam} gpu golden {path.split('/')[-1]}", flush=True)
    _check(fam, d, _library(fam, d["scores"], d["lengths"], float(d["gap"]), float(d["temp"]),
                            device="cuda"))


@pytest.mark.gpu
@pytest.mark.parametrize("fam", ["nw", "sv"])
def test_gpu_linear_equals_affine_call_with_one_gap(fam):
    rng = np.random.default_rng(12)
    s = rng.standard_normal((3, 200, 150)).astype(np.float32)
    ln = np.array([[200, 150], [77, 140], [130, 3]], np.int32)
    print(f"linear {fam} gpu vs affine call 3x200x150", flush=True)
    lin = _library(fam, s, ln, -1.2, 0.5, device="cuda")
    st = torch.from_numpy(s).cuda()
    lnd = torch.from_numpy(ln).cuda()
    out = [torch.empty(3, device="cuda"), torch.empty_like(st)] + \
        [torch.empty(3, device="cuda") for _ in range(3)]
    getattr(_lib, f"{fam}_affine_forward")(st, lnd, None, -1.2, -1.2, 0.5, value=out[0],
                                           map=out[1], grad_open=out[2], grad_ext=out[3],
                                           grad_temp=out[4])
    aff = [t.cpu().numpy() for t in out]
    assert np.array_equal(lin[0], aff[0]) and np.array_equal(lin[1], aff[1])
    assert np.array_equal(lin[2], aff[2] + aff[3]) and np.array_equal(lin[3], aff[4])
    ref = getattr(O, f"{fam}_forward")(s, -1.2, 0.5, ln, prec="f64")
    assert _err(lin[0], ref[0]) < 1e-5 and abs_err(lin[1], ref[1]) < 1e-4
    assert _err(lin[2], ref[2]) < 1e-4


def test_nn_modules_and_vmap_cover_the_nw_and_sv_families():
    from orihime_amd import nn as onn

    rng = np.random.default_rng(6)
    x = torch.from_numpy(rng.standard_normal((2, 6, 7)).astype(np.float32))
    for cls, fam, kw in ((onn.NW, "nw", {"gap_score": -0.5}),
                         (onn.SVLinear, "sv", {"gap_score": -0.5}),
                         (onn.NWAffine, "nw_affine", {"gap_open_score": -1.0, "gap_extend_score": -0.3}),
                         (onn.SVAffine, "sv_affine", {"gap_open_score": -1.0, "gap_extend_score": -0.3})):
        m = cls(**kw, temperature=0.7, learnable=("temperature",))
        out = m(x)
        ref = getattr(ohm, fam)(x, **kw, temperature=0.7)
        assert torch.allclose(out, ref)
        out.sum().backward()
        assert m.temperature.grad is not None and torch.isfinite(m.temperature.grad).all()
    v = torch.from_numpy(rng.standard_normal((3, 2, 6, 7)).astype(np.float32))
    batched = torch.vmap(lambda vv: torch.ops.orihime.nw_affine_marginals_hvp(
        x, vv, -1.0, -0.3, 0.7, None))(v)
    for k in range(3):
        one = torch.ops.orihime.nw_affine_marginals_hvp(x, v[k], -1.0, -0.3, 0.7, None)
        assert torch.allclose(batched[k], one, atol=1e-6)
    assert ohm.sv_linear is ohm.sv


def test_raw_tier_vjps_for_the_nw_and_sv_families():
    rng = np.random.default_rng(9)
    x = torch.from_numpy(rng.standard_normal((2, 5, 6)).astype(np.float32))
    G = torch.from_numpy(rng.standard_normal((2, 5, 6)).astype(np.float32))
    x64 = x.numpy().astype(np.float64)
    h = 1e-5
    for fam in ("nw_affine", "sv_affine"):
        out = getattr(ohm.raw, fam).vjp(x, wrt=("gap_open_score", "temperature"), cotangent=G,
                                        gap_open_score=-1.0, gap_extend_score=-0.3,
                                        temperature=0.7)
        f = getattr(O, f"{fam}_forward")
        fd_go = ((f(x64, -1.0 + h, -0.3, 0.7, prec="f64")[1] -
                  f(x64, -1.0 - h, -0.3, 0.7, prec="f64")[1]) / (2 * h) * G.numpy()).sum()
        fd_t = ((f(x64, -1.0, -0.3, 0.7 + h, prec="f64")[1] -
                 f(x64, -1.0, -0.3, 0.7 - h, prec="f64")[1]) / (2 * h) * G.numpy()).sum()
        assert abs(out["gap_open_score"].item() - fd_go) < 1e-4 * max(1.0, abs(fd_go))
        assert abs(out["temperature"].item() - fd_t) < 1e-4 * max(1.0, abs(fd_t))
    for fam, f in (("nw", O.nw_forward), ("sv_linear", O.sv_forward)):
        got = getattr(ohm.raw, fam).vjp_one(x, wrt="gap_score", cotangent=G, gap_score=-0.5,
                                            temperature=0.7)
        fd = ((f(x64, -0.5 + h, 0.7, prec="f64")[1] - f(x64, -0.5 - h, 0.7, prec="f64")[1]) /
              (2 * h) * G.numpy()).sum()
        assert abs(got.item() - fd) < 1e-4 * max(1.0, abs(fd))


def test_masks_compile_and_autocast_for_the_nw_and_sv_families():
    rng = np.random.default_rng(13)
    x = torch.from_numpy(rng.standard_normal((2, 6, 7)).astype(np.float32))
    m = torch.zeros_like(x, dtype=torch.bool)
    m[:, 2, 3] = True
    kw_aff = dict(gap_open_score=-1.0, gap_extend_score=-0.3, temperature=0.7)
    for f, kw in ((ohm.nw_affine, kw_aff), (ohm.sv_affine, kw_aff),
                  (ohm.nw, dict(gap_score=-0.5, temperature=0.7)),
                  (ohm.sv, dict(gap_score=-0.5, temperature=0.7))):
        out = f(x, mask=m, **kw)
        assert out[m].abs().max() == 0
        compiled = torch.compile(lambda s, f=f, kw=kw: f(s, **kw), backend="aot_eager")
        assert torch.allclose(compiled(x), f(x, **kw))
    with torch.autocast("cpu", dtype=torch.bfloat16):
        assert ohm.nw(x.bfloat16(), gap_score=-0.5).dtype == torch.float32
