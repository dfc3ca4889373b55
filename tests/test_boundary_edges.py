"""Boundary cases of the drop-in schemas that the reference's bindings pin:

* batches whose score tensor has an empty dimension, (B, 0, N) and (B, M, 0),
  called without ``lengths`` -- the reference returns the empty alignment
  (0; global alignment: the gap run over the other side) for every family;
  expected outputs are the golden fixtures written by the reference's own CPU
  kernels (tests/golden/*_empty_*.npz);
* the reference's pre-allocation shape guards and their messages
  (tests/test_soft_sw_affine.py:781-788 and :899-905 of the reference suite),
  reached with zero-storage tensors;
* (GPU) a packed-layout call with a workspace sized for fewer cells than its
  lengths ask for: the pairs that do not fit come back as NaN instead of
  passing for empty alignments.
"""

from __future__ import annotations

import os

import numpy as np
import pytest
import torch

import orihime_amd  # noqa: F401  (registers torch.ops.orihime)
from orihime_amd import _lib
from helpers import GOLDEN, load

EMPTY = sorted(f for f in os.listdir(GOLDEN) if "_empty_" in f)
OPS = {
    "affine": ("soft_sw_affine_with_grads", ("grad_open", "grad_ext", "grad_temp")),
    "linear": ("soft_sw_with_grads", ("grad_gap", "grad_temp")),
    "nw_affine": ("soft_nw_affine_with_grads", ("grad_open", "grad_ext", "grad_temp")),
    "sv_affine": ("soft_sv_affine_with_grads", ("grad_open", "grad_ext", "grad_temp")),
    "nw_linear": ("soft_nw_with_grads", ("grad_gap", "grad_temp")),
    "sv_linear": ("soft_sv_linear_with_grads", ("grad_gap", "grad_temp")),
}


def _params(d):
    if "gap" in d:
        return float(d["gap"]), float(d["temp"])
    return float(d["gap_open"]), float(d["gap_ext"]), float(d["temp"])


def _check_empty(name, device):
    d = load(os.path.join(GOLDEN, name))
    kind = str(d["kind"])
    op, grads = OPS[kind]
    s = torch.from_numpy(d["scores"]).to(device)
    assert s.numel() == 0 and s.shape[0] == 2
    out = getattr(torch.ops.orihime, op)(s, *_params(d), None)
    if device != "cpu":
        torch.cuda.synchronize()
    v, m, gr = out[0].cpu().numpy(), out[1], [x.cpu().numpy() for x in out[2:]]
    assert tuple(m.shape) == d["scores"].shape
    np.testing.assert_array_equal(v, d["value"])
    for got, key in zip(gr, grads):
        np.testing.assert_array_equal(got, d[key], err_msg=key)
    if "hvp" in d:  # second order through the same shape-empty input
        ns = {"affine": "sw_affine", "linear": "sw"}.get(kind, kind)
        p = _params(d)
        hv = getattr(torch.ops.orihime, f"{ns}_marginals_hvp")(s, s.clone(), *p, None)
        assert tuple(hv.shape) == d["scores"].shape
        mb = getattr(torch.ops.orihime, f"{ns}_marginals_backward")(s, s.clone(), *p, None)
        assert tuple(mb[0].shape) == d["scores"].shape
        assert all(float(x) == 0.0 for x in mb[1:])


@pytest.mark.parametrize("name", EMPTY)
def test_empty_dimension_matches_reference_without_lengths_cpu(name):
    _check_empty(name, "cpu")


@pytest.mark.gpu
@pytest.mark.parametrize("name", EMPTY)
def test_empty_dimension_matches_reference_without_lengths_gpu(name):
    print(f"[case] empty-dimension fixture {name} on cuda:0", flush=True)
    _check_empty(name, "cuda")


def test_empty_fixture_set_covers_every_family():
    kinds = {str(load(os.path.join(GOLDEN, f))["kind"]) for f in EMPTY}
    assert kinds == set(OPS)


I32 = 2**31 - 1
GUARDS = [
    # (schema, params, shape, CPU message, GPU message)
    ("sw_affine_forward", (-2.0, -0.5, 1.0), (1, 0, 715_827_882),
     "sw_affine CPU DP table is too large", "sw_affine DP table too large"),
    ("sv_affine_forward", (-2.0, -0.5, 1.0), (1, 0, 715_827_882),
     "sv_affine CPU DP table is too large", "sv_affine DP table too large"),
    ("nw_affine_forward", (-2.0, -0.5, 1.0), (1, 0, 715_827_882),
     "nw_affine CPU DP workspace is too large", "nw_affine alpha workspace size must fit in int32"),
    ("sw_forward", (-1.0, 1.0), (1, 0, I32),
     "SW CPU DP table size exceeds supported int32 range",
     "SW CUDA DP table is too large for 32-bit kernel indexing"),
    ("sv_linear_forward", (-1.0, 1.0), (1, 0, 715_827_882),
     "SV linear CPU DP table size exceeds supported int32 range",
     "SV linear HIP DP table is too large for 32-bit kernel indexing"),
]


@pytest.mark.parametrize("schema,params,shape,cpu_msg,gpu_msg", GUARDS,
                         ids=[g[0] for g in GUARDS])
def test_reference_shape_guards_cpu(schema, params, shape, cpu_msg, gpu_msg):
    scores = torch.empty(shape, dtype=torch.float32)  # zero storage
    with pytest.raises(RuntimeError, match=cpu_msg):
        getattr(torch.ops.orihime, schema)(scores, *params, None)


def test_shapes_just_inside_the_guards_run_cpu():
    # 3 * (0 + 1) * (715827881 + 1) = 2147483646 <= INT32_MAX: accepted, empty result
    v, m = torch.ops.orihime.sw_affine_forward(torch.empty((1, 0, 715_827_881)), -2.0, -0.5,
                                               1.0, None)
    assert float(v) == 0.0 and m.numel() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("schema,params,shape,cpu_msg,gpu_msg", GUARDS,
                         ids=[g[0] for g in GUARDS])
def test_reference_shape_guards_gpu(schema, params, shape, cpu_msg, gpu_msg):
    print(f"[case] shape guard {schema} {shape} on cuda:0", flush=True)
    scores = torch.empty(shape, dtype=torch.float32, device="cuda")  # zero storage
    with pytest.raises(RuntimeError, match=gpu_msg):
        getattr(torch.ops.orihime, schema)(scores, *params, None)


@pytest.mark.gpu
def test_undersized_packed_workspace_makes_lost_pairs_nan():
    """The packed layout's workspace depends on the lengths, which live on the
    device: a caller that sizes it for too few cells must not get silent
    empty-alignment results for the pairs the plan kernel could not place."""
    print("[case] undersized packed workspace, 3 pairs 100x100", flush=True)
    g = torch.Generator().manual_seed(11)
    L = [(100, 100), (100, 100), (100, 100)]
    n = sum(a * b for a, b in L)
    scores = torch.randn(n, generator=g).cuda()
    lengths = torch.tensor(L, dtype=torch.int32).cuda()
    offsets = torch.tensor([0, 10_000, 20_000], dtype=torch.int64).cuda()

    def run(total_cells):
        need = _lib.workspace_bytes(True, _lib.OP_FORWARD, 3, 100, 100, total_cells)
        ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        out = dict(value=torch.empty(3, device="cuda"), map=torch.full((n,), 7.0, device="cuda"),
                   grad_open=torch.empty(3, device="cuda"), grad_ext=torch.empty(3, device="cuda"),
                   grad_temp=torch.empty(3, device="cuda"))
        _lib.affine_forward(scores, lengths, offsets, -8.0, -0.5, 0.5, L1max=100, L2max=100,
                            workspace=ws, **out)
        torch.cuda.synchronize()
        return {k: v.cpu() for k, v in out.items()}

    full, short = run(n), run(50)
    lost = torch.isnan(short["value"])
    assert lost.any() and not lost.all(), "expected a fitting prefix and at least one lost pair"
    for b in range(3):
        blk = slice(10_000 * b, 10_000 * (b + 1))
        if lost[b]:
            for k in ("grad_open", "grad_ext", "grad_temp"):
                assert torch.isnan(short[k][b])
            assert torch.isnan(short["map"][blk]).all()
        else:
            for k in ("value", "grad_open", "grad_ext", "grad_temp"):
                assert torch.equal(short[k][b], full[k][b]), k
            assert torch.equal(short["map"][blk], full["map"][blk])
    assert not torch.isnan(full["value"]).any()


@pytest.mark.gpu
def test_lost_pair_nan_stays_inside_its_clamped_block():
    """A lost pair whose lengths exceed (L1max, L2max) is clamped like every
    other pair, so its NaN fill covers only the L1max x L2max cells it owns."""
    print("[case] lost pair with lengths above the bounds, 3 pairs 100x100", flush=True)
    g = torch.Generator().manual_seed(12)
    n, tail = 30_000, 2_000
    scores = torch.randn(n + tail, generator=g).cuda()
    lengths = torch.tensor([(100, 100), (100, 100), (120, 100)], dtype=torch.int32).cuda()
    offsets = torch.tensor([0, 10_000, 20_000], dtype=torch.int64).cuda()
    need = _lib.workspace_bytes(True, _lib.OP_FORWARD, 3, 100, 100, 50)
    ws = torch.empty(need, dtype=torch.uint8, device="cuda")
    out = dict(value=torch.empty(3, device="cuda"), map=torch.full((n + tail,), 7.0, device="cuda"))
    _lib.affine_forward(scores, lengths, offsets, -8.0, -0.5, 0.5, L1max=100, L2max=100,
                        workspace=ws, **out)
    torch.cuda.synchronize()
    value, mp = out["value"].cpu(), out["map"].cpu()
    assert torch.isnan(value[2])
    assert torch.isnan(mp[20_000:n]).all()
    assert (mp[n:] == 7.0).all(), "NaN fill ran past the pair's clamped block"
