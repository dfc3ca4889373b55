"""CPU tests: the C ABI library, the host implementation behind it, and the
drop-in Python surface (engine, named primitives, ops, nn, raw, ragged).

Numeric expectations come from the golden fixtures (reference CPU kernels)
and from the FP64 oracle; tolerances follow the reference suite
(tests/test_soft_sw_affine.py:29-30, :178-195): values 1e-5 relative, maps
1e-4 absolute.
"""

import ctypes
import os
import re
import warnings

import numpy as np
import pytest
import torch

import orihime_amd as ohm
from orihime_amd import _lib
from helpers import abs_err, affine_params, close_or_nearer_truth, golden_files, load, rel_err
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    declared = set()
    for h in ("ohm_sw.h", "ohm_embed.h"):
        header = open(os.path.join(ROOT, "include", h)).read()
        declared |= set(re.findall(r"^\S.*?\b(ohm_\w+)\(", header, re.M))
    assert len(declared) >= 12
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) == declared
    assert ohm.__version__ == _lib.version()


def test_probe_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "ohm_probe.h")).read()
    declared = set(re.findall(r"^\S.*?\b(ohm_\w+)\(", header, re.M))
    assert declared == {"ohm_probe_ops_per_thread", "ohm_probe_launch"}
    lib = ctypes.CDLL(os.path.join(ROOT, "orihime_amd", "libohm_probe.so"))
    for name in declared:
        assert hasattr(lib, name), name
    # host-side argument checks (no device needed)
    lib.ohm_probe_ops_per_thread.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    ops = ctypes.c_int64()
    assert lib.ohm_probe_ops_per_thread(4, 10, ctypes.byref(ops)) == 0 and ops.value == 400
    assert lib.ohm_probe_ops_per_thread(9, 10, ctypes.byref(ops)) == 1


def test_launch_config_hook_validates_and_resets():
    with pytest.raises(RuntimeError):
        _lib.set_launch_config(waves=3)
    with pytest.raises(RuntimeError):
        _lib.set_launch_config(teams=5)
    _lib.set_launch_config(4, 2, 2)
    _lib.set_launch_config()


def test_checkpoint_stride_hook_and_workspace_sizes():
    """The tile hook accepts the strides that divide the chunk width, and the
    affine workspace query follows it: a shorter stride stores more
    checkpoints, a longer one never less than the log-domain sweeps need."""
    assert _lib.checkpoint_stride() == 8
    for bad in (1, 2, 3, 12, 32, -4):
        with pytest.raises(RuntimeError, match="checkpoint stride"):
            _lib.set_checkpoint_stride(bad)
    sizes = {}
    try:
        for ck in (4, 8, 16):
            _lib.set_checkpoint_stride(ck)
            assert _lib.checkpoint_stride() == ck
            sizes[ck] = _lib.workspace_bytes(True, _lib.OP_FORWARD, 64, 700, 900)
            assert _lib.workspace_bytes(False, _lib.OP_FORWARD, 64, 700, 900) == \
                _lib.workspace_bytes(False, _lib.OP_FORWARD, 64, 700, 900)
    finally:
        _lib.set_checkpoint_stride()
    assert _lib.checkpoint_stride() == 8
    assert sizes[4] > sizes[8] == sizes[16]
    # the table part (6 planes per stored step) doubles from stride 8 to 4
    assert sizes[4] - sizes[8] > 0.4 * sizes[8]


def test_gpu_entry_points_reject_bad_arguments_without_a_device():
    # argument checks run before any HIP call
    lib = _lib._lib
    value = np.zeros(1, np.float32)  # never written: every call below fails its checks
    vp = value.ctypes.data
    st = lib.ohm_sw_affine_forward(None, None, None, 1, 4, 4, -1.0, -1.0, 0.0, vp, None, None,
                                   None, None, None, 0, None)
    assert st == 1 and b"temperature" in lib.ohm_last_error()
    st = lib.ohm_sw_affine_forward(None, None, None, 1, 4, 4, -1.0, -1.0, 1.0, vp, None, None,
                                   None, None, None, 0, None)
    assert st == 1 and b"scores" in lib.ohm_last_error()
    st = lib.ohm_sw_affine_tangent(None, None, None, 1, 4, 4, -1.0, -1.0, 1.0, None, 1.0, 0.0,
                                   0.0, None, vp, None, None, None, 0, None)
    assert st == 1 and b"pure score direction" in lib.ohm_last_error()
    # value only < forward + reverse (checkpoints every 8 steps, ~2.5 B per
    # cell) < tangent passes (full tables, 24 B per cell)
    ws = [lib.ohm_sw_affine_workspace(op, 2, 512, 512, -1) for op in range(3)]
    assert ws[0] < ws[1] < ws[2]
    assert ws[1] > 2 * 512 * 512 * 2.5 and ws[2] > 2 * 512 * 512 * 24


# --------------------------------------------------------------------------
# Host implementation vs golden fixtures and the FP64 oracle
# --------------------------------------------------------------------------


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def entropy_tol(value, temp):
    """grad_T = (S - E[score]) / T subtracts two numbers of size |S|; FP32
    rounding of either leaves ~|S| * 2^-23 / T of noise, which dominates at
    low temperature (the reference's own FP32 result carries it too)."""
    return 1e-4 + 8.0 * (np.abs(np.asarray(value, np.float64)) + 1.0) * 2.0 ** -23 / temp


@pytest.mark.parametrize("path", golden_files("affine"), ids=lambda p: p.split("/")[-1])
def test_affine_cpu_path_matches_golden(path):
    d = load(path)
    go, ge, T = affine_params(d)
    s, ln = _t(d["scores"]), _t(d["lengths"])
    v, m, a, b, c = torch.ops.orihime.soft_sw_affine_with_grads(s, go, ge, T, ln)
    assert rel_err(v, d["value"], 1e-3) < 1e-5
    assert abs_err(m, d["map"]) < 1e-4
    for got, ref in ((a, "grad_open"), (b, "grad_ext")):
        assert rel_err(got, d[ref], 1.0) < 1e-4, ref
    assert np.all(np.abs(c.numpy() - d["grad_temp"]) <= entropy_tol(d["value"], T) * np.maximum(1, np.abs(d["grad_temp"])))
    if "hvp" in d:
        V = _t(d["V"])
        s64 = d["scores"].astype(np.float64)
        lnn = d["lengths"]
        # second-order outputs scale like 1/T: tolerance relative to their size
        tol = lambda r: 1e-4 * max(1.0, np.abs(r).max(initial=0.0))
        got = torch.ops.orihime.sw_affine_marginals_hvp(s, V, go, ge, T, ln)
        assert close_or_nearer_truth(got, d["hvp"], lambda: O.sw_affine_hvp(s64, d["V"], go, ge, T, lnn, prec="f64"), tol(d["hvp"]))
        for p, n in enumerate(("gap_open", "gap_ext", "temp")):
            got = getattr(torch.ops.orihime, f"sw_affine_marginals_grad_{n}")(s, go, ge, T, ln)
            truth = lambda p=p: O.sw_affine_param_jacobian(s64, p, go, ge, T, lnn, prec="f64")
            assert close_or_nearer_truth(got, d[f"sens_{n}"], truth, tol(d[f"sens_{n}"])), n
        gs, x, y, z = torch.ops.orihime.sw_affine_marginals_backward(s, _t(d["G"]), go, ge, T, ln)
        truth = lambda: O.sw_affine_backward_full(s64, d["G"], go, ge, T, lnn, prec="f64")[0]
        assert close_or_nearer_truth(gs, d["mb_scores"], truth, tol(d["mb_scores"]))
        full = lambda: O.sw_affine_backward_full(s64, d["G"], go, ge, T, lnn, prec="f64")
        for k, (got, ref) in enumerate(((x, "mb_gap_open"), (y, "mb_gap_ext"), (z, "mb_temp"))):
            truth = lambda k=k: full()[k + 1]
            assert close_or_nearer_truth(float(got), float(d[ref]), truth, 1e-3 * max(1.0, abs(float(d[ref])))), ref


@pytest.mark.parametrize("path", golden_files("linear"), ids=lambda p: p.split("/")[-1])
def test_linear_cpu_path_matches_golden(path):
    d = load(path)
    g, T = float(d["gap"]), float(d["temp"])
    s, ln = _t(d["scores"]), _t(d["lengths"])
    v, m, a, b = torch.ops.orihime.soft_sw_with_grads(s, g, T, ln)
    assert rel_err(v, d["value"], 1e-3) < 1e-5
    assert abs_err(m, d["map"]) < 1e-4
    assert rel_err(a, d["grad_gap"], 1.0) < 1e-4
    assert np.all(np.abs(b.numpy() - d["grad_temp"]) <= entropy_tol(d["value"], T) * np.maximum(1, np.abs(d["grad_temp"])))
    if "hvp" in d:
        s64, lnn = d["scores"].astype(np.float64), d["lengths"]
        tol = lambda r: 1e-4 * max(1.0, np.abs(r).max(initial=0.0))
        got = torch.ops.orihime.sw_marginals_hvp(s, _t(d["V"]), g, T, ln)
        assert close_or_nearer_truth(got, d["hvp"], lambda: O.sw_hvp(s64, d["V"], g, T, lnn, prec="f64"), tol(d["hvp"]))
        for p, n in enumerate(("gap", "temp")):
            got = getattr(torch.ops.orihime, f"sw_marginals_grad_{n}")(s, g, T, ln)
            truth = lambda p=p: O.sw_param_jacobian(s64, p, g, T, lnn, prec="f64")
            assert close_or_nearer_truth(got, d[f"sens_{n}"], truth, tol(d[f"sens_{n}"])), n


def test_cpu_path_is_thread_count_invariant():
    x = torch.randn(6, 21, 17, generator=torch.Generator().manual_seed(5))
    outs = []
    old = torch.get_num_threads()
    try:
        for n in (1, 2, 4):
            torch.set_num_threads(n)
            outs.append(torch.ops.orihime.soft_sw_affine_with_grads(x, -1.0, -0.3, 0.8, None))
    finally:
        torch.set_num_threads(old)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert torch.equal(a, b)


# --------------------------------------------------------------------------
# Public API behaviour
# --------------------------------------------------------------------------

KW = dict(gap_open_score=-1.0, gap_extend_score=-0.3, temperature=0.9)


def test_public_functions_and_autograd():
    g = torch.Generator().manual_seed(0)
    x = torch.randn(3, 9, 7, generator=g)
    ref = O.sw_affine_forward(x.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs_err(ohm.sw_affine(x, **KW), ref[1]) < 1e-5
    assert rel_err(ohm.sw_affine_value(x, **KW), ref[0]) < 1e-5
    assert abs_err(ohm.sw_affine_entropy(x, **KW), ref[4]) < 1e-4
    xr = x.clone().requires_grad_(True)
    ohm.sw_affine_value(xr, **KW).sum().backward()
    assert abs_err(xr.grad, ref[1]) < 1e-5
    G = torch.randn(x.shape, generator=g)
    xr = x.clone().requires_grad_(True)
    (ohm.sw_affine(xr, **KW) * G).sum().backward()
    assert abs_err(xr.grad, O.sw_affine_hvp(x.numpy(), G.numpy(), -1.0, -0.3, 0.9, prec="f64")) < 1e-5
    xr = x.clone().requires_grad_(True)
    ohm.sw_affine_entropy(xr, **KW).sum().backward()
    dT = O.sw_affine_param_jacobian(x.numpy(), 2, -1.0, -0.3, 0.9, prec="f64")
    assert abs_err(xr.grad, dT) < 1e-4


def test_tensor_parameter_gradients():
    x = torch.randn(2, 8, 6, generator=torch.Generator().manual_seed(1))
    go = torch.tensor(-1.0, requires_grad=True)
    ge = torch.tensor([-0.3], requires_grad=True)
    T = torch.tensor(0.9, requires_grad=True)
    ohm.sw_affine_value(x, gap_open_score=go, gap_extend_score=ge, temperature=T).sum().backward()
    _, _, gO, gE, gT = O.sw_affine_forward(x.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs(go.grad.item() - gO.sum()) < 1e-4
    assert ge.grad.shape == (1,) and abs(ge.grad.item() - gE.sum()) < 1e-4
    assert abs(T.grad.item() - gT.sum()) < 1e-3
    G = torch.randn(x.shape, generator=torch.Generator().manual_seed(2))
    go.grad = None
    (ohm.sw_affine(x, gap_open_score=go, gap_extend_score=-0.3, temperature=0.9) * G).sum().backward()
    mb = O.sw_affine_backward_full(x.numpy(), G.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs(go.grad.item() - mb[1]) < 1e-4


def test_entropy_parameter_direction_is_rejected():
    x = torch.randn(1, 4, 4)
    T = torch.tensor(1.0, requires_grad=True)
    with pytest.raises(NotImplementedError, match="temperature"):
        ohm.sw_affine_entropy(x, temperature=T).sum().backward()


def test_linear_public_surface():
    x = torch.randn(2, 7, 9, generator=torch.Generator().manual_seed(3))
    ref = O.sw_forward(x.numpy(), -0.7, 0.9, prec="f64")
    assert abs_err(ohm.sw(x, gap_score=-0.7, temperature=0.9), ref[1]) < 1e-5
    assert rel_err(ohm.sw_value(x, gap_score=-0.7, temperature=0.9), ref[0]) < 1e-5
    assert abs_err(ohm.sw_entropy(x, gap_score=-0.7, temperature=0.9), ref[3]) < 1e-4


def test_ops_named_operations():
    x = torch.randn(2, 6, 5, generator=torch.Generator().manual_seed(4))
    G = torch.randn(x.shape, generator=torch.Generator().manual_seed(5))
    out = ohm.ops.sw_affine.forward(x, **KW)
    assert set(out) == {"map", "value", "entropy"}
    assert isinstance(ohm.ops.sw_affine.forward(x, output="value", **KW), torch.Tensor)
    bw = ohm.ops.sw_affine.backward(x, grad_map=G, **KW)
    assert set(bw) == {"pair_scores", "gap_open_score", "gap_extend_score", "temperature"}
    mb = O.sw_affine_backward_full(x.numpy(), G.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs_err(bw["pair_scores"], mb[0]) < 1e-5
    for k, ref in zip(("gap_open_score", "gap_extend_score", "temperature"), mb[1:]):
        assert bw[k].shape == () and abs(bw[k].item() - ref) < 1e-3 * max(1, abs(ref))
    sens = ohm.ops.sw_affine.sensitivity(x, **KW)
    for p, k in enumerate(("gap_open_score", "gap_extend_score", "temperature")):
        ref = O.sw_affine_param_jacobian(x.numpy(), p, -1.0, -0.3, 0.9, prec="f64")
        assert abs_err(sens[k], ref) < 1e-4
    with pytest.raises(ValueError, match="unknown output field"):
        ohm.ops.sw_affine.forward(x, output="nope")
    assert set(ohm.ops.sw.backward(x, grad_map=G, gap_score=-0.5)) == {"pair_scores", "gap_score", "temperature"}


def test_nn_modules():
    x = torch.randn(2, 5, 6, generator=torch.Generator().manual_seed(6))
    mod = ohm.nn.SmithWatermanAffine(gap_open_score=-1.0, gap_extend_score=-0.3, temperature=0.9,
                                     learnable=("gap_open_score", "temperature"))
    assert isinstance(mod.gap_open_score, torch.nn.Parameter)
    assert "gap_extend_score" in dict(mod.named_buffers())
    mod.value(x).sum().backward()
    _, _, gO, _, gT = O.sw_affine_forward(x.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs(mod.gap_open_score.grad.item() - gO.sum()) < 1e-4
    assert abs(mod.temperature.grad.item() - gT.sum()) < 1e-3
    assert mod(x).shape == x.shape and mod.entropy(x).shape == (2,)
    lin = ohm.nn.SmithWaterman(gap_score=-0.5, learnable=("gap_score",))
    lin.value(x).sum().backward()
    assert lin.gap_score.grad is not None
    with pytest.raises(ValueError, match="unknown learnable"):
        ohm.nn.SmithWaterman(learnable=("nope",))


def test_raw_tier():
    x = torch.randn(2, 5, 4, generator=torch.Generator().manual_seed(7))
    G = torch.randn(x.shape, generator=torch.Generator().manual_seed(8))
    out = ohm.raw.sw_affine.vjp(x, wrt=("gap_open_score", "temperature"), cotangent=G,
                                gap_open_score=-1.0, gap_extend_score=-0.3, temperature=0.9)
    mb = O.sw_affine_backward_full(x.numpy(), G.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs(out["gap_open_score"].item() - mb[1]) < 1e-4
    assert abs(out["temperature"].item() - mb[3]) < 1e-3
    v, m = ohm.raw.sw_affine.forward(x, -1.0, -0.3, 0.9)
    assert m.shape == x.shape and v.shape == (2,)
    with pytest.raises(TypeError, match="float32"):
        ohm.raw.sw_affine.vjp_one(x, wrt="temperature", cotangent=G.double())
    with pytest.raises(ValueError, match="contiguous"):
        ohm.raw.sw.vjp_one(x, wrt="gap_score", cotangent=G.transpose(1, 2).contiguous().transpose(1, 2))


def test_lengths_and_prefix_masks():
    x = torch.randn(3, 9, 7, generator=torch.Generator().manual_seed(9))
    ln = torch.tensor([[9, 7], [5, 4], [0, 3]], dtype=torch.int32)
    m = ohm.sw_affine(x, lengths=ln, **KW)
    assert m[1, 5:].abs().max() == 0 and m[1, :, 4:].abs().max() == 0 and m[2].abs().max() == 0
    v = ohm.sw_affine_value(x, lengths=ln, **KW)
    v1 = ohm.sw_affine_value(x[1:2, :5, :4].contiguous(), **KW)
    assert abs(v[1] - v1[0]) < 1e-5 and v[2] == 0
    mask1 = torch.arange(9)[None, :] < ln[:, 0:1]
    mask2 = torch.arange(7)[None, :] < ln[:, 1:2]
    from orihime_amd.families import engine; _sw_affine_operator = engine("sw_affine")

    vm = _sw_affine_operator.value(x, -1.0, -0.3, 0.9, mask1=mask1, mask2=mask2)
    assert torch.allclose(vm, v)
    with pytest.raises(ValueError, match="prefix-only"):
        bad = mask1.clone()
        bad[0, 0] = False
        _sw_affine_operator.value(x, -1.0, -0.3, 0.9, mask1=bad)


def test_mask_matches_manual_sentinel():
    x = torch.randn(2, 6, 6, generator=torch.Generator().manual_seed(10))
    mask = torch.zeros_like(x, dtype=torch.bool)
    mask[:, 2, 3] = True
    mask[1, 4] = True
    a = ohm.sw_affine(x, mask=mask, **KW)
    b = ohm.sw_affine(x.masked_fill(mask, float("-inf")), **KW)
    assert torch.equal(a, b)
    assert a[mask].abs().max() == 0


def test_validation_messages():
    x = torch.randn(2, 5, 5)
    with pytest.raises(ValueError, match=r"must have shape \[B, L1, L2\]"):
        ohm.sw_affine(torch.randn(5, 5))
    with pytest.raises(ValueError, match="B=0"):
        ohm.sw_affine(torch.randn(0, 5, 5))
    with pytest.raises(TypeError, match="float32"):
        ohm.sw_affine(x.double())
    with pytest.raises(ValueError, match="strictly positive"):
        ohm.sw_affine(x, temperature=0.0)
    with pytest.raises(RuntimeError, match="numerical domain"):
        ohm.sw_affine(x * 100, temperature=1.0)
    with pytest.raises(RuntimeError, match="unsupported infinity"):
        ohm.sw_affine(x.masked_fill(x > 1, float("inf")))
    with pytest.raises(RuntimeError, match="within the padded input shape"):
        ohm.sw_affine(x, lengths=torch.tensor([[6, 5], [5, 5]], dtype=torch.int32))
    with pytest.raises(TypeError, match="int32"):
        ohm.sw_affine(x, lengths=torch.tensor([[5, 5], [5, 5]]))
    with pytest.raises(ValueError, match=r"shape \[\] or \[1\]"):
        ohm.sw_affine(x, gap_open_score=torch.zeros(2))
    with pytest.raises(ValueError, match="only torch.float32"):
        ohm.sw_affine(x, dtype=torch.float64)
    with pytest.raises(TypeError, match="mask must be a boolean"):
        ohm.sw_affine(x, mask=torch.zeros_like(x))
    with pytest.warns(UserWarning, match="positive"):
        ohm.sw_affine(x, gap_open_score=0.5)
    # native boundary messages (src/sw_affine/torch_hip.cpp:35-99)
    with pytest.raises(RuntimeError, match=r"lengths\[1,0\] must be between 0 and 5, got 9"):
        torch.ops.orihime.sw_affine_forward(x, -1.0, -1.0, 1.0,
                                            torch.tensor([[5, 5], [9, 5]], dtype=torch.int32))
    with pytest.raises(RuntimeError, match="scores must be 3D"):
        torch.ops.orihime.sw_affine_forward(torch.randn(5, 5), -1.0, -1.0, 1.0, None)
    with pytest.raises(RuntimeError, match="same shape"):
        torch.ops.orihime.sw_affine_marginals_hvp(x, torch.randn(2, 5, 4), -1.0, -1.0, 1.0, None)
    with pytest.raises(RuntimeError, match="must be contiguous"):
        ohm.ops._kernels["sw_affine"].marginals_hvp(x, torch.randn(2, 5, 5).transpose(1, 2), -1, -1, 1)


def test_dtype_escape_hatch_and_autocast():
    x = torch.randn(2, 5, 6, generator=torch.Generator().manual_seed(11))
    a = ohm.sw_affine_value(x.bfloat16(), dtype=torch.float32, **KW)
    b = ohm.sw_affine_value(x.bfloat16().float(), **KW)
    assert torch.equal(a, b)
    xr = x.bfloat16().requires_grad_(True)
    with torch.autocast("cpu", dtype=torch.bfloat16):
        v = ohm.sw_affine_value(xr, **KW)
    assert v.dtype == torch.float32
    v.sum().backward()
    assert xr.grad.dtype == torch.bfloat16


def test_vmap_over_named_primitives():
    x = torch.randn(3, 2, 5, 4, generator=torch.Generator().manual_seed(12))
    vm = torch.vmap(lambda s: torch.ops.orihime.sw_affine_forward(s, -1.0, -0.3, 0.9, None)[0])(x)
    for k in range(3):
        assert torch.allclose(vm[k], torch.ops.orihime.sw_affine_forward(x[k], -1.0, -0.3, 0.9, None)[0])
    G = torch.randn(x.shape, generator=torch.Generator().manual_seed(13))
    out = torch.vmap(lambda s, g: torch.ops.orihime.sw_affine_marginals_backward(s, g, -1.0, -0.3, 0.9, None))(x, G)
    for k in range(3):
        ref = torch.ops.orihime.sw_affine_marginals_backward(x[k], G[k].contiguous(), -1.0, -0.3, 0.9, None)
        assert torch.allclose(out[0][k], ref[0], atol=1e-6)
        for j in range(1, 4):
            assert torch.allclose(out[j][k], ref[j], atol=1e-5, rtol=1e-5)


def test_fake_tensor_shapes():
    from torch._subclasses.fake_tensor import FakeTensorMode

    with FakeTensorMode():
        x = torch.empty(4, 7, 9)
        v, m = torch.ops.orihime.sw_affine_forward(x, -1.0, -1.0, 1.0, None)
        assert v.shape == (4,) and m.shape == (4, 7, 9)
        gs, a, b, c = torch.ops.orihime.sw_affine_marginals_backward(x, x, -1.0, -1.0, 1.0, None)
        assert gs.shape == x.shape and a.shape == (1,)
        assert torch.ops.orihime.sw_value_grad_params(x, -1.0, 1.0, None)[0].shape == (4,)


def test_forward_primitive_autograd():
    x = torch.randn(2, 6, 5, generator=torch.Generator().manual_seed(14)).requires_grad_(True)
    v, m = torch.ops.orihime.sw_affine_forward(x, -1.0, -0.3, 0.9, None)
    G = torch.randn(x.shape, generator=torch.Generator().manual_seed(15))
    (v.sum() + (m * G).sum()).backward()
    ref = O.sw_affine_forward(x.detach().numpy(), -1.0, -0.3, 0.9, prec="f64")[1]
    H = O.sw_affine_hvp(x.detach().numpy(), G.numpy(), -1.0, -0.3, 0.9, prec="f64")
    assert abs_err(x.grad, ref + H) < 1e-5
    go = torch.tensor([-1.0], requires_grad=True)
    ln = torch.tensor([[6, 5], [4, 5]], dtype=torch.int32)
    v, m = torch.ops.orihime.sw_affine_forward_t(x.detach(), go, torch.tensor([-0.3]),
                                                 torch.tensor([0.9]), ln)
    v.sum().backward()
    gO = O.sw_affine_forward(x.detach().numpy(), -1.0, -0.3, 0.9, ln.numpy(), prec="f64")[2]
    assert abs(go.grad.item() - gO.sum()) < 1e-4


def test_ragged_packed_matches_padded_lengths():
    g = torch.Generator().manual_seed(16)
    pairs = [torch.randn(a, b, generator=g) for a, b in ((9, 7), (5, 4), (1, 8), (12, 3))]
    flat, ln, off = ohm.ragged.pack(pairs)
    pad = torch.zeros(4, 12, 8)
    for k, p in enumerate(pairs):
        pad[k, :p.shape[0], :p.shape[1]] = p
    v = ohm.ragged.sw_affine_value(flat, ln, off, **KW)
    vp = ohm.sw_affine_value(pad, lengths=ln, **KW)
    assert torch.allclose(v, vp, rtol=1e-6)
    m = ohm.ragged.sw_affine(flat, ln, off, **KW)
    mp = ohm.sw_affine(pad, lengths=ln, **KW)
    for k, mk in enumerate(ohm.ragged.unpack(m, ln, off)):
        assert torch.allclose(mk, mp[k, :mk.shape[0], :mk.shape[1]], atol=1e-7)
    fr = flat.clone().requires_grad_(True)
    ohm.ragged.sw_affine_value(fr, ln, off, **KW).sum().backward()
    assert torch.allclose(fr.grad, m, atol=1e-7)


def _embed_truth(Q, K, ln, go, ge, T):
    """FP64 oracle of the embedding-scored alignment: scores Q K^T in float64,
    the oracle's value and map, dQ = P K and dK = P^T Q per pair."""
    Q64, K64 = Q.double().numpy(), K.double().numpy()
    S = np.einsum("bmd,bnd->bmn", Q64, K64)
    out = O.sw_affine_forward(S, go, ge, T, ln.numpy(), prec="f64")
    P = out[1]
    return out[0], np.einsum("bmn,bnd->bmd", P, K64), np.einsum("bmn,bmd->bnd", P, Q64)


def test_embedding_scored_alignment_cpu_matches_fp64_oracle():
    from orihime_amd.embed import soft_align_embeddings
    g = torch.Generator().manual_seed(4)
    B, M, N, D = 3, 40, 50, 32
    go, ge, T = -4.0, -0.5, 0.25
    Q = (torch.randn(B, M, D, generator=g) / D ** 0.5).bfloat16().requires_grad_()
    K = (torch.randn(B, N, D, generator=g) / D ** 0.5).bfloat16().requires_grad_()
    ln = torch.tensor([[40, 50], [31, 22], [7, 50]], dtype=torch.int32)
    v = soft_align_embeddings(Q, K, ln, go, ge, T)
    v.sum().backward()
    val, dQ, dK = _embed_truth(Q.detach(), K.detach(), ln, go, ge, T)
    assert rel_err(v.detach().numpy(), val, 1e-3) < 1e-5
    # gradients are BF16 like the inputs: half an ulp is 2^-9 of the entry
    gq, gk = Q.grad.float().numpy(), K.grad.float().numpy()
    assert np.all(np.abs(gq - dQ) <= 2.0 ** -8 * np.abs(dQ) + 1e-6)
    assert np.all(np.abs(gk - dK) <= 2.0 ** -8 * np.abs(dK) + 1e-6)
    assert float(Q.grad[1, 31:].abs().max()) == 0.0 and float(K.grad[1, 22:].abs().max()) == 0.0
    with pytest.raises(TypeError, match="bfloat16"):
        soft_align_embeddings(Q.detach().float(), K.detach().float(), ln, go, ge, T)
    # the in-sweep score generation is a GPU path of its own
    with pytest.raises(ValueError, match="fused sweeps take GPU tensors"):
        soft_align_embeddings(Q, K, ln, go, ge, T, fused=True)


@pytest.mark.parametrize("affine", [True, False], ids=["affine", "linear"])
def test_tensor_parameter_gradients_reuse_the_forward_sweeps(affine, monkeypatch):
    """The tensor-parameter forwards keep the parameter gradients of their own
    sweeps, so backward runs no second forward for them (the reference
    recomputes value_grad_params, orihime/ops/sw_affine.py:133-146); the
    result equals the explicit value_grad_params."""
    g = torch.Generator().manual_seed(11)
    x = torch.randn(3, 20, 30, generator=g)
    ln = torch.tensor([[20, 30], [7, 30], [20, 4]], dtype=torch.int32)
    w = torch.tensor([1.0, -2.0, 0.5])
    if affine:
        params = [torch.tensor([-2.0], requires_grad=True), torch.tensor([-0.5], requires_grad=True),
                  torch.tensor([0.7], requires_grad=True)]
        op, vgp, name = torch.ops.orihime.sw_affine_forward_t, torch.ops.orihime.sw_affine_value_grad_params, "affine_forward"
    else:
        params = [torch.tensor([-1.0], requires_grad=True), torch.tensor([0.7], requires_grad=True)]
        op, vgp, name = torch.ops.orihime.sw_forward_t, torch.ops.orihime.sw_value_grad_params, "linear_forward"
    calls = []
    orig = getattr(_lib, name)

    def counted(*a, **k):
        calls.append(1)
        return orig(*a, **k)

    monkeypatch.setattr(_lib, name, counted)
    v, _ = op(x, *params, ln)
    n = len(calls)
    (v * w).sum().backward()
    assert len(calls) == n
    ref = vgp(x, *[float(p.detach()) for p in params], ln)
    for p, r in zip(params, ref):
        assert torch.allclose(p.grad, (w * r).sum().reshape(1), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("family", ["sw_affine", "sw", "nw_affine", "sv_affine"])
def test_value_only_backward_runs_no_second_order_pass(family, monkeypatch):
    """A backward through the value alone of ``<family>_forward`` runs no
    tangent (second-order) pass; a backward that also goes through the map
    runs one (reference: tests/test_soft_sw_affine.py:1088-1115)."""
    name = {"sw_affine": "affine_tangent", "sw": "linear_tangent",
            "nw_affine": "_three_state_tangent", "sv_affine": "_three_state_tangent"}[family]
    calls = []
    orig = getattr(_lib, name)

    def counted(*a, **k):
        calls.append(1)
        return orig(*a, **k)

    monkeypatch.setattr(_lib, name, counted)
    kern = ohm.ops._kernels[family]
    params = (-1.0, 0.7) if family == "sw" else (-1.0, -0.3, 0.7)
    g = torch.Generator().manual_seed(776)
    x = torch.randn(1, 8, 9, generator=g).requires_grad_(True)
    value, _ = kern.forward(x, *params, None)
    value.sum().backward()
    assert calls == [] and x.grad is not None
    x.grad = None
    value, posteriors = kern.forward(x, *params, None)
    (value.sum() + 0.0 * posteriors.sum()).backward()
    assert len(calls) >= 1


def test_public_surface_signatures_docs_and_selection():
    """Public functions keep the reference's frozen signatures and documented
    arguments (tests/test_v3_contract.py:280-310, tests/test_public_docs.py:283-293);
    named operations validate ``output`` (tests/test_public_operations.py:106-120)."""
    import inspect

    kw = inspect.Parameter.KEYWORD_ONLY
    for fam, gaps in (("sw", ("gap_score",)), ("sw_affine", ("gap_open_score", "gap_extend_score")),
                      ("nw", ("gap_score",)), ("nw_affine", ("gap_open_score", "gap_extend_score")),
                      ("sv", ("gap_score",)), ("sv_affine", ("gap_open_score", "gap_extend_score"))):
        for suffix in ("", "_value", "_entropy"):
            fn = getattr(ohm, fam + suffix)
            sig = inspect.signature(fn)
            assert tuple(sig.parameters) == ("pair_scores", *gaps, "temperature", "lengths", "mask",
                                             "dtype")
            assert sig.parameters["pair_scores"].default is inspect.Parameter.empty
            for g in gaps:
                assert sig.parameters[g].kind is kw and sig.parameters[g].default == 0.0
            assert sig.parameters["temperature"].default == 1.0
            assert all(sig.parameters[n].default is None for n in ("lengths", "mask", "dtype"))
            doc = inspect.getdoc(fn)
            for part in ("Args:", "Returns:", "Raises:", *(f"{n}:" for n in sig.parameters)):
                assert part in doc, (fam + suffix, part)
            assert fn.__name__ == fam + suffix
        op = getattr(ohm.ops, fam)
        assert op.backward_fields == ("pair_scores", *gaps, "temperature")
        assert tuple(inspect.signature(op.backward).parameters)[-2:] == ("grad_map", "output")
    assert tuple(ohm.ops.__all__) == ("sw", "sw_affine", "nw_affine", "nw", "sv", "sv_affine")
    x = torch.randn(1, 3, 3)
    with pytest.raises(ValueError, match="unknown output field"):
        ohm.ops.sw.forward(x, output="scores")
    with pytest.raises(ValueError, match="at least one"):
        ohm.ops.sw.sensitivity(x, output=())
    with pytest.raises(ValueError, match="duplicate"):
        ohm.ops.sw.backward(x, grad_map=torch.ones_like(x), output=("temperature", "temperature"))
    with pytest.raises(TypeError, match="output must be"):
        ohm.ops.sw.forward(x, output=object())
    with pytest.raises(TypeError, match="unexpected keyword"):
        ohm.sw(x, gap_open_score=-1.0)
    assert tuple(inspect.signature(ohm.nn.SmithWatermanAffine).parameters) == (
        "gap_open_score", "gap_extend_score", "temperature", "learnable", "device", "dtype")
    layer = ohm.nn.NW(gap_score=-0.5)
    assert "gap_score" in dict(layer.named_buffers()) and "learnable=()" in repr(layer)
