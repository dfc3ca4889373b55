"""The drop-in Python path under HIP graph capture.

The reference bindings copy lengths to the host to range-check them and read
tensor parameters with ``.item()`` (src/sw_affine/torch_hip.cpp:69-99), so
they cannot be captured.  Here every check that reads device data back is
skipped while the stream is capturing (the kernels clamp lengths into
[0, max]), and a tensor parameter, which would need such a read, is refused
with a clear error.

CPU tests: the deferral logic, with the capture state simulated, and the
packed layout's offset and length checks.  GPU tests: ``sw_affine`` (map and
value, forward and backward), the packed ``ragged`` form and the
embedding-scored path captured with ``torch.cuda.graph``, replayed several
times on new inputs with eager work and synchronisations between replays,
every replay equal bit for bit to the eager result on the same inputs.
"""

from __future__ import annotations

import pytest
import torch

import orihime_amd as ohm
from orihime_amd import _checks, _ops, embed, ragged

DEV = torch.device("cuda:0")


# --------------------------------------------------------------------------
# CPU: deferral logic and packed-layout checks
# --------------------------------------------------------------------------


def test_checks_run_outside_capture():
    assert not _checks.deferred(torch.zeros(1))
    with pytest.raises(RuntimeError, match="bad"):
        _checks.require(torch.tensor([True, False]), "bad")
    assert _checks.host_scalar(torch.tensor([-2.5]), "gap") == -2.5


def test_capture_skips_read_backs_and_refuses_tensor_parameters(monkeypatch):
    monkeypatch.setattr(_checks, "deferred", lambda t: isinstance(t, torch.Tensor))
    _checks.require(torch.tensor([False]), "must not raise while capturing")
    scores = torch.zeros(2, 4, 5)
    bad = torch.tensor([[9, 1], [1, -3]], dtype=torch.int32)
    assert _ops._check_lengths(bad, scores) is bad  # left to the kernels' clamp
    with pytest.raises(RuntimeError, match="Python number"):
        _checks.host_scalar(torch.tensor(0.5), "temperature")
    flat, ln, off = ragged.pack([torch.zeros(2, 3), torch.zeros(4, 1)])
    with pytest.raises(ValueError, match="L1max and L2max"):
        ragged.sw_affine_value(flat, ln, off, gap_open_score=-1.0, gap_extend_score=-0.5)
    # with explicit bounds the call proceeds to the (CPU) kernels
    v = ragged.sw_affine_value(flat, ln, off, gap_open_score=-1.0, gap_extend_score=-0.5,
                               L1max=4, L2max=3)
    assert v.shape == (2,)


@pytest.mark.parametrize("offsets,lengths,msg", [
    ([0, 6], [[2, 3], [4, 1]], "within the packed scores"),      # block 1 ends at 10 > 9
    ([-1, 5], [[2, 3], [4, 1]], "within the packed scores"),     # negative offset
    ([0, 5], [[2, 3], [5, 1]], "lengths must lie in"),           # L1 above L1max
    ([0, 5], [[2, -1], [4, 1]], "lengths must lie in"),          # negative length
])
def test_packed_offsets_and_lengths_are_checked(offsets, lengths, msg):
    flat = torch.zeros(9)
    ln = torch.tensor(lengths, dtype=torch.int32)
    off = torch.tensor(offsets, dtype=torch.int64)
    with pytest.raises(RuntimeError, match=msg):
        ragged.sw_affine_value(flat, ln, off, gap_open_score=-1.0, gap_extend_score=-0.5,
                               L1max=4, L2max=3)


def test_packed_blocks_that_fit_pass():
    flat = torch.randn(10)
    ln = torch.tensor([[2, 3], [4, 1]], dtype=torch.int32)
    off = torch.tensor([0, 6], dtype=torch.int64)
    v = ragged.sw_affine_value(flat, ln, off, gap_open_score=-1.0, gap_extend_score=-0.5)
    assert torch.isfinite(v).all()


# --------------------------------------------------------------------------
# GPU: captured graphs against eager calls
# --------------------------------------------------------------------------


def _capture(fn, *static):
    """Warm up on a side stream, then capture fn(*static) into a graph."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            fn(*static)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = fn(*static)
    return graph, out


def _replay_matches(graph, out, static, fresh, eager, replays=3):
    for r in range(replays):
        new = fresh(r)
        for s, n in zip(static, new):
            s.copy_(n)
        graph.replay()
        want = eager(*new)  # eager work on the same stream between replays
        torch.cuda.synchronize()
        for k, (a, b) in enumerate(zip(out, want)):
            assert torch.equal(a, b), f"replay {r}, output {k}"


def _padded(seed, B=6, L1=150, L2=230):
    g = torch.Generator().manual_seed(seed)
    scores = torch.randn(B, L1, L2, generator=g)
    lengths = torch.stack([torch.randint(0, L1 + 1, (B,), generator=g),
                           torch.randint(0, L2 + 1, (B,), generator=g)], dim=1)
    lengths[0] = torch.tensor([L1, L2])
    return scores.to(DEV), lengths.to(torch.int32).to(DEV)


P = dict(gap_open_score=-3.0, gap_extend_score=-0.5, temperature=0.7)


@pytest.mark.gpu
def test_sw_affine_map_and_value_captured():
    print("[gpu case] captured sw_affine map + value, B=6 150x230", flush=True)

    def fn(scores, lengths):
        return (ohm.sw_affine(scores, lengths=lengths, **P),
                ohm.sw_affine_value(scores, lengths=lengths, **P))

    static = _padded(0)
    graph, out = _capture(fn, *static)
    _replay_matches(graph, out, static, lambda r: _padded(10 + r), fn)


@pytest.mark.gpu
def test_sw_affine_value_backward_captured():
    print("[gpu case] captured sw_affine_value forward + backward, B=6 150x230", flush=True)

    def fn(scores, lengths):
        x = scores.detach().requires_grad_(True)
        v = ohm.sw_affine_value(x, lengths=lengths, **P)
        (g,) = torch.autograd.grad(v.sum(), x)
        return v.detach(), g

    static = _padded(1)
    graph, out = _capture(fn, *static)
    _replay_matches(graph, out, static, lambda r: _padded(20 + r), fn)


@pytest.mark.gpu
def test_ragged_value_captured_with_explicit_bounds():
    print("[gpu case] captured ragged sw_affine_value, 5 pairs", flush=True)
    shapes = [(100, 130), (1, 7), (64, 64), (0, 9), (129, 33)]

    def packed(seed):
        g = torch.Generator().manual_seed(seed)
        return ragged.pack([torch.randn(a, b, generator=g).to(DEV) for a, b in shapes])

    def fn(flat, lengths, offsets):
        return (ragged.sw_affine_value(flat, lengths, offsets, L1max=129, L2max=130, **P),)

    static = packed(2)
    graph, out = _capture(fn, *static)
    _replay_matches(graph, out, static, lambda r: packed(30 + r), fn)


@pytest.mark.gpu
def test_embedding_path_captured_forward_backward():
    print("[gpu case] captured soft_align_embeddings fwd + bwd, B=4 200x180 D=64", flush=True)
    B, M, N, D = 4, 200, 180, 64

    def inputs(seed):
        g = torch.Generator().manual_seed(seed)
        Q = (torch.randn(B, M, D, generator=g) / 8).bfloat16().to(DEV)
        K = (torch.randn(B, N, D, generator=g) / 8).bfloat16().to(DEV)
        ln = torch.tensor([[200, 180], [17, 150], [199, 1], [64, 100]], dtype=torch.int32)
        return Q, K, ln.to(DEV)

    def fn(Q, K, ln):
        q, k = Q.detach().requires_grad_(True), K.detach().requires_grad_(True)
        v = embed.soft_align_embeddings(q, k, ln)
        gq, gk = torch.autograd.grad(v.sum(), (q, k))
        return v.detach(), gq, gk

    static = inputs(3)
    graph, out = _capture(fn, *static)
    _replay_matches(graph, out, static, lambda r: inputs(40 + r), fn)


@pytest.mark.gpu
def test_tensor_parameter_refused_while_capturing():
    print("[gpu case] tensor temperature under capture raises", flush=True)
    scores, lengths = _padded(4, B=2, L1=40, L2=50)
    temp = torch.tensor(0.7, device=DEV)
    ohm.sw_affine_value(scores, lengths=lengths, gap_open_score=-3.0, gap_extend_score=-0.5,
                        temperature=temp)  # fine eagerly
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with pytest.raises(RuntimeError, match="Python number"):
        with torch.cuda.graph(graph):
            ohm.sw_affine_value(scores, lengths=lengths, gap_open_score=-3.0,
                                gap_extend_score=-0.5, temperature=temp)
    torch.cuda.synchronize()
