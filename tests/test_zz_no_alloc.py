"""The allocation-free entry point (ohm_sw_affine_forward_noalloc) against
ohm_sw_affine_forward with a caller workspace, through the C ABI: two calls
in a row, the second on a smaller batch (the library's array keeps the first
call's contents), padded and packed layouts; and the rejection of a batch
whose workspace exceeds the array."""

import ctypes

import numpy as np
import pytest
import torch

from helpers import abs_err, rel_err
from oracle import oracle as O
from orihime_amd import _lib, ragged

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float


def _abi():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    lib.ohm_sw_affine_forward.argtypes = [vp, vp, vp, i32, i32, i32, f32, f32, f32, vp, vp, vp,
                                          vp, vp, vp, ctypes.c_size_t, vp]
    lib.ohm_sw_affine_forward.restype = ctypes.c_int
    lib.ohm_sw_affine_forward_noalloc.argtypes = [vp, vp, vp, i64, i32, i32, i32, f32, f32, f32,
                                                  vp, vp, vp, vp, vp, vp]
    lib.ohm_sw_affine_forward_noalloc.restype = ctypes.c_int
    lib.ohm_sw_affine_workspace.argtypes = [i32, i32, i32, i32, i64]
    lib.ohm_sw_affine_workspace.restype = ctypes.c_size_t
    lib.ohm_scratch_bytes.restype = ctypes.c_size_t
    lib.ohm_last_error.restype = ctypes.c_char_p
    return lib


def _p(t):
    return None if t is None else t.data_ptr()


def _outputs(B, like):
    return (torch.empty(B, device=DEV), torch.empty_like(like),
            *(torch.empty(B, device=DEV) for _ in range(3)))


def _both(lib, scores, lengths, offsets, B, L1, L2, params, total):
    stream = torch.cuda.current_stream().cuda_stream
    ref = _outputs(B, scores)
    need = lib.ohm_sw_affine_workspace(1, B, L1, L2, total if offsets is not None else -1)
    ws = torch.empty(need, dtype=torch.uint8, device=DEV)
    st = lib.ohm_sw_affine_forward(_p(scores), _p(lengths), _p(offsets), B, L1, L2, *params,
                                   *(_p(t) for t in ref), ws.data_ptr(), need, stream)
    assert st == 0, lib.ohm_last_error()
    got = _outputs(B, scores)
    st = lib.ohm_sw_affine_forward_noalloc(_p(scores), _p(lengths), _p(offsets), total, B, L1, L2,
                                           *params, *(_p(t) for t in got), stream)
    assert st == 0, lib.ohm_last_error()
    torch.cuda.synchronize()
    return ref, got


def test_noalloc_matches_workspace_entry_point_padded_twice():
    lib = _abi()
    params = (-8.0, -0.5, 0.5)
    rng = np.random.default_rng(5)
    for B, L1, L2 in ((3, 200, 300), (2, 70, 90)):  # second call smaller
        print(f"[gpu case] noalloc padded {B}x{L1}x{L2}", flush=True)
        s = rng.normal(size=(B, L1, L2)).astype(np.float32)
        ln = np.array([[L1, L2]] + [[L1 - 7, L2 // 2]] * (B - 1), np.int32)
        st, lt = torch.from_numpy(s).to(DEV), torch.from_numpy(ln).to(DEV)
        ref, got = _both(lib, st, lt, None, B, L1, L2, params, -1)
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        oracle = O.sw_affine_forward(s, *params, ln)
        assert rel_err(got[0].cpu().numpy(), oracle[0], 1e-3) < 1e-5
        assert abs_err(got[1].cpu().numpy(), oracle[1]) < 1e-4


def test_noalloc_matches_workspace_entry_point_packed_twice():
    lib = _abi()
    params = (-1.0, -0.3, 1.0)
    rng = np.random.default_rng(6)
    for n, hi in ((24, 400), (9, 120)):  # second call smaller
        pairs = [rng.normal(size=(int(rng.integers(1, hi)), int(rng.integers(1, hi)))).astype(np.float32)
                 for _ in range(n)]
        print(f"[gpu case] noalloc packed {n} pairs up to {hi}", flush=True)
        flat, ln, off = ragged.pack([torch.from_numpy(p).to(DEV) for p in pairs])
        L1 = int(ln[:, 0].max())
        L2 = int(ln[:, 1].max())
        ref, got = _both(lib, flat, ln, off, n, L1, L2, params, flat.numel())
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        r = O.sw_affine_forward(pairs[0][None], *params)
        assert rel_err(got[0][:1].cpu().numpy(), r[0], 1e-3) < 1e-5


def test_noalloc_rejects_a_batch_beyond_its_array():
    lib = _abi()
    cap = lib.ohm_scratch_bytes()
    B, L = 64, 4096
    assert lib.ohm_sw_affine_workspace(1, B, L, L, -1) > cap
    print(f"[gpu case] noalloc rejection {B}x{L}x{L} (host check only)", flush=True)
    out = torch.empty(B, device=DEV)
    st = lib.ohm_sw_affine_forward_noalloc(None, None, None, -1, B, L, L, -1.0, -0.5, 1.0,
                                           out.data_ptr(), None, None, None, None,
                                           torch.cuda.current_stream().cuda_stream)
    assert st == 2  # OHM_EWORKSPACE, before any launch
    assert b"allocation-free" in lib.ohm_last_error()
