"""ctypes binding of libohm_sw.so (C ABI in include/ohm_sw.h).

This is the only module that touches the native library.  It raises
ImportError when the library is missing — there is no Python fallback — and
passes raw data pointers, sizes and the current HIP stream to the C entry
points.  Tensors on a GPU go to the HIP kernels, CPU tensors to the host
implementation in the same library.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libohm_sw.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"orihime_amd native library not found at {LIB_PATH}; build it with "
        "`python orihime_amd/build_native.py` (requires ROCm hipcc)"
    )

_lib = ctypes.CDLL(LIB_PATH)

_f = ctypes.c_float
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

EXPORTS = (
    "ohm_last_error",
    "ohm_version",
    "ohm_set_stage_events",
    "ohm_set_launch_config",
    "ohm_set_checkpoint_stride",
    "ohm_checkpoint_stride",
    "ohm_scan_inputs",
    "ohm_embed_offsets",
    "ohm_embed_sw_affine_forward",
    "ohm_embed_sw_affine_forward_cpu",
    "ohm_sw_affine_workspace",
    "ohm_sw_workspace",
    "ohm_sw_affine_forward",
    "ohm_sw_affine_tangent",
    "ohm_sw_forward",
    "ohm_sw_tangent",
    "ohm_sw_affine_forward_cpu",
    "ohm_sw_affine_tangent_cpu",
    "ohm_sw_forward_cpu",
    "ohm_sw_tangent_cpu",
    "ohm_nw_affine_workspace",
    "ohm_nw_affine_forward",
    "ohm_nw_affine_forward_cpu",
    "ohm_sv_affine_workspace",
    "ohm_sv_affine_forward",
    "ohm_sv_affine_forward_cpu",
    "ohm_nw_workspace",
    "ohm_sv_workspace",
    "ohm_nw_forward",
    "ohm_sv_forward",
    "ohm_nw_forward_cpu",
    "ohm_sv_forward_cpu",
    "ohm_nw_affine_tangent",
    "ohm_nw_affine_tangent_cpu",
    "ohm_sv_affine_tangent",
    "ohm_sv_affine_tangent_cpu",
    "ohm_scratch_bytes",
    "ohm_sw_affine_forward_noalloc",
    "ohm_embed_scores",
    "ohm_embed_scores_cpu",
    "ohm_embed_grads_workspace",
    "ohm_embed_grads",
    "ohm_embed_grads_cpu",
)


def _sig(name, restype, *argtypes):
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_last_error = _sig("ohm_last_error", ctypes.c_char_p)
_version = _sig("ohm_version", ctypes.c_char_p)
_stage_events = _sig("ohm_set_stage_events", ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), _i32)
_launch_config = _sig("ohm_set_launch_config", ctypes.c_int, _i32, _i32, _i32)
_set_ckpt = _sig("ohm_set_checkpoint_stride", ctypes.c_int, _i32)
_scan = _sig("ohm_scan_inputs", ctypes.c_int, _vp, _vp, _i32, _i32, _i32, ctypes.c_float,
             ctypes.c_float, _vp, _vp, _vp)
_get_ckpt = _sig("ohm_checkpoint_stride", _i32)
_ws_affine = _sig("ohm_sw_affine_workspace", _sz, _i32, _i32, _i32, _i32, _i64)
_ws_linear = _sig("ohm_sw_workspace", _sz, _i32, _i32, _i32, _i32, _i64)
_ws_nw_affine = _sig("ohm_nw_affine_workspace", _sz, _i32, _i32, _i32, _i32, _i64)
_ws_sv_affine = _sig("ohm_sv_affine_workspace", _sz, _i32, _i32, _i32, _i32, _i64)
_ws_nw = _sig("ohm_nw_workspace", _sz, _i32, _i32, _i32, _i32, _i64)
_ws_sv = _sig("ohm_sv_workspace", _sz, _i32, _i32, _i32, _i32, _i64)

_PROBLEM = (_vp, _vp, _vp, _i32, _i32, _i32)  # scores, lengths, offsets, B, L1max, L2max
_aff_fwd = _sig("ohm_sw_affine_forward", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_aff_tan = _sig("ohm_sw_affine_tangent", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_lin_fwd = _sig("ohm_sw_forward", _i32, *_PROBLEM, _f, _f,
                _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_lin_tan = _sig("ohm_sw_tangent", _i32, *_PROBLEM, _f, _f,
                _vp, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_aff_fwd_cpu = _sig("ohm_sw_affine_forward_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _vp, _vp, _vp, _vp, _i32)
_aff_tan_cpu = _sig("ohm_sw_affine_tangent_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _i32)
_lin_fwd_cpu = _sig("ohm_sw_forward_cpu", _i32, *_PROBLEM, _f, _f,
                    _vp, _vp, _vp, _vp, _i32)
_lin_tan_cpu = _sig("ohm_sw_tangent_cpu", _i32, *_PROBLEM, _f, _f,
                    _vp, _f, _f, _vp, _vp, _vp, _vp, _i32)
_nwa_fwd = _sig("ohm_nw_affine_forward", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_nwa_fwd_cpu = _sig("ohm_nw_affine_forward_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _vp, _vp, _vp, _vp, _i32)
_sva_fwd = _sig("ohm_sv_affine_forward", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_sva_fwd_cpu = _sig("ohm_sv_affine_forward_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _vp, _vp, _vp, _vp, _i32)
_nwa_tan = _sig("ohm_nw_affine_tangent", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_nwa_tan_cpu = _sig("ohm_nw_affine_tangent_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _i32)
_sva_tan = _sig("ohm_sv_affine_tangent", _i32, *_PROBLEM, _f, _f, _f,
                _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_sva_tan_cpu = _sig("ohm_sv_affine_tangent_cpu", _i32, *_PROBLEM, _f, _f, _f,
                    _vp, _f, _f, _f, _vp, _vp, _vp, _vp, _i32)
_nwl_fwd = _sig("ohm_nw_forward", _i32, *_PROBLEM, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_svl_fwd = _sig("ohm_sv_forward", _i32, *_PROBLEM, _f, _f, _vp, _vp, _vp, _vp, _vp, _sz, _vp)
_nwl_fwd_cpu = _sig("ohm_nw_forward_cpu", _i32, *_PROBLEM, _f, _f, _vp, _vp, _vp, _vp, _i32)
_svl_fwd_cpu = _sig("ohm_sv_forward_cpu", _i32, *_PROBLEM, _f, _f, _vp, _vp, _vp, _vp, _i32)

OP_VALUE, OP_FORWARD, OP_TANGENT = 0, 1, 2


def version() -> str:
    return _version().decode()


def _check(status: int, what: str) -> None:
    if status != 0:
        msg = _last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


# --------------------------------------------------------------------------
# Device workspace: one growing byte buffer per (device, stream).
# --------------------------------------------------------------------------

_ws_lock = threading.Lock()
_ws_cache: dict[tuple[int, int], torch.Tensor] = {}


def set_stage_events(events) -> None:
    """Record the given hipEvent_t handles (ints, up to 5; None skips one)
    after the planning launch, the forward sweep, its log-domain launch, the
    reverse sweep and its log-domain launch of every GPU call; an empty
    sequence clears the hook."""
    arr = (ctypes.c_void_p * 5)(*[(e or None) for e in events])
    _check(_stage_events(arr, len(events)), "ohm_set_stage_events")


def set_launch_config(waves: int = 0, groups_per_cu: int = 0, teams: int = 0) -> None:
    """Launch shape of the GPU kernels (0 = automatic; see include/ohm_sw.h)."""
    _check(_launch_config(waves, groups_per_cu, teams), "ohm_set_launch_config")


def set_checkpoint_stride(steps: int = 0) -> None:
    """Steps between the affine forward sweep's checkpoints (4, 8 or 16; 0 =
    automatic).  Results do not depend on it; workspace sizes do."""
    _check(_set_ckpt(steps), "ohm_set_checkpoint_stride")


def checkpoint_stride() -> int:
    return int(_get_ckpt())


def scan_inputs(scores: torch.Tensor, lengths, bound: float, marker: float,
                counts: torch.Tensor, pair_max: torch.Tensor) -> None:
    """Device scan of padded scores [B, L1, L2] and lengths (see
    include/ohm_sw.h, ohm_scan_inputs): counts [4] int32, pair_max [B]."""
    B, L1, L2 = scores.shape
    stream = torch.cuda.current_stream(scores.device).cuda_stream
    _check(_scan(_ptr(scores), _ptr(lengths), B, L1, L2, float(bound), float(marker),
                 _ptr(counts), _ptr(pair_max), stream), "ohm_scan_inputs")


def workspace_bytes(affine: bool, op: int, B: int, L1: int, L2: int, total_cells: int = -1) -> int:
    fn = _ws_affine if affine else _ws_linear
    return int(fn(op, B, L1, L2, total_cells))


def _workspace(device: torch.device, nbytes: int, stream: int) -> torch.Tensor:
    key = (device.index if device.index is not None else torch.cuda.current_device(), stream)
    with _ws_lock:
        buf = _ws_cache.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=device)
            _ws_cache[key] = buf
        return buf


def release_workspaces() -> None:
    with _ws_lock:
        _ws_cache.clear()


def _cpu_threads() -> int:
    return int(torch.get_num_threads())


# --------------------------------------------------------------------------
# Entry points.  All tensors must already be validated, contiguous FP32 (int32
# lengths, int64 offsets) on one device; outputs are allocated by the caller.
# --------------------------------------------------------------------------


def _dims(scores: torch.Tensor, lengths, offsets, L1max, L2max):
    if offsets is None:
        B, L1, L2 = scores.shape
        return B, L1, L2
    return int(offsets.shape[0]), int(L1max), int(L2max)


def affine_forward(scores, lengths, offsets, gap_open, gap_ext, temp, *, value, map=None,
                   grad_open=None, grad_ext=None, grad_temp=None, L1max=None, L2max=None,
                   workspace=None):
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2,
            float(gap_open), float(gap_ext), float(temp),
            _ptr(value), _ptr(map), _ptr(grad_open), _ptr(grad_ext), _ptr(grad_temp))
    if scores.is_cuda:
        op = OP_FORWARD if (map is not None or grad_open is not None or grad_ext is not None
                            or grad_temp is not None) else OP_VALUE
        total = -1 if offsets is None else scores.numel()
        need = workspace_bytes(True, op, B, L1, L2, total)
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_aff_fwd(*args, ws.data_ptr(), ws.numel(), stream), "ohm_sw_affine_forward")
    else:
        _check(_aff_fwd_cpu(*args, _cpu_threads()), "ohm_sw_affine_forward_cpu")


def affine_tangent(scores, lengths, offsets, gap_open, gap_ext, temp, *, direction=None,
                   d_gap_open=0.0, d_gap_ext=0.0, d_temp=0.0, map_tangent=None, contract=None,
                   value=None, map=None, L1max=None, L2max=None, workspace=None):
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2,
            float(gap_open), float(gap_ext), float(temp), _ptr(direction),
            float(d_gap_open), float(d_gap_ext), float(d_temp),
            _ptr(map_tangent), _ptr(contract), _ptr(value), _ptr(map))
    if scores.is_cuda:
        total = -1 if offsets is None else scores.numel()
        need = workspace_bytes(True, OP_TANGENT, B, L1, L2, total)
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_aff_tan(*args, ws.data_ptr(), ws.numel(), stream), "ohm_sw_affine_tangent")
    else:
        _check(_aff_tan_cpu(*args, _cpu_threads()), "ohm_sw_affine_tangent_cpu")


def linear_forward(scores, lengths, offsets, gap, temp, *, value, map=None, grad_gap=None,
                   grad_temp=None, L1max=None, L2max=None, workspace=None):
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2, float(gap), float(temp),
            _ptr(value), _ptr(map), _ptr(grad_gap), _ptr(grad_temp))
    if scores.is_cuda:
        op = OP_FORWARD if (map is not None or grad_gap is not None or grad_temp is not None) \
            else OP_VALUE
        total = -1 if offsets is None else scores.numel()
        need = workspace_bytes(False, op, B, L1, L2, total)
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_lin_fwd(*args, ws.data_ptr(), ws.numel(), stream), "ohm_sw_forward")
    else:
        _check(_lin_fwd_cpu(*args, _cpu_threads()), "ohm_sw_forward_cpu")


def linear_tangent(scores, lengths, offsets, gap, temp, *, direction=None, d_gap=0.0,
                   d_temp=0.0, map_tangent=None, contract=None, value=None, map=None,
                   L1max=None, L2max=None, workspace=None):
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2, float(gap), float(temp),
            _ptr(direction), float(d_gap), float(d_temp), _ptr(map_tangent), _ptr(contract),
            _ptr(value), _ptr(map))
    if scores.is_cuda:
        total = -1 if offsets is None else scores.numel()
        need = workspace_bytes(False, OP_TANGENT, B, L1, L2, total)
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_lin_tan(*args, ws.data_ptr(), ws.numel(), stream), "ohm_sw_tangent")
    else:
        _check(_lin_tan_cpu(*args, _cpu_threads()), "ohm_sw_tangent_cpu")


def nw_affine_forward(scores, lengths, offsets, gap_open, gap_ext, temp, *, value, map=None,
                      grad_open=None, grad_ext=None, grad_temp=None, L1max=None, L2max=None,
                      workspace=None):
    """Needleman-Wunsch affine: value, and map + parameter gradients when
    ``map`` is given."""
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2,
            float(gap_open), float(gap_ext), float(temp),
            _ptr(value), _ptr(map), _ptr(grad_open), _ptr(grad_ext), _ptr(grad_temp))
    if scores.is_cuda:
        op = OP_FORWARD if map is not None else OP_VALUE
        total = -1 if offsets is None else scores.numel()
        need = int(_ws_nw_affine(op, B, L1, L2, total))
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_nwa_fwd(*args, ws.data_ptr(), ws.numel(), stream), "ohm_nw_affine_forward")
    else:
        _check(_nwa_fwd_cpu(*args, _cpu_threads()), "ohm_nw_affine_forward_cpu")


def sv_affine_forward(scores, lengths, offsets, gap_open, gap_ext, temp, *, value, map=None,
                      grad_open=None, grad_ext=None, grad_temp=None, L1max=None, L2max=None,
                      workspace=None):
    """Saigo-Vert affine: value, and map + parameter gradients when ``map`` is
    given."""
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2,
            float(gap_open), float(gap_ext), float(temp),
            _ptr(value), _ptr(map), _ptr(grad_open), _ptr(grad_ext), _ptr(grad_temp))
    if scores.is_cuda:
        op = OP_FORWARD if map is not None else OP_VALUE
        total = -1 if offsets is None else scores.numel()
        need = int(_ws_sv_affine(op, B, L1, L2, total))
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(_sva_fwd(*args, ws.data_ptr(), ws.numel(), stream), "ohm_sv_affine_forward")
    else:
        _check(_sva_fwd_cpu(*args, _cpu_threads()), "ohm_sv_affine_forward_cpu")


def _linear_global_local(family):
    gpu, cpu, ws_fn = {"nw": (_nwl_fwd, _nwl_fwd_cpu, _ws_nw),
                       "sv": (_svl_fwd, _svl_fwd_cpu, _ws_sv)}[family]

    def run(scores, lengths, offsets, gap, temp, *, value, map=None, grad_gap=None,
            grad_temp=None, L1max=None, L2max=None, workspace=None):
        B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
        args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2, float(gap), float(temp),
                _ptr(value), _ptr(map), _ptr(grad_gap), _ptr(grad_temp))
        if scores.is_cuda:
            op = OP_FORWARD if map is not None else OP_VALUE
            total = -1 if offsets is None else scores.numel()
            need = int(ws_fn(op, B, L1, L2, total))
            stream = torch.cuda.current_stream(scores.device).cuda_stream
            ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
            _check(gpu(*args, ws.data_ptr(), ws.numel(), stream), f"ohm_{family}_forward")
        else:
            _check(cpu(*args, _cpu_threads()), f"ohm_{family}_forward_cpu")

    run.__doc__ = (f"Linear-gap {'Needleman-Wunsch' if family == 'nw' else 'Saigo-Vert'}: "
                   "value, and map + dS/dgap, dS/dT when ``map`` is given.")
    return run


nw_forward = _linear_global_local("nw")
sv_forward = _linear_global_local("sv")


def _three_state_tangent(gpu, cpu, ws_fn, name, scores, lengths, offsets, gap_open, gap_ext, temp, *, direction=None,
                      d_gap_open=0.0, d_gap_ext=0.0, d_temp=0.0, map_tangent=None,
                      contract=None, value=None, map=None, L1max=None, L2max=None,
                      workspace=None):
    B, L1, L2 = _dims(scores, lengths, offsets, L1max, L2max)
    args = (_ptr(scores), _ptr(lengths), _ptr(offsets), B, L1, L2,
            float(gap_open), float(gap_ext), float(temp), _ptr(direction),
            float(d_gap_open), float(d_gap_ext), float(d_temp),
            _ptr(map_tangent), _ptr(contract), _ptr(value), _ptr(map))
    if scores.is_cuda:
        total = -1 if offsets is None else scores.numel()
        need = int(ws_fn(OP_TANGENT, B, L1, L2, total))
        stream = torch.cuda.current_stream(scores.device).cuda_stream
        ws = workspace if workspace is not None else _workspace(scores.device, need, stream)
        _check(gpu(*args, ws.data_ptr(), ws.numel(), stream), name)
    else:
        _check(cpu(*args, _cpu_threads()), name + "_cpu")


def nw_affine_tangent(*args, **kw):
    """Needleman-Wunsch affine forward-over-reverse pass (arguments as
    ``affine_tangent``)."""
    return _three_state_tangent(_nwa_tan, _nwa_tan_cpu, _ws_nw_affine, "ohm_nw_affine_tangent",
                                *args, **kw)


def sv_affine_tangent(*args, **kw):
    """Saigo-Vert affine forward-over-reverse pass (arguments as
    ``affine_tangent``)."""
    return _three_state_tangent(_sva_tan, _sva_tan_cpu, _ws_sv_affine, "ohm_sv_affine_tangent",
                                *args, **kw)


# --------------------------------------------------------------------------
# Embedding-scored alignment (include/ohm_embed.h)
# --------------------------------------------------------------------------

_emb_scores = _sig("ohm_embed_scores", ctypes.c_int, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _vp,
                   _vp, _vp)
_emb_scores_cpu = _sig("ohm_embed_scores_cpu", ctypes.c_int, _vp, _vp, _vp, _i32, _i32, _i32, _i32,
                       _vp, _vp)
_emb_offsets = _sig("ohm_embed_offsets", ctypes.c_int, _vp, _i32, _i32, _i32, _vp, _vp)
_emb_sw = _sig("ohm_embed_sw_affine_forward", ctypes.c_int, _vp, _vp, _vp, _i32, _i32, _i32, _i32,
               ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
               _vp, _sz, _vp)
_emb_sw_cpu = _sig("ohm_embed_sw_affine_forward_cpu", ctypes.c_int, _vp, _vp, _vp, _i32, _i32, _i32,
                   _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp, _vp, _vp,
                   _vp, _vp, _i32)
_emb_ws = _sig("ohm_embed_grads_workspace", _sz, _i32, _i32, _i32, _i32)
_emb_grads = _sig("ohm_embed_grads", ctypes.c_int, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _vp,
                  _vp, _vp, _vp, _vp, _sz, _vp)
_emb_grads_cpu = _sig("ohm_embed_grads_cpu", ctypes.c_int, _vp, _vp, _vp, _vp, _i32, _i32, _i32,
                      _i32, _vp, _vp, _vp, _vp)


def embed_scores(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor, offsets: torch.Tensor,
                 scores: torch.Tensor) -> None:
    """Packed ``Q_b K_b^T`` into ``scores`` (capacity B * Lq * Lk floats) and
    the packed offsets.  Q [B, Lq, D], K [B, Lk, D] BF16, contiguous."""
    B, Lq, D = Q.shape
    Lk = K.shape[1]
    args = (_ptr(Q), _ptr(K), _ptr(lengths), B, Lq, Lk, D, _ptr(offsets), _ptr(scores))
    if Q.is_cuda:
        stream = torch.cuda.current_stream(Q.device).cuda_stream
        _check(_emb_scores(*args, stream), "ohm_embed_scores")
    else:
        _check(_emb_scores_cpu(*args), "ohm_embed_scores_cpu")


def embed_sw_affine(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor, gap_open: float,
                    gap_ext: float, temp: float, *, offsets: torch.Tensor, scores: torch.Tensor,
                    value: torch.Tensor, map=None, grad_open=None, grad_ext=None, grad_temp=None,
                    workspace: torch.Tensor | None = None) -> None:
    """Affine soft-SW of every pair scored by ``Q_b K_b^T``, the scores
    generated inside the sweeps (include/ohm_embed.h,
    ohm_embed_sw_affine_forward).  ``scores`` (capacity B * Lq * Lk) is
    written only for pairs the log-domain sweeps redo; ``offsets`` is
    written."""
    B, Lq, D = Q.shape
    Lk = K.shape[1]
    args = (_ptr(Q), _ptr(K), _ptr(lengths), B, Lq, Lk, D, float(gap_open), float(gap_ext),
            float(temp), _ptr(offsets), _ptr(scores), _ptr(value), _ptr(map), _ptr(grad_open),
            _ptr(grad_ext), _ptr(grad_temp))
    if Q.is_cuda:
        stream = torch.cuda.current_stream(Q.device).cuda_stream
        need = workspace_bytes(True, OP_FORWARD, B, Lq, Lk, B * Lq * Lk)
        ws = workspace if workspace is not None else _workspace(Q.device, need, stream)
        _check(_emb_sw(*args, _ptr(ws), ws.numel(), stream), "ohm_embed_sw_affine_forward")
    else:
        _check(_emb_sw_cpu(*args, _cpu_threads()), "ohm_embed_sw_affine_forward_cpu")


def embed_grads(Q: torch.Tensor, K: torch.Tensor, lengths: torch.Tensor, offsets: torch.Tensor,
                map: torch.Tensor, grad_value: torch.Tensor, dQ: torch.Tensor, dK: torch.Tensor,
                workspace: torch.Tensor | None = None) -> None:
    """dQ, dK (BF16, like Q and K) of sum_b g_b S_b through the packed map."""
    B, Lq, D = Q.shape
    Lk = K.shape[1]
    args = (_ptr(Q), _ptr(K), _ptr(lengths), _ptr(offsets), B, Lq, Lk, D, _ptr(map),
            _ptr(grad_value), _ptr(dQ), _ptr(dK))
    if Q.is_cuda:
        stream = torch.cuda.current_stream(Q.device).cuda_stream
        need = int(_emb_ws(B, Lq, Lk, D))
        ws = workspace if workspace is not None else _workspace(Q.device, need, stream)
        _check(_emb_grads(*args, _ptr(ws), ws.numel(), stream), "ohm_embed_grads")
    else:
        _check(_emb_grads_cpu(*args), "ohm_embed_grads_cpu")
