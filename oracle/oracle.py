"""ORACLE — test infrastructure only, never part of the product.

ctypes front end for the plain-C restatement in ``oracle/sw_oracle.c`` (FP32
and FP64 builds, ``oracle/libswo.so``) and, where it was built, for the
reference's own CPU kernels in ``oracle/_ref/``.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.

The composites below restate which kernel passes each reference primitive
runs and how it contracts them:

* ``forward``            -> src/sw*/torch_cpu.cpp ``soft_sw*_cpu_with_grads``
                            (forward + backward: value, map, parameter grads)
* ``hvp``                -> ``soft_sw*_hvp_cpu`` (forward + HVP)
* ``param_jacobian``     -> ``soft_sw*_param_jacobian_cpu`` (forward + backward +
                            U/W sensitivity pass)
* ``backward_full``      -> ``soft_sw*_backward_full_cpu`` (forward + backward +
                            HVP + one sensitivity pass per parameter, each
                            contracted with the cotangent and summed)
* ``nw_affine_forward``  -> src/nw_affine/torch_cpu.cpp ``soft_nw_affine_cpu_with_grads``
                            (global alignment: forward + backward)
* ``sv_affine_forward``  -> src/sv_affine/torch_cpu.cpp ``soft_sv_affine_cpu_with_grads``
                            (Saigo-Vert: forward + backward)
* ``nw_forward``, ``sv_forward`` -> the linear-gap twins (src/nw, src/sv_linear)

All arrays are numpy; ``lengths`` defaults to the full shape.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(ref: bool = False) -> None:
    """Compile the restatement (and the reference kernels when ``ref``)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libswo.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def ref_available() -> bool:
    return all(
        os.path.exists(os.path.join(_HERE, "_ref", f"lib{n}_ref.so"))
        for n in ("sw", "sw_affine")
    )


def nw_ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libnw_affine_ref.so"))


def sv_ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libsv_affine_ref.so"))


def linear_ref_available() -> bool:
    return all(os.path.exists(os.path.join(_HERE, "_ref", f"lib{n}_ref.so"))
               for n in ("nw", "sv_linear"))


def _ref_lib(name: str):
    if name not in _REF:
        _REF[name] = ctypes.CDLL(os.path.join(_HERE, "_ref", f"lib{name}_ref.so"))
    return _REF[name]


class _Kernels:
    """Resolve the eight kernels of one implementation and precision."""

    def __init__(self, impl: str, prec: str):
        if impl == "ref":
            if prec != "f32":
                raise ValueError("the reference kernels are FP32 only")
            la, ll = _ref_lib("sw_affine"), _ref_lib("sw")
            self.aff_forward = la.sw_affine_forward_cpu
            self.aff_backward = la.sw_affine_backward_cpu
            self.aff_hvp = la.sw_affine_hvp_cpu
            self.aff_param_grad = la.sw_affine_param_grad_cpu
            self.lin_forward = ll.sw_regular_forward_cpu
            self.lin_backward = ll.sw_regular_backward_cpu
            self.lin_hvp = ll.sw_regular_hvp_cpu
            self.lin_param_grad = ll.sw_regular_param_grad_cpu
        else:
            lib = _lib()
            for name in (
                "aff_forward", "aff_backward", "aff_hvp", "aff_param_grad",
                "lin_forward", "lin_backward", "lin_hvp", "lin_param_grad",
            ):
                setattr(self, name, getattr(lib, f"oracle_{prec}_{name}"))
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.ctype = ctypes.c_float if prec == "f32" else ctypes.c_double


def _kernels(impl: str, prec: str) -> _Kernels:
    return _Kernels(impl, prec)


def _arr(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def _ptr(a):
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    return a.ctypes.data_as(_ip)


def _lens(scores, lengths):
    B, L1, L2 = scores.shape
    if lengths is None:
        lengths = np.tile(np.array([L1, L2], dtype=np.int32), (B, 1))
    return np.ascontiguousarray(lengths, dtype=np.int32)


def _setup(scores, lengths, impl, prec):
    k = _kernels(impl, prec)
    s = _arr(scores, k.dtype)
    ln = _lens(s, lengths)
    B, L1, L2 = s.shape
    return k, s, ln, B, L1, L2


# --------------------------------------------------------------------------
# Affine gap (src/sw_affine/torch_cpu.cpp)
# --------------------------------------------------------------------------

def _aff_fwd_bwd(k, s, ln, B, L1, L2, go, ge, T):
    C = k.ctype
    asz = 3 * (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), k.dtype)
    part = np.zeros(B, k.dtype)
    k.aff_forward(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(go), C(ge), C(T))
    beta = np.zeros((B, asz), k.dtype)
    post = np.zeros((B, L1, L2), k.dtype)
    g = [np.zeros(B, k.dtype) for _ in range(3)]
    k.aff_backward(_ptr(alpha), _ptr(s), _ptr(part), _ptr(beta), _ptr(post),
                   _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(ln), B, L1, L2,
                   C(go), C(ge), C(T))
    return alpha, part, post, g


def sw_affine_forward(scores, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_open[B], grad_ext[B], grad_temp[B])."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    _, part, post, g = _aff_fwd_bwd(k, s, ln, B, L1, L2, gap_open, gap_ext, temp)
    return part, post, g[0], g[1], g[2]


def sw_affine_value(scores, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    """value[B] alone: the forward pass (src/sw_affine/kernels_cpu.cpp:94-195)."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    C = k.ctype
    alpha = np.zeros((B, 3 * (L1 + 1) * (L2 + 1)), k.dtype)
    part = np.zeros(B, k.dtype)
    k.aff_forward(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(gap_open),
                  C(gap_ext), C(temp))
    return part


def over_pairs(fn, pairs, *args, threads=None, chunk=8, impl="oracle", **kw):
    """Run ``fn(scores, *args, lengths=..., impl=..., **kw)`` (one of the
    functions above) over a list of 2-D per-pair score arrays: chunks of
    ``chunk`` pairs, each padded to its own largest pair, on a thread pool
    (the restatement is single-threaded; ctypes releases the GIL).  The
    reference build parallelises over its batch itself, so it runs one chunk
    at a time.  Returns, per output of ``fn``, a list with one entry per
    pair (maps cropped to the pair)."""
    from concurrent.futures import ThreadPoolExecutor

    def run(lo):
        part = pairs[lo:lo + chunk]
        L1 = max(p.shape[0] for p in part)
        L2 = max(p.shape[1] for p in part)
        dt = np.float64 if kw.get("prec") == "f64" else np.float32
        pad = np.zeros((len(part), L1, L2), dt)
        ln = np.zeros((len(part), 2), np.int32)
        for i, p in enumerate(part):
            pad[i, :p.shape[0], :p.shape[1]] = p
            ln[i] = p.shape
        out = fn(pad, *args, lengths=ln, impl=impl, **kw)
        out = out if isinstance(out, tuple) else (out,)
        res = []
        for o in out:
            if o.ndim == 3:
                res.append([o[i, :p.shape[0], :p.shape[1]].copy() for i, p in enumerate(part)])
            else:
                res.append(list(o))
        return res

    starts = range(0, len(pairs), chunk)
    workers = 1 if impl == "ref" else (threads or min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        parts = list(ex.map(run, starts))
    return tuple([x for p in parts for x in p[k]] for k in range(len(parts[0])))


def _aff_hvp_from(k, s, ln, B, L1, L2, alpha, part, V, go, ge, T):
    C = k.ctype
    asz = 3 * (L1 + 1) * (L2 + 1)
    v = _arr(V, k.dtype)
    da = np.zeros((B, asz), k.dtype)
    dp = np.zeros(B, k.dtype)
    db = np.zeros((B, asz), k.dtype)
    H = np.zeros((B, L1, L2), k.dtype)
    k.aff_hvp(_ptr(alpha), _ptr(s), _ptr(part), _ptr(v), _ptr(da), _ptr(dp), _ptr(db),
              _ptr(H), _ptr(ln), B, L1, L2, C(go), C(ge), C(T))
    return H


def sw_affine_hvp(scores, V, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    C = k.ctype
    asz = 3 * (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), k.dtype)
    part = np.zeros(B, k.dtype)
    k.aff_forward(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2,
                  C(gap_open), C(gap_ext), C(temp))
    return _aff_hvp_from(k, s, ln, B, L1, L2, alpha, part, V, gap_open, gap_ext, temp)


def _aff_pj_from(k, s, ln, B, L1, L2, alpha, part, dS, ptype, go, ge, T):
    C = k.ctype
    asz = 3 * (L1 + 1) * (L2 + 1)
    U = np.zeros((B, asz), k.dtype)
    be = np.zeros((B, asz), k.dtype)
    W = np.zeros((B, asz), k.dtype)
    dP = np.zeros((B, L1, L2), k.dtype)
    dS = _arr(dS, k.dtype)
    k.aff_param_grad(_ptr(alpha), _ptr(s), _ptr(part), _ptr(dS), _ptr(U), _ptr(be), _ptr(W),
                     _ptr(dP), _ptr(ln), B, L1, L2, C(go), C(ge), C(T), int(ptype))
    return dP


def sw_affine_param_jacobian(scores, ptype, gap_open, gap_ext, temp, lengths=None,
                             impl="oracle", prec="f32"):
    """d(map)/d(theta), theta = gap_open (0), gap_ext (1), temperature (2)."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    alpha, part, _, g = _aff_fwd_bwd(k, s, ln, B, L1, L2, gap_open, gap_ext, temp)
    return _aff_pj_from(k, s, ln, B, L1, L2, alpha, part, g[ptype], ptype,
                        gap_open, gap_ext, temp)


def sw_affine_backward_full(scores, G, gap_open, gap_ext, temp, lengths=None,
                            impl="oracle", prec="f32"):
    """(grad_scores[B,L1,L2], grad_open, grad_ext, grad_temp) for cotangent G."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    alpha, part, _, g = _aff_fwd_bwd(k, s, ln, B, L1, L2, gap_open, gap_ext, temp)
    Gc = _arr(G, k.dtype)
    H = _aff_hvp_from(k, s, ln, B, L1, L2, alpha, part, Gc, gap_open, gap_ext, temp)
    out = [H]
    for p in range(3):
        dP = _aff_pj_from(k, s, ln, B, L1, L2, alpha, part, g[p], p, gap_open, gap_ext, temp)
        out.append(np.sum(Gc.astype(np.float64) * dP.astype(np.float64)))
    return tuple(out)


# --------------------------------------------------------------------------
# Linear gap (src/sw/torch_cpu.cpp)
# --------------------------------------------------------------------------

def _lin_fwd_bwd(k, s, ln, B, L1, L2, gap, T):
    C = k.ctype
    asz = (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), k.dtype)
    part = np.zeros(B, k.dtype)
    k.lin_forward(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(gap), C(T))
    beta = np.zeros((B, asz), k.dtype)
    post = np.zeros((B, L1, L2), k.dtype)
    gg = np.zeros(B, k.dtype)
    gT = np.zeros(B, k.dtype)
    k.lin_backward(_ptr(alpha), _ptr(s), _ptr(part), _ptr(beta), _ptr(post), _ptr(gg),
                   _ptr(gT), _ptr(ln), B, L1, L2, C(gap), C(T))
    return alpha, part, post, [gg, gT]


def sw_forward(scores, gap, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_gap[B], grad_temp[B])."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    _, part, post, g = _lin_fwd_bwd(k, s, ln, B, L1, L2, gap, temp)
    return part, post, g[0], g[1]


def _lin_hvp_from(k, s, ln, B, L1, L2, alpha, part, V, gap, T):
    C = k.ctype
    asz = (L1 + 1) * (L2 + 1)
    v = _arr(V, k.dtype)
    da = np.zeros((B, asz), k.dtype)
    dp = np.zeros(B, k.dtype)
    be = np.zeros((B, asz), k.dtype)
    db = np.zeros((B, asz), k.dtype)
    H = np.zeros((B, L1, L2), k.dtype)
    k.lin_hvp(_ptr(alpha), _ptr(s), _ptr(part), _ptr(v), _ptr(da), _ptr(dp), _ptr(be), _ptr(db),
              _ptr(H), _ptr(ln), B, L1, L2, C(gap), C(T))
    return H


def sw_hvp(scores, V, gap, temp, lengths=None, impl="oracle", prec="f32"):
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    C = k.ctype
    alpha = np.zeros((B, (L1 + 1) * (L2 + 1)), k.dtype)
    part = np.zeros(B, k.dtype)
    k.lin_forward(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(gap), C(temp))
    return _lin_hvp_from(k, s, ln, B, L1, L2, alpha, part, V, gap, temp)


def _lin_pj_from(k, s, ln, B, L1, L2, alpha, part, dS, ptype, gap, T):
    C = k.ctype
    asz = (L1 + 1) * (L2 + 1)
    U = np.zeros((B, asz), k.dtype)
    be = np.zeros((B, asz), k.dtype)
    W = np.zeros((B, asz), k.dtype)
    dP = np.zeros((B, L1, L2), k.dtype)
    dS = _arr(dS, k.dtype)
    k.lin_param_grad(_ptr(alpha), _ptr(s), _ptr(part), _ptr(dS), _ptr(U), _ptr(be), _ptr(W),
                     _ptr(dP), _ptr(ln), B, L1, L2, C(gap), C(T), int(ptype))
    return dP


def sw_param_jacobian(scores, ptype, gap, temp, lengths=None, impl="oracle", prec="f32"):
    """d(map)/d(theta), theta = gap (0), temperature (1)."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    alpha, part, _, g = _lin_fwd_bwd(k, s, ln, B, L1, L2, gap, temp)
    return _lin_pj_from(k, s, ln, B, L1, L2, alpha, part, g[ptype], ptype, gap, temp)


def sw_backward_full(scores, G, gap, temp, lengths=None, impl="oracle", prec="f32"):
    """(grad_scores[B,L1,L2], grad_gap, grad_temp) for cotangent G."""
    k, s, ln, B, L1, L2 = _setup(scores, lengths, impl, prec)
    alpha, part, _, g = _lin_fwd_bwd(k, s, ln, B, L1, L2, gap, temp)
    Gc = _arr(G, k.dtype)
    H = _lin_hvp_from(k, s, ln, B, L1, L2, alpha, part, Gc, gap, temp)
    out = [H]
    for p in range(2):
        dP = _lin_pj_from(k, s, ln, B, L1, L2, alpha, part, g[p], p, gap, temp)
        out.append(np.sum(Gc.astype(np.float64) * dP.astype(np.float64)))
    return tuple(out)


# --------------------------------------------------------------------------
# Needleman-Wunsch, affine gap (src/nw_affine/torch_cpu.cpp)
# --------------------------------------------------------------------------

def _three_state_forward(family, scores, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    dtype = np.float32 if prec == "f32" else np.float64
    C = ctypes.c_float if prec == "f32" else ctypes.c_double
    if impl == "ref":
        if prec != "f32":
            raise ValueError("the reference kernels are FP32 only")
        lib = _ref_lib(f"{family}_affine")
        fwd = getattr(lib, f"{family}_affine_forward_cpu")
        bwd = getattr(lib, f"{family}_affine_backward_cpu")
    else:
        lib = _lib()
        fwd = getattr(lib, f"oracle_{prec}_{family}a_forward")
        bwd = getattr(lib, f"oracle_{prec}_{family}a_backward")
    s = _arr(scores, dtype)
    ln = _lens(s, lengths)
    B, L1, L2 = s.shape
    asz = 3 * (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), dtype)
    part = np.zeros(B, dtype)
    fwd(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(gap_open), C(gap_ext), C(temp))
    beta = np.zeros((B, asz), dtype)
    post = np.zeros((B, L1, L2), dtype)
    g = [np.zeros(B, dtype) for _ in range(3)]
    bwd(_ptr(alpha), _ptr(s), _ptr(part), _ptr(beta), _ptr(post), _ptr(g[0]), _ptr(g[1]),
        _ptr(g[2]), _ptr(ln), B, L1, L2, C(gap_open), C(gap_ext), C(temp))
    return part, post, g[0], g[1], g[2]


def nw_affine_forward(scores, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_open[B], grad_ext[B], grad_temp[B]) of
    the global (Needleman-Wunsch) affine alignment."""
    return _three_state_forward("nw", scores, gap_open, gap_ext, temp, lengths, impl, prec)


def sv_affine_forward(scores, gap_open, gap_ext, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_open[B], grad_ext[B], grad_temp[B]) of
    the Saigo-Vert affine alignment."""
    return _three_state_forward("sv", scores, gap_open, gap_ext, temp, lengths, impl, prec)


def _linear_forward(family, scores, gap, temp, lengths, impl, prec):
    dtype = np.float32 if prec == "f32" else np.float64
    C = ctypes.c_float if prec == "f32" else ctypes.c_double
    states = 1 if family == "nw" else 3
    if impl == "ref":
        if prec != "f32":
            raise ValueError("the reference kernels are FP32 only")
        name = "nw" if family == "nw" else "sv_linear"
        lib = _ref_lib(name)
        fwd = getattr(lib, f"{name}_forward_cpu")
        bwd = getattr(lib, f"{name}_backward_cpu")
    else:
        lib = _lib()
        fwd = getattr(lib, f"oracle_{prec}_{family}l_forward")
        bwd = getattr(lib, f"oracle_{prec}_{family}l_backward")
    s = _arr(scores, dtype)
    ln = _lens(s, lengths)
    B, L1, L2 = s.shape
    asz = states * (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), dtype)
    part = np.zeros(B, dtype)
    fwd(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B, L1, L2, C(gap), C(temp))
    beta = np.zeros((B, asz), dtype)
    post = np.zeros((B, L1, L2), dtype)
    g = [np.zeros(B, dtype) for _ in range(2)]
    bwd(_ptr(alpha), _ptr(s), _ptr(part), _ptr(beta), _ptr(post), _ptr(g[0]), _ptr(g[1]),
        _ptr(ln), B, L1, L2, C(gap), C(temp))
    return part, post, g[0], g[1]


def nw_forward(scores, gap, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_gap[B], grad_temp[B]) of the linear-gap
    global (Needleman-Wunsch) alignment."""
    return _linear_forward("nw", scores, gap, temp, lengths, impl, prec)


def sv_forward(scores, gap, temp, lengths=None, impl="oracle", prec="f32"):
    """(value[B], map[B,L1,L2], grad_gap[B], grad_temp[B]) of the linear-gap
    Saigo-Vert alignment."""
    return _linear_forward("sv", scores, gap, temp, lengths, impl, prec)


# Second order of the three-state global / Saigo-Vert models: the reference's
# own kernels only (no restatement; the tests pin them, and the library,
# against central differences of the FP64 restatement above).  Composites as
# src/{nw,sv}_affine/torch_cpu.cpp: hvp = forward + HVP pass; the map's
# parameter sensitivity = forward + backward (dS/dtheta) + parameter pass.

def _ref_second_order_setup(family, scores, go, ge, temp, lengths):
    lib = _ref_lib(f"{family}_affine")
    s = _arr(scores, np.float32)
    ln = _lens(s, lengths)
    B, L1, L2 = s.shape
    C = ctypes.c_float
    asz = 3 * (L1 + 1) * (L2 + 1)
    alpha = np.zeros((B, asz), np.float32)
    part = np.zeros(B, np.float32)
    getattr(lib, f"{family}_affine_forward_cpu")(_ptr(s), _ptr(alpha), _ptr(part), _ptr(ln), B,
                                                 L1, L2, C(go), C(ge), C(temp))
    return lib, s, ln, B, L1, L2, alpha, part, asz


def three_state_hvp_ref(family, scores, V, gap_open, gap_ext, temp, lengths=None):
    """H.V of the reference's ``{family}_affine_hvp_cpu`` (family "nw" or "sv")."""
    C = ctypes.c_float
    lib, s, ln, B, L1, L2, alpha, part, asz = _ref_second_order_setup(
        family, scores, gap_open, gap_ext, temp, lengths)
    v = _arr(V, np.float32)
    ws = [np.zeros((B, asz), np.float32) for _ in range(3)]
    dpart = np.zeros(B, np.float32)
    H = np.zeros((B, L1, L2), np.float32)
    if family == "nw":
        lib.nw_affine_hvp_cpu(_ptr(alpha), _ptr(s), _ptr(part), _ptr(v), _ptr(ws[0]), _ptr(dpart),
                              _ptr(ws[1]), _ptr(ws[2]), _ptr(H), _ptr(ln), B, L1, L2,
                              C(gap_open), C(gap_ext), C(temp))
    else:
        lib.sv_affine_hvp_cpu(_ptr(alpha), _ptr(s), _ptr(part), _ptr(v), _ptr(ws[0]), _ptr(dpart),
                              _ptr(ws[1]), _ptr(H), _ptr(ln), B, L1, L2, C(gap_open), C(gap_ext),
                              C(temp))
    return H


def three_state_param_jacobian_ref(family, scores, ptype, gap_open, gap_ext, temp, lengths=None):
    """d(map)/d(theta) of the reference's ``{family}_affine_param_grad_cpu``
    (ptype 0 gap_open, 1 gap_ext, 2 temperature)."""
    C = ctypes.c_float
    lib, s, ln, B, L1, L2, alpha, part, asz = _ref_second_order_setup(
        family, scores, gap_open, gap_ext, temp, lengths)
    beta = np.zeros((B, asz), np.float32)
    post = np.zeros((B, L1, L2), np.float32)
    g = [np.zeros(B, np.float32) for _ in range(3)]
    getattr(lib, f"{family}_affine_backward_cpu")(
        _ptr(alpha), _ptr(s), _ptr(part), _ptr(beta), _ptr(post), _ptr(g[0]), _ptr(g[1]),
        _ptr(g[2]), _ptr(ln), B, L1, L2, C(gap_open), C(gap_ext), C(temp))
    ws = [np.zeros((B, asz), np.float32) for _ in range(3)]
    dP = np.zeros((B, L1, L2), np.float32)
    getattr(lib, f"{family}_affine_param_grad_cpu")(
        _ptr(alpha), _ptr(s), _ptr(part), _ptr(g[ptype]), _ptr(ws[0]), _ptr(ws[1]), _ptr(ws[2]),
        _ptr(dP), _ptr(ln), B, L1, L2, C(gap_open), C(gap_ext), C(temp), int(ptype))
    return dP
