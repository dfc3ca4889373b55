"""Argument checks shared by every alignment family.

What is checked follows the reference engine (orihime/operator.py:159-331:
scalar parameter shapes, the temperature, derivative vectors, the FP32
numerical domain |x| / T <= 80 with -inf as the mask marker) and so do the
messages that callers match on; the mask sentinel follows
orihime/operator.py:334-386.

Checks on device tensors reduce on the device and read back one bool, so an
invalid input raises an ordinary Python exception (never a device assert).
While the tensor's stream is being captured into a HIP graph nothing may be
read back, so those checks are skipped (``deferred``): the kernels clamp
lengths into [0, max] and never address outside the inputs, and a value out
of the numerical domain gives wrong numbers, not a fault.
"""

from __future__ import annotations

import math

import torch
from torch import Tensor

Param = float | Tensor

#: Largest |score or gap| / temperature the FP32 kernels accept.
DOMAIN_LIMIT = 80.0
#: Finite stand-in some callers write for an excluded cell; accepted as is.
LEGACY_MASK_VALUE = -1.0e4
#: Extra temperature units a masked cell sits below every reachable path.
MASK_MARGIN = 200.0

_DOMAIN_TEXT = ("orihime's supported FP32 numerical domain requires every finite "
                f"|score or cost| / temperature to be <= {DOMAIN_LIMIT:g}")


def _plain(t: Tensor) -> Tensor:
    """Strip functorch wrappers (vmap / grad levels) from a bool condition."""
    fx = torch._C._functorch
    while fx.is_batchedtensor(t) or fx.is_gradtrackingtensor(t):
        t = fx.get_unwrapped(t)
    return t


def deferred(t) -> bool:
    """True while ``t`` is a device tensor whose current stream is being
    captured into a HIP graph, where reading device data back is forbidden."""
    return isinstance(t, Tensor) and t.is_cuda and torch.cuda.is_current_stream_capturing()


def host_scalar(t: Tensor, name: str) -> float:
    """A one-element tensor parameter as a Python float (the kernels take
    scoring parameters by value)."""
    if deferred(t):
        raise RuntimeError(f"{name} is a device tensor, read on the host before the launch; "
                           "pass it as a Python number while capturing a HIP graph")
    return float(t.detach().reshape(-1)[0].item())


def require(condition: Tensor, message: str) -> None:
    """RuntimeError(message) unless every element of the bool ``condition``
    holds; skipped while capturing (``deferred``)."""
    if not isinstance(condition, Tensor) or condition.dtype is not torch.bool:
        raise TypeError("internal validation condition must be a bool tensor")
    if deferred(condition):
        return
    if not bool(_plain(condition).all().cpu()):
        raise RuntimeError(message)


def scalar_shape(value: Param, name: str) -> None:
    """A tensor parameter holds one number: shape [] or [1]."""
    if isinstance(value, Tensor) and value.numel() == 1 and value.ndim <= 1:
        return
    if isinstance(value, Tensor):
        raise ValueError(
            f"{name} must be a scalar tensor with shape [] or [1], got shape {tuple(value.shape)}")


def temperature(value: Param) -> None:
    """Finite and strictly positive, as a number or a one-element tensor."""
    scalar_shape(value, "temperature")
    text = "temperature must be finite and strictly positive"
    if isinstance(value, Tensor):
        if not value.is_floating_point():
            raise TypeError("temperature tensor must have a floating-point dtype")
        require(torch.isfinite(value) & (value > 0), text)
        return
    try:
        t = float(value)
    except (TypeError, ValueError) as exc:
        raise TypeError("temperature must be a real number or tensor") from exc
    if not (math.isfinite(t) and t > 0.0):
        raise ValueError(f"{text}, got {t}")


def temperature_like(value: Param, like: Tensor) -> float | Tensor:
    """The temperature as a float, or as a tensor on ``like``'s device/dtype."""
    if isinstance(value, Tensor):
        return value.to(device=like.device, dtype=like.dtype)
    return float(value)


def derivative_vector(primary: Tensor, vector: Tensor, *, name: str, primary_name: str,
                      error_type: type[Exception] | None = RuntimeError) -> Tensor:
    """A tangent or cotangent: FP32, contiguous, with ``primary``'s shape and
    device.  Every failure raises ``error_type`` when one is given, else the
    natural kind (TypeError for kind/dtype, ValueError for the rest)."""
    if not isinstance(primary, Tensor):
        raise TypeError(f"{primary_name} must be a tensor")
    problems = (
        (lambda: not isinstance(vector, Tensor), TypeError,
         lambda: f"{name} must be a tensor"),
        (lambda: vector.shape != primary.shape, ValueError,
         lambda: (f"{name} must have same shape as {primary_name}; got {tuple(vector.shape)} "
                  f"and {tuple(primary.shape)}")),
        (lambda: vector.dtype != torch.float32, TypeError,
         lambda: f"{name} must have dtype torch.float32, got {vector.dtype}"),
        (lambda: vector.device != primary.device, ValueError,
         lambda: (f"{name} must be on same device as {primary_name}; got {vector.device} "
                  f"and {primary.device}")),
        (lambda: not vector.is_contiguous(), ValueError, lambda: f"{name} must be contiguous"),
    )
    for failed, kind, text in problems:
        if failed():
            raise (error_type or kind)(text())
    return vector


def _param_in_domain(name: str, p: Param, temp: Param) -> None:
    scalar_shape(p, name)
    bad = f"{_DOMAIN_TEXT}; {name} is out of domain"
    if isinstance(p, Tensor):
        if not p.is_floating_point():
            raise TypeError(f"{name} tensor must have a floating-point dtype")
        require(torch.isfinite(p) & (p.abs() <= DOMAIN_LIMIT * temperature_like(temp, p)), bad)
        return
    g = float(p)
    if not math.isfinite(g):
        raise ValueError(f"{_DOMAIN_TEXT}; {name} must be finite")
    if isinstance(temp, Tensor):
        require(DOMAIN_LIMIT * temp >= abs(g), bad)
    elif abs(g) > DOMAIN_LIMIT * float(temp):
        raise ValueError(bad)


def numerical_domain(input_name: str, scores: Tensor, names: tuple[str, ...],
                     params: tuple[Param, ...]) -> None:
    """Every finite score and gap is within DOMAIN_LIMIT temperatures of 0;
    -inf (and the legacy finite marker) flags a masked cell; +inf and NaN are
    rejected.  ``names`` are the internal parameter names, "temp" among them."""
    temp = params[names.index("temp")]
    temperature(temp)
    if scores.is_floating_point() and not deferred(scores):
        require(~torch.isposinf(scores),
                f"{input_name} contains an unsupported infinity; use -inf score masks")
        bound = DOMAIN_LIMIT * temperature_like(temp, scores)
        inside = (scores.abs() <= bound) | torch.isneginf(scores) | (scores == LEGACY_MASK_VALUE)
        require(inside & ~torch.isnan(scores), f"{_DOMAIN_TEXT}; {input_name} is out of domain")
    for name, p in zip(names, params):
        if name != "temp":
            _param_in_domain(name, p, temp)


def _sentinel_from(scores: Tensor, top: Tensor, t: float) -> Tensor:
    span = sum(scores.shape[1:]) * DOMAIN_LIMIT + MASK_MARGIN
    top = torch.where(torch.isfinite(top), top, torch.zeros_like(top))
    return torch.where(torch.isneginf(scores), (top - span * t).to(scores.dtype), scores)


def mask_to_sentinel(scores: Tensor, temp: Param) -> Tensor:
    """Replace -inf cells by a finite score so low that no alignment through
    it can compete: the largest retained score minus (L1 + L2) * 80 + 200
    temperatures (the most any path can gain, plus a margin).  The maximum
    and the temperature are taken without gradient."""
    t = temperature_like(temp, scores)
    if isinstance(t, Tensor):
        t = t.detach()
    top = scores.amax(dim=tuple(range(1, scores.ndim)), keepdim=True).detach()
    return _sentinel_from(scores, top, t)


# --------------------------------------------------------------------------
# Device fast path: the checks above in one native scan
# --------------------------------------------------------------------------

_RANGE_OK = "_ohm_range_ok"


def lengths_checked(lengths: Tensor, L1: int, L2: int) -> bool:
    """True when ``scan`` already checked these lengths against (L1, L2) and
    they have not been modified in place since."""
    return getattr(lengths, _RANGE_OK, None) == (L1, L2, lengths._version)


def scan_applies(scores: Tensor, temp: Param, lengths: Tensor | None) -> bool:
    """Whether ``scan`` can stand in for the torch checks: a plain contiguous
    FP32 device tensor [B, L1, L2], a Python-number temperature, int32
    lengths on the same device, outside graph capture."""
    fx = torch._C._functorch
    return (isinstance(scores, Tensor) and scores.is_cuda and scores.dim() == 3
            and scores.dtype == torch.float32 and scores.is_contiguous()
            and not fx.is_batchedtensor(scores) and not fx.is_gradtrackingtensor(scores)
            and not isinstance(temp, Tensor) and not deferred(scores)
            and (lengths is None or (lengths.is_cuda and lengths.dtype == torch.int32
                                     and lengths.is_contiguous())))


def scan(input_name: str, scores: Tensor, names: tuple[str, ...], params: tuple[Param, ...],
         lengths: Tensor | None, length_error: str) -> Tensor:
    """``numerical_domain`` + ``mask_to_sentinel`` + the lengths range check
    for a device tensor (``scan_applies``) in one native pass over the
    scores and one read-back: the same errors in the same order (lengths,
    temperature, +inf, scores domain, parameters), and the same normalised
    scores.  Without -inf cells the scores are returned as they are."""
    from . import _lib

    temp = params[names.index("temp")]
    temperature(temp)
    t = float(temp)
    B = scores.shape[0]
    counts = torch.empty(4, dtype=torch.int32, device=scores.device)
    pair_max = torch.empty(B, dtype=torch.float32, device=scores.device)
    # the bound in FP32, as a tensor comparison against the Python float rounds it
    bound = float(torch.tensor(DOMAIN_LIMIT * t, dtype=torch.float32))
    _lib.scan_inputs(scores.detach(), lengths, bound, LEGACY_MASK_VALUE, counts, pair_max)
    pinf, bad, ninf, bad_len = (int(v) for v in counts.cpu().tolist())
    if bad_len:
        raise RuntimeError(length_error)
    if lengths is not None:
        setattr(lengths, _RANGE_OK, (scores.shape[1], scores.shape[2], lengths._version))
    if pinf:
        raise RuntimeError(f"{input_name} contains an unsupported infinity; use -inf score masks")
    if bad:
        raise RuntimeError(f"{_DOMAIN_TEXT}; {input_name} is out of domain")
    for name, p in zip(names, params):
        if name != "temp":
            _param_in_domain(name, p, temp)
    if not ninf:
        return scores
    return _sentinel_from(scores, pair_max.view(B, 1, 1), t)


__all__ = ["DOMAIN_LIMIT", "deferred", "derivative_vector", "host_scalar", "lengths_checked",
           "mask_to_sentinel", "numerical_domain", "require", "scalar_shape", "scan",
           "scan_applies", "temperature", "temperature_like"]
