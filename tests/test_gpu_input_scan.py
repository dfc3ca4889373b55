"""The native input scan (ohm_scan_inputs) that stands in, on the GPU, for
the drop-in engine's torch checks: the numerical domain and mask marker
(orihime/operator.py:249-386 of the reference), the mask sentinel's per-pair
maximum and the lengths range (src/sw_affine/torch_hip.cpp:69-99).

Each case compares the scan path with the torch path on the same input: the
same error (message and order) or the same normalised scores, bit for bit.
CPU tests cover the parts that need no device: when the scan applies, and
the in-place-modification guard on the lengths mark.
"""

from __future__ import annotations

import pytest
import torch

import orihime_amd as ohm
from orihime_amd import _checks, _lib

DEV = torch.device("cuda:0")
NAMES = ("gap_open", "gap_ext", "temp")
LEN_MSG = "lengths entries must lie within the padded input shape"


def _torch_path(scores, params, lengths):
    """The engine's checks without the scan, in the engine's order."""
    if lengths is not None:
        ok = (lengths[:, 0] >= 0) & (lengths[:, 0] <= scores.shape[1]) & \
             (lengths[:, 1] >= 0) & (lengths[:, 1] <= scores.shape[2])
        _checks.require(ok, LEN_MSG)
    _checks.numerical_domain("scores", scores, NAMES, params)
    return _checks.mask_to_sentinel(scores, params[2])


def _both(scores, params, lengths=None):
    """(result or exception text) of the scan path and of the torch path."""
    out = []
    for fn in (lambda: _checks.scan("scores", scores, NAMES, params, lengths, LEN_MSG),
               lambda: _torch_path(scores, params, lengths)):
        try:
            out.append(fn())
        except (RuntimeError, ValueError) as exc:
            out.append(f"{type(exc).__name__}: {exc}")
    return out


# --------------------------------------------------------------------------
# CPU
# --------------------------------------------------------------------------


def test_scan_applies_only_to_plain_device_tensors():
    x = torch.zeros(2, 3, 4)
    assert not _checks.scan_applies(x, 1.0, None)  # CPU tensor


def test_lengths_mark_expires_on_in_place_change():
    ln = torch.tensor([[3, 4], [1, 1]], dtype=torch.int32)
    setattr(ln, _checks._RANGE_OK, (3, 4, ln._version))
    assert _checks.lengths_checked(ln, 3, 4)
    assert not _checks.lengths_checked(ln, 3, 5)  # another shape
    ln[0, 0] = 9
    assert not _checks.lengths_checked(ln, 3, 4)


# --------------------------------------------------------------------------
# GPU
# --------------------------------------------------------------------------


def _scores(seed, shape=(3, 70, 90), scale=3.0):
    g = torch.Generator().manual_seed(seed)
    return (torch.randn(shape, generator=g) * scale).to(DEV)


@pytest.mark.gpu
def test_scan_counts_and_pair_maxima():
    print("[gpu case] ohm_scan_inputs counts and maxima, planted cells", flush=True)
    x = _scores(0, (4, 33, 1000), scale=1.0)
    x[0, 0, 0] = float("inf")
    x[1, 5, 7] = float("nan")
    x[1, 6, 7] = 80.5          # above the bound 80
    x[2, 3, 3] = -80.0         # on the bound: inside
    x[2, 4, 4] = -1.0e4        # legacy marker: inside
    x[3, :, 500:] = float("-inf")
    x[3, 2, 2] = 7.25          # pair 3's maximum
    ln = torch.tensor([[33, 1000], [34, 10], [-1, 0], [0, 1001]], dtype=torch.int32, device=DEV)
    counts = torch.empty(4, dtype=torch.int32, device=DEV)
    pmax = torch.empty(4, device=DEV)
    _lib.scan_inputs(x, ln, 80.0, -1.0e4, counts, pmax)
    torch.cuda.synchronize()
    assert counts.tolist() == [1, 2, 33 * 500, 3]
    inf = float("inf")
    want = x.cpu().nan_to_num(nan=-inf, posinf=inf, neginf=-inf).amax(dim=(1, 2))
    got = pmax.cpu()
    assert torch.equal(got, want)  # NaN cells aside; pair 0's +inf is its maximum
    assert got[0] == inf and got[3] == 7.25


@pytest.mark.gpu
def test_scan_of_empty_boxes():
    print("[gpu case] ohm_scan_inputs on (2, 0, 5)", flush=True)
    x = torch.empty(2, 0, 5, device=DEV)
    counts = torch.full((4,), 7, dtype=torch.int32, device=DEV)
    pmax = torch.zeros(2, device=DEV)
    _lib.scan_inputs(x, None, 80.0, -1.0e4, counts, pmax)
    torch.cuda.synchronize()
    assert counts.tolist() == [0, 0, 0, 0]
    assert (pmax.cpu() == float("-inf")).all()


CASES = {
    "clean": lambda x: None,
    "masked_rows": lambda x: x[1, 40:, :].fill_(float("-inf")),
    "fully_masked_pair": lambda x: x[2].fill_(float("-inf")),
    "legacy_marker": lambda x: x[0, :5, :5].fill_(-1.0e4),
    "plus_inf": lambda x: x[0, 1, 1].fill_(float("inf")),
    "nan": lambda x: x[2, 1, 1].fill_(float("nan")),
    "out_of_domain": lambda x: x[1, 2, 3].fill_(2000.0),
    "inf_and_nan": lambda x: (x[0, 1, 1].fill_(float("nan")), x[1, 1, 1].fill_(float("inf"))),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("lengths_ok", [True, False])
def test_scan_matches_torch_checks(case, lengths_ok):
    print(f"[gpu case] scan vs torch checks: {case}, lengths_ok={lengths_ok}", flush=True)
    x = _scores(1)
    CASES[case](x)
    ln = torch.tensor([[70, 90], [50, 2], [1, 90 if lengths_ok else 91]], dtype=torch.int32,
                      device=DEV)
    params = (-3.0, -0.5, 5.0)
    fast, ref = _both(x, params, ln)
    if isinstance(ref, str):
        assert fast == ref
    else:
        assert torch.equal(fast, ref)
        if case == "clean":
            assert fast is x  # nothing to replace: no copy


@pytest.mark.gpu
def test_scan_parameter_domain_and_temperature_errors():
    print("[gpu case] scan: parameter and temperature errors", flush=True)
    x = _scores(2, scale=1.0)
    for params in ((-500.0, -0.5, 1.0), (-3.0, float("inf"), 1.0), (-3.0, -0.5, 0.0),
                   (-3.0, -0.5, -1.0)):
        fast, ref = _both(x, params)
        assert isinstance(ref, str) and fast == ref, params
    # a scores error comes before a parameter error
    x[0, 0, 0] = 1e6
    fast, ref = _both(x, (-500.0, -0.5, 1.0))
    assert fast == ref and "scores is out of domain" in ref


@pytest.mark.gpu
def test_public_calls_equal_with_and_without_scan(monkeypatch):
    print("[gpu case] public sw_affine / nw_affine with masks: scan vs torch path", flush=True)
    x = _scores(3, (4, 130, 70), scale=1.0)
    x[1, 100:, :] = float("-inf")
    x[3, :, 60:] = float("-inf")
    ln = torch.tensor([[130, 70], [129, 70], [17, 3], [130, 65]], dtype=torch.int32, device=DEV)
    kw = dict(gap_open_score=-2.0, gap_extend_score=-0.5, temperature=0.7, lengths=ln)
    outs = []
    for use_scan in (True, False):
        if not use_scan:
            monkeypatch.setattr(_checks, "scan_applies", lambda *a: False)
        xr = x.clone().requires_grad_(True)
        v = ohm.sw_affine_value(xr, **kw)
        v.sum().backward()
        outs.append((v.detach(), xr.grad, ohm.sw_affine(x, **kw), ohm.nw_affine_value(x, **kw)))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.gpu
def test_schema_recheck_after_in_place_lengths_change():
    print("[gpu case] lengths mark: in-place change re-checked by the schema", flush=True)
    x = _scores(4, (2, 40, 50), scale=1.0)
    ln = torch.tensor([[40, 50], [10, 10]], dtype=torch.int32, device=DEV)
    ohm.sw_affine_value(x, gap_open_score=-2.0, gap_extend_score=-0.5, temperature=1.0, lengths=ln)
    assert _checks.lengths_checked(ln, 40, 50)
    ln[1, 0] = 41
    with pytest.raises(RuntimeError, match=r"lengths\[1,0\] must be between 0 and 40, got 41"):
        torch.ops.orihime.soft_sw_affine(x, torch.tensor(-2.0, device=DEV),
                                         torch.tensor(-0.5, device=DEV),
                                         torch.tensor(1.0, device=DEV), ln)
    torch.cuda.synchronize()
