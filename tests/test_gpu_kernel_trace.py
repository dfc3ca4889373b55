"""Kernel-trace pruning contract on the GPU (reference:
tests/test_soft_sw_affine.py:1088-1115 with its helper :865-895).

A backward through the value alone must launch no second-order kernel; a
backward through the map must launch the HVP sweeps and the parameter
derivative sweeps.  The kernels are told apart by name: the HVP passes are
named sw_affine_hvp_*, the parameter-derivative passes sw_affine_param_grad_*
(including the reverse sweep of the map VJP, which contracts the parameter
derivatives), the first-order sweeps neither.
"""

from __future__ import annotations

import pytest
import torch

import orihime_amd

sw_affine_ops = orihime_amd.ops._kernels["sw_affine"]

pytestmark = pytest.mark.gpu

GAP_OPEN, GAP_EXT, TEMPERATURE = -1.0, -0.3, 0.7


def _kernel_names(call):
    from torch.profiler import ProfilerActivity, profile

    torch.cuda.synchronize()
    with profile(activities=[ProfilerActivity.CPU, ProfilerActivity.CUDA]) as prof:
        call()
        torch.cuda.synchronize()
    return {getattr(e, "name", "") or getattr(e, "key", "") for e in prof.events()}


def _scores(seed):
    g = torch.Generator().manual_seed(seed)
    return torch.randn((1, 8, 9), generator=g).cuda().requires_grad_(True)


def test_score_backward_skips_second_order_kernels():
    print("[case] kernel trace: value-only backward (1, 8, 9)", flush=True)

    def score_only_backward():
        scores = _scores(776)
        score, _ = sw_affine_ops.forward(scores, GAP_OPEN, GAP_EXT, TEMPERATURE, None)
        score.sum().backward()

    names = _kernel_names(score_only_backward)
    sweeps = [n for n in names if "sw_affine_" in n]
    assert any("sw_affine_scaled_" in n for n in sweeps), sorted(sweeps)[:20]
    assert not any("sw_affine_hvp_" in n for n in names), sorted(sweeps)
    assert not any("sw_affine_param_grad_" in n for n in names), sorted(sweeps)


def test_map_backward_runs_hvp_and_param_grad_kernels():
    print("[case] kernel trace: map backward (1, 8, 9)", flush=True)

    def explicit_zero_map_backward():
        scores = _scores(777)
        score, posteriors = sw_affine_ops.forward(scores, GAP_OPEN, GAP_EXT, TEMPERATURE, None)
        (score.sum() + 0.0 * posteriors.sum()).backward()

    names = _kernel_names(explicit_zero_map_backward)
    sweeps = sorted(n for n in names if "sw_affine_" in n)
    assert any("sw_affine_hvp_" in n for n in names), sweeps
    assert any("sw_affine_param_grad_" in n for n in names), sweeps


def test_parameter_sensitivity_runs_param_grad_kernels_only():
    print("[case] kernel trace: map sensitivity to gap_open (1, 8, 9)", flush=True)
    scores = _scores(778).detach()
    names = _kernel_names(
        lambda: sw_affine_ops.marginals_grad_gap_open(scores, GAP_OPEN, GAP_EXT, TEMPERATURE))
    sweeps = sorted(n for n in names if "sw_affine_" in n)
    assert any("sw_affine_param_grad_scaled_" in n for n in names), sweeps
    assert not any("sw_affine_hvp_scaled_" in n for n in names), sweeps
