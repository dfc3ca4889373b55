"""Write the golden fixtures in tests/golden/*.npz.

Every expected output comes from the reference's own CPU kernels
(src/sw/kernels_cpu.cpp, src/sw_affine/kernels_cpu.cpp,
src/nw_affine/kernels_cpu.cpp, src/sv_affine/kernels_cpu.cpp, src/nw/kernels_cpu.cpp,
src/sv_linear/kernels_cpu.cpp) compiled from their
sources into oracle/_ref/ by ``make -C oracle ref``; the composites follow
src/sw*/torch_cpu.cpp (which passes each primitive runs).  The fixtures hold
inputs and expected outputs only.  Regenerate with

    make -C oracle ref && python tests/golden/make_golden.py

Inputs use one RNG convention: ``torch.Generator().manual_seed(seed)`` and
``torch.randn(shape, generator=g, dtype=torch.float32)``, drawn in the order
scores, then V (HVP direction), then G (map cotangent).
"""

from __future__ import annotations

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def _draw(seed, shape, scale=1.0, shift=0.0, extra=2):
    g = torch.Generator().manual_seed(seed)
    s = torch.randn(shape, generator=g, dtype=torch.float32) * scale + shift
    rest = [torch.randn(shape, generator=g, dtype=torch.float32) for _ in range(extra)]
    return [s.numpy()] + [r.numpy() for r in rest]


def _sentinel(scores, lengths, T):
    """Padded-and-masked input: cells beyond each pair's lengths get the
    finite exclusion sentinel amax - ((L1+L2)*80 + 200)*T of
    orihime/operator.py:334-386 (amax over the retained cells)."""
    B, L1, L2 = scores.shape
    out = scores.copy()
    for b in range(B):
        mask = np.ones((L1, L2), bool)
        mask[: lengths[b, 0], : lengths[b, 1]] = False
        kept = scores[b][~mask]
        amax = kept.max() if kept.size else 0.0
        out[b][mask] = np.float32(amax - ((L1 + L2) * 80.0 + 200.0) * T)
    return out


def affine_case(name, s, V, G, go, ge, T, lengths=None, full=True):
    B, L1, L2 = s.shape
    if lengths is None:
        lengths = np.tile(np.array([L1, L2], np.int32), (B, 1))
    d = dict(kind="affine", scores=s, lengths=lengths, gap_open=np.float32(go),
             gap_ext=np.float32(ge), temp=np.float32(T))
    v, m, go_, ge_, gt = O.sw_affine_forward(s, go, ge, T, lengths, impl="ref")
    d.update(value=v, map=m, grad_open=go_, grad_ext=ge_, grad_temp=gt)
    if full:
        d["V"] = V
        d["hvp"] = O.sw_affine_hvp(s, V, go, ge, T, lengths, impl="ref")
        for p, n in enumerate(("gap_open", "gap_ext", "temp")):
            d[f"sens_{n}"] = O.sw_affine_param_jacobian(s, p, go, ge, T, lengths, impl="ref")
        d["G"] = G
        gs, a, b, c = O.sw_affine_backward_full(s, G, go, ge, T, lengths, impl="ref")
        d.update(mb_scores=gs, mb_gap_open=np.float64(a), mb_gap_ext=np.float64(b),
                 mb_temp=np.float64(c))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
    return d


def linear_case(name, s, V, G, gap, T, lengths=None, full=True):
    B, L1, L2 = s.shape
    if lengths is None:
        lengths = np.tile(np.array([L1, L2], np.int32), (B, 1))
    d = dict(kind="linear", scores=s, lengths=lengths, gap=np.float32(gap), temp=np.float32(T))
    v, m, gg, gt = O.sw_forward(s, gap, T, lengths, impl="ref")
    d.update(value=v, map=m, grad_gap=gg, grad_temp=gt)
    if full:
        d["V"] = V
        d["hvp"] = O.sw_hvp(s, V, gap, T, lengths, impl="ref")
        for p, n in enumerate(("gap", "temp")):
            d[f"sens_{n}"] = O.sw_param_jacobian(s, p, gap, T, lengths, impl="ref")
        d["G"] = G
        gs, a, b = O.sw_backward_full(s, G, gap, T, lengths, impl="ref")
        d.update(mb_scores=gs, mb_gap=np.float64(a), mb_temp=np.float64(b))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
    return d


def nw_affine_case(name, s, go, ge, T, lengths=None, family="nw"):
    B, L1, L2 = s.shape
    if lengths is None:
        lengths = np.tile(np.array([L1, L2], np.int32), (B, 1))
    d = dict(kind=f"{family}_affine", scores=s, lengths=lengths, gap_open=np.float32(go),
             gap_ext=np.float32(ge), temp=np.float32(T))
    fwd = O.nw_affine_forward if family == "nw" else O.sv_affine_forward
    v, m, go_, ge_, gt = fwd(s, go, ge, T, lengths, impl="ref")
    d.update(value=v, map=m, grad_open=go_, grad_ext=ge_, grad_temp=gt)
    if B * L1 * L2 <= 8192:
        # second order: the reference's HVP along V (N(0,1), seed 7 + B) and
        # its parameter sensitivities of the map
        V = torch.randn((B, L1, L2), generator=torch.Generator().manual_seed(7 + B),
                        dtype=torch.float32).numpy()
        d["V"] = V
        d["hvp"] = O.three_state_hvp_ref(family, s, V, go, ge, T, lengths)
        for p, n in enumerate(("gap_open", "gap_ext", "temp")):
            d[f"sens_{n}"] = O.three_state_param_jacobian_ref(family, s, p, go, ge, T, lengths)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
    return d


def linear_gl_case(name, s, gap, T, lengths=None, family="nw"):
    B, L1, L2 = s.shape
    if lengths is None:
        lengths = np.tile(np.array([L1, L2], np.int32), (B, 1))
    d = dict(kind=f"{family}_linear", scores=s, lengths=lengths, gap=np.float32(gap),
             temp=np.float32(T))
    fwd = O.nw_forward if family == "nw" else O.sv_forward
    v, m, gg, gt = fwd(s, gap, T, lengths, impl="ref")
    d.update(value=v, map=m, grad_gap=gg, grad_temp=gt)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **d)
    return d


def main():
    if not O.ref_available():
        raise SystemExit("build the reference kernels first: make -C oracle ref")
    # Linear gap, B=8, 128x128, N(0,1) seed 0, gap=-1, T=1: value, map and
    # the parameter gradients at full size; second-order outputs on a crop.
    s, V, G = _draw(0, (8, 128, 128))
    linear_case("linear_b8_128x128_seed0", s, V, G, -1.0, 1.0, full=False)
    linear_case("linear_b2_32x32_seed0_crop", s[:2, :32, :32].copy(), V[:2, :32, :32].copy(),
                G[:2, :32, :32].copy(), -1.0, 1.0)
    # Affine, N(0,1)*2-0.5 seed 1, go=-11, ge=-1, T=1 (crop of B=256 512x512).
    s, V, G = _draw(1, (2, 48, 40), 2.0, -0.5)
    affine_case("affine_b2_48x40_seed1", s, V, G, -11.0, -1.0, 1.0)
    # Ragged affine, go=-8, ge=-0.5, T=0.5: lengths form and padded+sentinel form.
    s, V, G = _draw(2, (4, 40, 36))
    ln = np.array([[40, 36], [17, 29], [33, 6], [1, 23]], np.int32)
    affine_case("affine_ragged_b4_40x36_seed2", s, V, G, -8.0, -0.5, 0.5, ln)
    affine_case("affine_padmask_b4_40x36_seed2", _sentinel(s, ln, 0.5), V, G, -8.0, -0.5, 0.5,
                full=False)
    # Long-sequence crop, N(0,1) seed 3, go=-11, ge=-1, T=1.
    s, V, G = _draw(3, (1, 64, 64))
    affine_case("affine_b1_64x64_seed3", s, V, G, -11.0, -1.0, 1.0)
    # Reference test-suite style edge cases (tests/test_soft_sw_affine.py:295-348).
    edge = [
        ("1x1", (1, 1, 1), -1.0, -0.3, 1.0),
        ("1x3", (1, 1, 3), -1.0, -0.3, 1.0),
        ("3x1", (1, 3, 1), -1.0, -0.3, 1.0),
        ("lowT", (2, 8, 10), -1.0, -0.3, 0.01),
        ("highT", (2, 8, 10), -1.0, -0.3, 10.0),
        ("gaps_m1m1", (2, 5, 20), -1.0, -1.0, 1.0),
        ("gaps_m2_0", (2, 16, 16), -2.0, 0.0, 1.0),
        ("gaps_pos", (2, 8, 10), 0.5, 0.2, 1.0),
    ]
    for i, (tag, shape, go, ge, T) in enumerate(edge):
        s, V, G = _draw(40 + i, shape)
        affine_case(f"affine_edge_{tag}", s, V, G, go, ge, T)
        linear_case(f"linear_edge_{tag}", s, V, G, go, T)
    # Zero lengths and boundary-heavy scores (test_soft_sw_affine.py:220-264).
    s, V, G = _draw(60, (3, 9, 11))
    s[:, 0, :] -= 2.5
    ln = np.array([[0, 11], [9, 0], [4, 7]], np.int32)
    affine_case("affine_zero_lengths", s, V, G, -1.0, -0.3, 0.9, ln)
    linear_case("linear_zero_lengths", s, V, G, -0.7, 0.9, ln)
    # Needleman-Wunsch affine (global): the reference test grid's shapes and
    # gaps (tests/test_soft_nw_affine.py:143-160), a multi-strip pair, ragged
    # lengths with empty sides, boundary-heavy scores (:248-258).
    if not O.nw_ref_available():
        raise SystemExit("build the reference kernels first: make -C oracle ref")
    for i, (shape, go, ge, T) in enumerate([((4, 8, 10), -2.0, -0.5, 1.0),
                                            ((4, 16, 16), -1.0, -0.3, 0.1),
                                            ((4, 5, 20), -3.0, -1.0, 2.0)]):
        s = _draw(70 + i, shape, extra=0)[0]
        nw_affine_case(f"nw_affine_b{shape[0]}_{shape[1]}x{shape[2]}", s, go, ge, T)
    s = _draw(80, (2, 150, 97), extra=0)[0]
    nw_affine_case("nw_affine_b2_150x97", s, -2.0, -0.5, 0.5)
    s = _draw(81, (5, 40, 36), extra=0)[0]
    ln = np.array([[40, 36], [17, 29], [0, 6], [33, 0], [1, 1]], np.int32)
    nw_affine_case("nw_affine_ragged_b5_40x36", s, -2.0, -0.5, 1.0, ln)
    s = _draw(7, (3, 5, 7), extra=0)[0]
    s[:, 0, :] -= 2.5
    s[:, :, 0] -= 2.5
    nw_affine_case("nw_affine_boundary_b3_5x7", s, -2.0, -0.5, 1.0)
    for i, (tag, shape, go, ge, T) in enumerate([("1x1", (1, 1, 1), -2.0, -0.5, 1.0),
                                                 ("1x3", (1, 1, 3), -2.0, -0.5, 1.0),
                                                 ("3x1", (1, 3, 1), -2.0, -0.5, 1.0),
                                                 ("gaps_pos", (2, 8, 10), 0.5, 0.2, 1.0)]):
        s = _draw(90 + i, shape, extra=0)[0]
        nw_affine_case(f"nw_affine_edge_{tag}", s, go, ge, T)
    # Saigo-Vert affine: the same shapes and edge cases (values are LSE(0, {M})).
    if not O.sv_ref_available():
        raise SystemExit("build the reference kernels first: make -C oracle ref")
    for i, (shape, go, ge, T) in enumerate([((4, 8, 10), -2.0, -0.5, 1.0),
                                            ((4, 16, 16), -1.0, -0.3, 0.1),
                                            ((4, 5, 20), -3.0, -1.0, 2.0)]):
        s = _draw(100 + i, shape, extra=0)[0]
        nw_affine_case(f"sv_affine_b{shape[0]}_{shape[1]}x{shape[2]}", s, go, ge, T, family="sv")
    s = _draw(110, (2, 150, 97), extra=0)[0]
    nw_affine_case("sv_affine_b2_150x97", s, -2.0, -0.5, 0.5, family="sv")
    s = _draw(111, (5, 40, 36), extra=0)[0]
    ln = np.array([[40, 36], [17, 29], [0, 6], [33, 0], [1, 1]], np.int32)
    nw_affine_case("sv_affine_ragged_b5_40x36", s, -8.0, -0.5, 0.5, ln, family="sv")
    for i, (tag, shape, go, ge, T) in enumerate([("1x1", (1, 1, 1), -2.0, -0.5, 1.0),
                                                 ("3x1", (1, 3, 1), -2.0, -0.5, 1.0),
                                                 ("gaps_pos", (2, 8, 10), 0.5, 0.2, 1.0)]):
        s = _draw(120 + i, shape, extra=0)[0]
        nw_affine_case(f"sv_affine_edge_{tag}", s, go, ge, T, family="sv")
    # Linear-gap Needleman-Wunsch and Saigo-Vert (src/nw, src/sv_linear).
    if not O.linear_ref_available():
        raise SystemExit("build the reference kernels first: make -C oracle ref")
    for fam in ("nw", "sv"):
        for i, (shape, gap, T) in enumerate([((4, 8, 10), -1.0, 1.0), ((4, 16, 16), -0.7, 0.2),
                                             ((2, 120, 77), -1.5, 0.5)]):
            s = _draw(130 + i, shape, extra=0)[0]
            linear_gl_case(f"{fam}_linear_b{shape[0]}_{shape[1]}x{shape[2]}", s, gap, T,
                           family=fam)
        s = _draw(140, (5, 40, 36), extra=0)[0]
        ln = np.array([[40, 36], [17, 29], [0, 6], [33, 0], [1, 1]], np.int32)
        linear_gl_case(f"{fam}_linear_ragged_b5_40x36", s, -1.0, 0.7, ln, family=fam)
    # Empty dimensions in the tensor's shape (not through lengths): the
    # reference accepts (B, 0, N) and (B, M, 0) and returns the empty
    # alignment (0; for global alignment the gap run over the other side).
    for shape in ((2, 0, 5), (2, 5, 0)):
        tag = "x".join(str(n) for n in shape)
        z = np.zeros(shape, np.float32)
        affine_case(f"affine_empty_{tag}", z, z, z, -2.0, -0.5, 1.0)
        linear_case(f"linear_empty_{tag}", z, z, z, -1.0, 1.0)
        nw_affine_case(f"nw_affine_empty_{tag}", z, -2.0, -0.5, 1.0)
        nw_affine_case(f"sv_affine_empty_{tag}", z, -2.0, -0.5, 1.0, family="sv")
        for fam in ("nw", "sv"):
            linear_gl_case(f"{fam}_linear_empty_{tag}", z, -1.0, 1.0, family=fam)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
