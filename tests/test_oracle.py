"""Pins the oracle (oracle/sw_oracle.c) before anything is checked against it.

* bit-for-bit against every golden fixture (written by the reference's own
  CPU kernels compiled from source, tests/golden/make_golden.py);
* bit-for-bit against oracle/_ref on fresh random inputs, when that build
  is present (this container only);
* the FP64 build against brute-force enumeration of every alignment path
  on tiny grids (no dynamic programming) and against closed forms;
* derivatives against central finite differences of the FP64 oracle.
"""

import math

import numpy as np
import pytest

from helpers import stable_seed, abs_err, affine_params, brute_affine, brute_linear, golden_files, load
from oracle import oracle as O

AFFINE = golden_files("affine")
LINEAR = golden_files("linear")


@pytest.mark.parametrize("path", AFFINE, ids=lambda p: p.split("/")[-1])
def test_affine_oracle_matches_golden_bitwise(path):
    d = load(path)
    go, ge, T = affine_params(d)
    s, ln = d["scores"], d["lengths"]
    v, m, a, b, c = O.sw_affine_forward(s, go, ge, T, ln)
    for name, got in zip(("value", "map", "grad_open", "grad_ext", "grad_temp"), (v, m, a, b, c)):
        assert np.array_equal(got, d[name]), name
    if "hvp" in d:
        assert np.array_equal(O.sw_affine_hvp(s, d["V"], go, ge, T, ln), d["hvp"])
        for p, n in enumerate(("gap_open", "gap_ext", "temp")):
            assert np.array_equal(O.sw_affine_param_jacobian(s, p, go, ge, T, ln), d[f"sens_{n}"])
        gs, x, y, z = O.sw_affine_backward_full(s, d["G"], go, ge, T, ln)
        assert np.array_equal(gs, d["mb_scores"])
        assert (x, y, z) == (float(d["mb_gap_open"]), float(d["mb_gap_ext"]), float(d["mb_temp"]))


@pytest.mark.parametrize("path", LINEAR, ids=lambda p: p.split("/")[-1])
def test_linear_oracle_matches_golden_bitwise(path):
    d = load(path)
    g, T = float(d["gap"]), float(d["temp"])
    s, ln = d["scores"], d["lengths"]
    v, m, a, b = O.sw_forward(s, g, T, ln)
    for name, got in zip(("value", "map", "grad_gap", "grad_temp"), (v, m, a, b)):
        assert np.array_equal(got, d[name]), name
    if "hvp" in d:
        assert np.array_equal(O.sw_hvp(s, d["V"], g, T, ln), d["hvp"])
        for p, n in enumerate(("gap", "temp")):
            assert np.array_equal(O.sw_param_jacobian(s, p, g, T, ln), d[f"sens_{n}"])
        gs, x, y = O.sw_backward_full(s, d["G"], g, T, ln)
        assert np.array_equal(gs, d["mb_scores"])
        assert (x, y) == (float(d["mb_gap"]), float(d["mb_temp"]))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (make -C oracle ref)")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(100 + seed)
    B, L1, L2 = 3, int(rng.integers(1, 20)), int(rng.integers(1, 20))
    s = (rng.normal(size=(B, L1, L2)) * 2 - 0.5).astype(np.float32)
    V = rng.normal(size=s.shape).astype(np.float32)
    ln = np.stack([rng.integers(0, L1 + 1, B), rng.integers(0, L2 + 1, B)], 1).astype(np.int32)
    go, ge, T = float(-rng.uniform(0, 3)), float(-rng.uniform(0, 1)), float(rng.uniform(0.2, 3))
    for x, y in zip(O.sw_affine_forward(s, go, ge, T, ln),
                    O.sw_affine_forward(s, go, ge, T, ln, impl="ref")):
        assert np.array_equal(x, y)
    assert np.array_equal(O.sw_affine_hvp(s, V, go, ge, T, ln),
                          O.sw_affine_hvp(s, V, go, ge, T, ln, impl="ref"))
    for p in range(3):
        assert np.array_equal(O.sw_affine_param_jacobian(s, p, go, ge, T, ln),
                              O.sw_affine_param_jacobian(s, p, go, ge, T, ln, impl="ref"))
    for x, y in zip(O.sw_forward(s, go, T, ln), O.sw_forward(s, go, T, ln, impl="ref")):
        assert np.array_equal(x, y)
    assert np.array_equal(O.sw_hvp(s, V, go, T, ln), O.sw_hvp(s, V, go, T, ln, impl="ref"))
    for p in range(2):
        assert np.array_equal(O.sw_param_jacobian(s, p, go, T, ln),
                              O.sw_param_jacobian(s, p, go, T, ln, impl="ref"))


@pytest.mark.parametrize("shape", [(1, 1), (1, 3), (3, 1), (2, 2), (2, 3), (3, 3), (3, 4)])
@pytest.mark.parametrize("T", [0.5, 1.0, 3.0])
def test_fp64_oracle_matches_path_enumeration(shape, T):
    rng = np.random.default_rng(stable_seed(shape, T))
    s = rng.normal(size=(1, *shape))
    go, ge, g = -0.9, -0.4, -0.7
    v = O.sw_affine_forward(s, go, ge, T, prec="f64")[0][0]
    assert abs(v - brute_affine(s[0], go, ge, T)) < 1e-12
    v = O.sw_forward(s, g, T, prec="f64")[0][0]
    assert abs(v - brute_linear(s[0], g, T)) < 1e-12


def test_one_by_one_closed_form():
    for sv in (-3.0, -0.2, 0.0, 1.7):
        for T in (0.1, 1.0, 4.0):
            s = np.array([[[sv]]])
            v, m, *_ = O.sw_affine_forward(s, -1.0, -0.5, T, prec="f64")
            assert abs(v[0] - T * math.log1p(math.exp(sv / T))) < 1e-12
            assert abs(m[0, 0, 0] - 1.0 / (1.0 + math.exp(-sv / T))) < 1e-12
            v, m, *_ = O.sw_forward(s, -1.0, T, prec="f64")
            assert abs(v[0] - T * math.log1p(math.exp(sv / T))) < 1e-12


def test_zero_length_pairs_are_the_empty_alignment():
    s = np.random.default_rng(0).normal(size=(2, 4, 5))
    ln = np.array([[0, 5], [4, 0]], np.int32)
    v, m, go_, ge_, gt = O.sw_affine_forward(s, -1.0, -0.3, 0.7, ln, prec="f64")
    assert np.all(v == 0) and np.all(m == 0) and np.all(gt == 0)


def _fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


@pytest.mark.parametrize("T", [0.5, 1.3])
def test_fp64_oracle_derivatives_match_finite_differences(T):
    rng = np.random.default_rng(7)
    s = rng.normal(size=(2, 5, 6))
    s[:, 0, :] -= 2.5  # boundary-heavy scores, as in the reference tests
    V = rng.normal(size=s.shape)
    go, ge = -1.1, -0.4
    h = 1e-5
    val = lambda s_, go_=go, ge_=ge, T_=T: O.sw_affine_forward(s_, go_, ge_, T_, prec="f64")[0]
    mp = lambda s_, go_=go, ge_=ge, T_=T: O.sw_affine_forward(s_, go_, ge_, T_, prec="f64")[1]
    v, m, gO, gE, gT = O.sw_affine_forward(s, go, ge, T, prec="f64")
    # map = dS/ds
    fdm = np.zeros_like(s)
    for idx in np.ndindex(*s.shape):
        e = np.zeros_like(s)
        e[idx] = 1
        fdm[idx] = _fd(lambda t: val(s + t * e)[idx[0]], 0.0, h)
    assert abs_err(m, fdm) < 1e-7
    assert abs_err(gO, _fd(lambda t: val(s, go_=t), go, h)) < 1e-7
    assert abs_err(gE, _fd(lambda t: val(s, ge_=t), ge, h)) < 1e-7
    assert abs_err(gT, _fd(lambda t: val(s, T_=t), T, h)) < 1e-6  # entropy = dS/dT
    # HVP = directional derivative of the map; sensitivities = dmap/dtheta
    H = O.sw_affine_hvp(s, V, go, ge, T, prec="f64")
    assert abs_err(H, _fd(lambda t: mp(s + t * V), 0.0, h)) < 1e-6
    for p, kw in enumerate(("go_", "ge_", "T_")):
        base = (go, ge, T)[p]
        fd = _fd(lambda t: mp(s, **{kw: t}), base, h)
        assert abs_err(O.sw_affine_param_jacobian(s, p, go, ge, T, prec="f64"), fd) < 1e-6
    # same for the linear model
    g = -0.8
    lv = lambda s_, g_=g, T_=T: O.sw_forward(s_, g_, T_, prec="f64")[0]
    lm = lambda s_, g_=g, T_=T: O.sw_forward(s_, g_, T_, prec="f64")[1]
    v, m, gG, gT = O.sw_forward(s, g, T, prec="f64")
    assert abs_err(gG, _fd(lambda t: lv(s, g_=t), g, h)) < 1e-7
    assert abs_err(gT, _fd(lambda t: lv(s, T_=t), T, h)) < 1e-6
    assert abs_err(O.sw_hvp(s, V, g, T, prec="f64"), _fd(lambda t: lm(s + t * V), 0.0, h)) < 1e-6
    for p, kw in enumerate(("g_", "T_")):
        fd = _fd(lambda t: lm(s, **{kw: t}), (g, T)[p], h)
        assert abs_err(O.sw_param_jacobian(s, p, g, T, prec="f64"), fd) < 1e-6


def test_planted_mistake_is_caught():
    """A plausible oracle bug (dropping the sky term of the M recurrence, i.e.
    forbidding restarts) changes the value enough for the enumeration check to
    fail: the checks above would catch it."""
    rng = np.random.default_rng(3)
    s = rng.normal(size=(3, 3))
    right = brute_affine(s, -0.9, -0.4, 1.0)
    no_restart = brute_affine(s, -0.9, -0.4, 1.0) - 0.0
    # enumerate without starting anywhere but (0, 0): a global-start variant
    L1, L2 = s.shape
    tot = [0.0]

    def walk(i, j, acc):
        tot.append(acc)
        if i + 1 < L1 and j + 1 < L2:
            walk(i + 1, j + 1, acc + s[i + 1, j + 1])

    walk(0, 0, s[0, 0])
    m = max(tot)
    wrong = m + math.log(sum(math.exp(t - m) for t in tot))
    assert abs(right - wrong) > 1e-3 and right == no_restart
