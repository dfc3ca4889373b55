"""Shared helpers for the tests: golden fixtures, brute-force enumeration and
FP64 references."""

from __future__ import annotations

import glob
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def golden_files(kind: str | None = None):
    files = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
    out = []
    for f in files:
        d = np.load(f, allow_pickle=False)
        if kind is None or str(d["kind"]) == kind:
            out.append(f)
    return out


def load(path):
    return dict(np.load(path, allow_pickle=False))


def affine_params(d):
    return float(d["gap_open"]), float(d["gap_ext"]), float(d["temp"])


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def abs_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b))) if a.size else 0.0


# --------------------------------------------------------------------------
# Brute force: enumerate every local alignment path of the soft SW models and
# sum exp(score / T).  No dynamic programming; exponential, tiny grids only.
# --------------------------------------------------------------------------


def _lse(vals, T):
    m = max(vals)
    return m + T * math.log(sum(math.exp((v - m) / T) for v in vals))


def brute_affine(s, go, ge, T):
    """Paths start in M at any cell with that cell's score; from M(i,j) they
    may go to M(i+1,j+1) (+s), I(i+1,j) (+go), D(i,j+1) (+go); from I to
    M(i+1,j+1) (+s) or I(i+1,j) (+ge); from D to M(i+1,j+1) (+s) or
    D(i,j+1) (+ge).  Every path prefix is an alignment; the empty alignment
    scores 0."""
    L1, L2 = s.shape
    totals = [0.0]

    def walk(i, j, st, acc):
        totals.append(acc)
        if i + 1 < L1 and j + 1 < L2:
            walk(i + 1, j + 1, "M", acc + s[i + 1, j + 1])
        if st in "MI" and i + 1 < L1:
            walk(i + 1, j, "I", acc + (go if st == "M" else ge))
        if st in "MD" and j + 1 < L2:
            walk(i, j + 1, "D", acc + (go if st == "M" else ge))

    for i in range(L1):
        for j in range(L2):
            walk(i, j, "M", s[i, j])
    return _lse(totals, T)


def brute_nw_affine(s, go, ge, T):
    """Global affine alignment: every path from the start (0, 0) in M to the
    cell (L1, L2) in any state; moves diag into M (+ the new cell's score),
    down into I (+ge from I, +go otherwise), right into D (+ge from D, +go
    otherwise).  An empty side leaves one gap run (or the empty path)."""
    L1, L2 = s.shape
    totals = []

    def walk(i, j, st, acc):
        if i == L1 and j == L2:
            totals.append(acc)
            return
        if i < L1 and j < L2:
            walk(i + 1, j + 1, "M", acc + s[i, j])
        if i < L1:
            walk(i + 1, j, "I", acc + (ge if st == "I" else go))
        if j < L2:
            walk(i, j + 1, "D", acc + (ge if st == "D" else go))

    walk(0, 0, "M", 0.0)
    return _lse(totals, T)


def brute_sv_affine(s, go, ge, T):
    """Saigo-Vert: paths start in M at any cell with that cell's score and end
    in M; from M(i,j) they go to M(i+1,j+1) (+s), I(i+1,j) (+go) or D(i,j+1)
    (+go); from I to M (+s), I(i+1,j) (+ge) or D(i,j+1) (+go); from D to M
    (+s) or D(i,j+1) (+ge).  The empty alignment scores 0."""
    L1, L2 = s.shape
    totals = [0.0]

    def walk(i, j, st, acc):
        if st == "M":
            totals.append(acc)
        if i + 1 < L1 and j + 1 < L2:
            walk(i + 1, j + 1, "M", acc + s[i + 1, j + 1])
        if st in "MI" and i + 1 < L1:
            walk(i + 1, j, "I", acc + (go if st == "M" else ge))
        if j + 1 < L2:
            walk(i, j + 1, "D", acc + (ge if st == "D" else go))

    for i in range(L1):
        for j in range(L2):
            walk(i, j, "M", s[i, j])
    return _lse(totals, T)


def brute_nw_linear(s, g, T):
    """Linear-gap global alignment: every path of diagonal (+ the new cell's
    score), down (+g) and right (+g) moves from (0, 0) to (L1, L2)."""
    L1, L2 = s.shape
    totals = []

    def walk(i, j, acc):
        if i == L1 and j == L2:
            totals.append(acc)
            return
        if i < L1 and j < L2:
            walk(i + 1, j + 1, acc + s[i, j])
        if i < L1:
            walk(i + 1, j, acc + g)
        if j < L2:
            walk(i, j + 1, acc + g)

    walk(0, 0, 0.0)
    return _lse(totals, T)


def brute_linear(s, g, T):
    """Linear gap: start at any cell with its score; moves diag (+s of the new
    cell), down (+g), right (+g); every prefix counts; empty alignment 0."""
    L1, L2 = s.shape
    totals = [0.0]

    def walk(i, j, acc):
        totals.append(acc)
        if i + 1 < L1 and j + 1 < L2:
            walk(i + 1, j + 1, acc + s[i + 1, j + 1])
        if i + 1 < L1:
            walk(i + 1, j, acc + g)
        if j + 1 < L2:
            walk(i, j + 1, acc + g)

    for i in range(L1):
        for j in range(L2):
            walk(i, j, s[i, j])
    return _lse(totals, T)


def close_or_nearer_truth(got, ref, truth, tol):
    """True when `got` matches the reference's FP32 result `ref` to `tol`, or
    else deviates from the FP64 result `truth` by at most twice what `ref`
    itself does.

    Second-order outputs (HVP, parameter sensitivities, marginals backward)
    carry FP32 rounding noise of order ulp(|A|) / T per cell in the reference
    as well; two FP32 implementations that evaluate the same exact quantity in
    different orders differ by about that noise, so the fair bar is distance
    to the exact value.  The reference's own CPU-vs-GPU bar for these outputs
    is far looser (rtol 1e-2, atol 5e-3: tests/test_soft_sw_affine.py:538-543).
    `truth` is a callable so the FP64 oracle only runs when needed."""
    if abs_err(got, ref) < tol:
        return True
    t = truth()
    return abs_err(got, t) <= max(tol, 2.0 * abs_err(ref, t))


def stable_seed(*parts) -> int:
    """A seed from the test's parameters that is the same in every process
    (Python's hash() of a str is salted per process)."""
    import zlib
    return zlib.crc32(repr(parts).encode()) & 0x7FFFFFFF
