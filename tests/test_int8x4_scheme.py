"""CPU checks of the INT8x4 leaf's slicing (fmmhip_int8x4.hip), restated in
float64/int64 numpy: the digits the slice pass writes, their range, the
representation error, and the classical-leaf error that the GPU tests'
bounds are derived from.  Also the leaf's enum value against the Python
constant.

The restatement follows the kernel step for step: per row, mx = f 2^E
(frexp), y = ldexp(x, -E) * (127 / f), four digits d = rint(r), r <- 254 (r
- d); ten slice products of weight 254^-(i+j), i + j <= 3, summed exactly
(int64 here, int32 on the device: exact while k * 4 * 127^2 < 2^31), then
t = G0 + (G1 + (G2 + G3/254)/254)/254 times the row and column inverse
scales mx / 127.
"""
import os
import re

import numpy as np

import fmm_blis_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def slices(X):
    """Digits (four int64 arrays) and inverse scales of the rows of X."""
    mx = np.abs(X).max(axis=1, keepdims=True)
    f, E = np.frexp(mx)
    f = np.where(mx > 0, f, 0.5)
    E = np.where(mx > 0, E, 0)
    r = np.ldexp(X, -E) * (127.0 / f)
    d = []
    for _ in range(4):
        ds = np.rint(r)
        d.append(ds.astype(np.int64))
        r = (r - ds) * 254.0
    return d, (mx / 127.0)[:, 0]


def product(A, B):
    da, sa = slices(A)
    db, sb = slices(np.ascontiguousarray(B.T))
    G = [np.zeros((A.shape[0], B.shape[1]), np.int64) for _ in range(4)]
    for i in range(4):
        for j in range(4 - i):
            G[i + j] += da[i] @ db[j].T
    inv = 1.0 / 254.0
    t = (((G[3] * inv + G[2]) * inv + G[1]) * inv + G[0])
    return t * sa[:, None] * sb[None, :]


def test_leaf_enum_matches_header():
    text = open(os.path.join(ROOT, "include", "fmm_blis_amd.h")).read()
    m = re.search(r"FMMHIP_LEAF_INT8X4\s*=\s*(\d+)", text)
    assert m and int(m.group(1)) == F.LEAF_INT8X4 == 5


def test_digits_in_int8_range_and_representation_error():
    rng = np.random.default_rng(3)
    # uniform rows, rows spanning many decades, a row at the largest finite
    # magnitudes and a tiny one: digits stay in [-127, 127] and the value
    # is represented to 0.5 / (127 * 254^3) of its row's maximum (half a
    # unit of the fourth digit: 2.4e-10, about 2^-32)
    X = np.vstack([rng.uniform(-1, 1, (8, 300)),
                   rng.uniform(-1, 1, (8, 300)) * 10.0 ** rng.uniform(-30, 30, (8, 300)),
                   np.full((1, 300), 1.7e308) * rng.uniform(-1, 1, 300),
                   np.full((1, 300), 1e-300) * rng.uniform(-1, 1, 300)])
    d, s = slices(X)
    for ds in d:
        assert np.abs(ds).max() <= 127
    y = d[0] + (d[1] + (d[2] + d[3] / 254.0) / 254.0) / 254.0
    rec = y * s[:, None]
    rowmax = np.abs(X).max(axis=1, keepdims=True)
    err = np.abs(rec - X) / rowmax
    assert err.max() <= 0.5 / (127 * 254.0 ** 3) * (1 + 1e-9), err.max()
    # the bound is attained to within a factor of a few: the last digit
    # really is the fourth (a dropped or misweighted digit fails above or here)
    assert err.max() > 2.0 ** -36


def test_integer_sums_stay_exact_in_int32():
    # the largest slice-product sum the kernel forms: four terms of 127^2
    # over k = 32768 (the host splits longer k)
    assert 4 * 127 * 127 * 32768 < 2 ** 31
    assert 4 * 127 * 127 * 65536 >= 2 ** 31


def test_classical_error_model():
    """The restated classical leaf at 512^3 and 1024^3 on uniform inputs:
    the normalised errors the GPU bounds scale from (1.7e-11 at 1024^3),
    falling about as 1/sqrt(k)."""
    rng = np.random.default_rng(0)
    errs = {}
    for n in (512, 1024):
        A = rng.uniform(-1, 1, (n, n))
        B = rng.uniform(-1, 1, (n, n))
        C = product(A, B)
        R = A @ B
        errs[n] = np.linalg.norm(C - R) / (np.linalg.norm(A) * np.linalg.norm(B))
    assert errs[1024] < 2.0e-11, errs
    assert 1.1 < errs[512] / errs[1024] < 2.0, errs
