"""GPU parity of the INT8x4 leaf (FMMHIP_LEAF_INT8X4): FP64 operand sums cut
into four int8 slices per value (base 254, one scale per row of the summed
operand), ten slice products on the int8 MFMA with exact int32 sums, FP64
recombination and FP64 W write-back.

Bars (stated per test):
  * uniform[-1,1] inputs at 1024^3: normalised error
    ||C - C_ref||_F/(||A||_F ||B||_F) against the oracle's FP64 classical
    product below TOL.  A float64/int64 restatement of the slicing (the same
    digits, exact slice products, the same recombination) gives 1.7e-11 for
    the classical leaf at 1024^3 (2.4e-11 at 512^3: the error falls as
    1/sqrt(k)); the bounds are 1.5x that, times the growth that one- and
    two-level Strassen's operand sums and W recombination bring to a
    slicing error of this kind (2.7x and 6.8x, measured on the split BF16
    leaf, whose error has the same origin); <3,3,3;23> (sums of up to five
    parts) takes the two-level bound (measured 1.03e-10, 6.1x);
  * ragged shapes, alpha/beta, beta = 0 with a NaN-filled C (C not read),
    small-integer and uniform inputs: normalised error below BOUND_TERM =
    5 e against the oracle, e = 0.5 / (127 * 254^3) = 2.4e-10 (about
    2^-32; the classical leaf and 1- and 2-level FMM): each slice product's
    term errs by at most e of the row
    maximum times the column maximum for each operand's representation and
    below 3 e for the dropped slice products; over k terms of random sign
    the normalised error falls about as 1/sqrt(k), so the precision target
    below holds at large k, while at k of a few it is this bound (3 x 13 x
    7 with small integers: 2.0e-10);
  * rows of A and columns of B scaled by 10^U(-4,4) (768^3): normalised error
    (restatement 2.4e-11 classical) and, for the classical leaf, elementwise
    error relative to the row/column scales (restatement 2.8e-8);
  * k above the int32-exact limit (32768 leaf columns) runs as consecutive
    k ranges: 64 x 40000 x 72 against the oracle;
  * the product split (G workgroups per tile) against the unsplit result,
    repeatable bit for bit;
  * the precision target at full size: the classical algorithm and one- and
    two-level Strassen at 8192^3 and 16384^3, uniform and stress inputs,
    within 1e-10 of the FP64 classical kernel (which the parity tests pin
    against the oracle);
  * a NaN or an infinity in A makes exactly its row of C NaN; a zero row of
    A leaves beta*C in its row.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu

MODEL_CLASSICAL = 1.7e-11          # restatement, classical leaf, 1024^3 uniform
TOL = {"classical": 1.5 * MODEL_CLASSICAL, "222x1": 1.5 * 2.7 * MODEL_CLASSICAL,
       "222x2": 1.5 * 6.8 * MODEL_CLASSICAL, "333x1": 1.5 * 6.8 * MODEL_CLASSICAL}
TOL_STRESS = {"classical": 1.5 * 2.4e-11, "222x2": 1.5 * 6.8 * 2.4e-11}
TOL_ELEM = 1.5 * 2.8e-8
TARGET = 1e-10     # the precision form's bound against the FP64 product
BOUND_TERM = 5 * 0.5 / (127 * 254.0 ** 3)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    rc = F.lib().fmmhip_init()
    if rc != 0:
        pytest.fail(f"no usable gfx950 device: {F.STATUS.get(rc)}")
    yield


def dev(X):
    D = F.DeviceMatrix(X.shape[0], X.shape[1], dtype=X.dtype)
    D.upload(X)
    return D


def run(alg, alpha, A, B, beta, C):
    m, k = A.shape
    n = B.shape[1]
    dA, dB, dC = dev(A), dev(B), dev(C)
    F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC, leaf=F.LEAF_INT8X4)
    F.sync()
    return dC.to_numpy()


def algs():
    s = F.strassen()
    return {"classical": (F.Alg(classical=True), None),
            "222x1": (F.Alg([s]), O.parse(alg_path("222"))),
            "222x2": (F.Alg([s, s]), O.compose(O.parse(alg_path("222")), O.parse(alg_path("222")))),
            "333x1": (F.Alg([F.new_fmm(alg_path("333"))]), O.parse(alg_path("333")))}


SHAPES = [(1, 1, 1), (3, 13, 7), (96, 96, 1), (129, 130, 257), (300, 280, 600), (64, 512, 300)]


@pytest.mark.parametrize("name", ["classical", "222x1", "222x2", "333x1"])
def test_int8x4_ragged_shapes(name):
    alg, og = algs()[name]
    for (m, n, k) in SHAPES:
        for kind in ("integers", "uniform"):
            say("int8x4 ragged", name, kind, m, n, k)
            if kind == "integers":
                A, B, C = O.integers(m, k, 5, 0, 4), O.integers(k, n, 5, 1, 4), O.integers(m, n, 5, 2, 4)
            else:
                A, B, C = O.uniform(m, k, 5, 0), O.uniform(k, n, 5, 1), O.uniform(m, n, 5, 2)
            for alpha, beta in [(1.0, 1.0), (-2.0, 0.5), (1.0, 0.0)]:
                ref = O.dgemm(alpha, A, B, beta, C.copy(order="F"))
                Cn = C.copy(order="F")
                if beta == 0.0:
                    Cn[:] = np.nan
                got = run(alg, alpha, A, B, beta, Cn)
                err = O.normalized_error(got, ref, A, B) / max(abs(alpha), 1.0)
                assert np.isfinite(got).all(), (m, n, k, alpha, beta)
                assert err < BOUND_TERM, (m, n, k, alpha, beta, err)


@pytest.mark.parametrize("name", ["classical", "222x1", "222x2", "333x1"])
def test_int8x4_random_error(name):
    alg, _ = algs()[name]
    m = n = k = 1024
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    got = run(alg, 1.0, A, B, 1.0, C)
    err = O.normalized_error(got, ref, A, B)
    say(f"int8x4 {name} 1024^3 normalised error {err:.3e} (bound {TOL[name]:.2e})")
    assert err < TOL[name]


@pytest.mark.parametrize("name", ["classical", "222x2"])
def test_int8x4_scaled_rows_and_columns(name):
    """Stress variant: rows of A and columns of B scaled by 10^U(-4,4)."""
    alg, _ = algs()[name]
    m = n = k = 768
    rng = np.random.default_rng(7)
    A = np.asfortranarray(O.uniform(m, k, 1, 0) * (10.0 ** rng.uniform(-4, 4, size=(m, 1))))
    B = np.asfortranarray(O.uniform(k, n, 1, 1) * (10.0 ** rng.uniform(-4, 4, size=(1, n))))
    C = np.zeros((m, n), order="F")
    ref = O.dgemm(1.0, A, B, 0.0, C.copy(order="F"))
    got = run(alg, 1.0, A, B, 0.0, C)
    err = O.normalized_error(got, ref, A, B)
    say(f"int8x4 {name} stress normalised error {err:.3e}")
    assert err < TOL_STRESS[name]
    if name == "classical":
        # elementwise, relative to the row/column scales: per-row scaling keeps
        # small rows as accurate as large ones (an FMM mixes rows of different
        # scales in its sums, so this holds for the classical leaf only)
        scale = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=0)[None, :] * np.sqrt(k)
        rel = np.abs(got - ref) / scale
        say(f"  max scaled elementwise error {rel.max():.3e}")
        assert rel.max() < TOL_ELEM


def test_int8x4_long_k_in_ranges():
    """k = 40000 > 32768: two k ranges, the second accumulating into C."""
    m, n, k = 64, 72, 40000
    A, B, C = O.uniform(m, k, 3, 0), O.uniform(k, n, 3, 1), O.uniform(m, n, 3, 2)
    ref = O.dgemm(1.5, A, B, -0.5, C.copy(order="F"))
    got = run(F.Alg(classical=True), 1.5, A, B, -0.5, C)
    err = O.normalized_error(got, ref, A, B)
    say(f"int8x4 classical 64x72x40000 normalised error {err:.3e}")
    assert err < TOL["classical"]


def test_int8x4_product_split_repeatable():
    """Product split (G workgroups per tile, pre-summed pack) against the
    unsplit call: within the random-error bound, repeatable bit for bit."""
    alg, _ = algs()["222x2"]
    m = n = k = 1024
    A, B, C = O.uniform(m, k, 8, 0), O.uniform(k, n, 8, 1), O.uniform(m, n, 8, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    for pack in (1, 2):
        alg.set_pack(pack)
        for g in (1, 2, 8):
            alg.set_split(g)
            outs = [run(alg, 1.0, A, B, 1.0, C) for _ in range(2)]
            err = O.normalized_error(outs[0], ref, A, B)
            say(f"int8x4 222x2 pack {pack} split {g} normalised error {err:.3e}")
            assert err < TOL["222x2"]
            assert np.array_equal(outs[0], outs[1])
    alg.set_split(-1)
    alg.set_pack(-1)


def test_int8x4_nonfinite_and_zero_rows():
    m, n, k = 200, 150, 300
    A, B, C = O.uniform(m, k, 2, 0), O.uniform(k, n, 2, 1), O.uniform(m, n, 2, 2)
    A[17, 123] = np.nan
    A[40, 5] = np.inf
    A[90, :] = 0.0
    got = run(F.Alg(classical=True), 1.0, A, B, 0.5, C)
    assert np.isnan(got[17, :]).all() and np.isnan(got[40, :]).all()
    assert np.array_equal(got[90, :], 0.5 * C[90, :])
    rest = np.delete(got, [17, 40], axis=0)
    assert np.isfinite(rest).all()
    Ar = np.asfortranarray(np.delete(A, [17, 40], axis=0))
    ref = O.dgemm(1.0, Ar, B, 0.5, np.asfortranarray(np.delete(C, [17, 40], axis=0)))
    assert O.normalized_error(rest, ref, Ar, B) < BOUND_TERM


def stress_scales(n, seed):
    rng = np.random.default_rng(seed)
    return 10.0 ** rng.uniform(-4, 4, size=n)


@pytest.mark.parametrize("n", [8192, 16384])
def test_int8x4_precision_target_at_full_size(n):
    """The classical algorithm and one- and two-level Strassen with the
    FMMHIP_LEAF_INT8X4 leaf at n^3 (the precision form's sizes),
    uniform[-1,1] inputs and the stress variant (rows of A, columns of B
    scaled by 10^U(-4,4)): normalised error against the FP64 classical
    kernel on the same operands <= 1e-10."""
    A = F.DeviceMatrix(n, n).fill_uniform(0, 0)
    B = F.DeviceMatrix(n, n).fill_uniform(0, 1)
    ref, C = F.DeviceMatrix(n, n), F.DeviceMatrix(n, n)
    s = F.strassen()
    for inputs in ("uniform", "stress"):
        if inputs == "stress":
            rs, cs = F.DeviceBuffer(8 * n), F.DeviceBuffer(8 * n)
            rs.upload(stress_scales(n, 1))
            cs.upload(stress_scales(n, 2))
            F.check(F.lib().fmmhip_scale_rows_cols(A.ptr, n, n, n, rs.ptr, None, None), "scale A rows")
            F.check(F.lib().fmmhip_scale_rows_cols(B.ptr, n, n, n, None, cs.ptr, None), "scale B cols")
        F.dgemm(n, n, n, 1.0, A, B, 0.0, ref)
        F.sync()
        nA, nB = F.fro(A), F.fro(B)
        for levels in (0, 1, 2):
            say(f"int8x4 precision target: {inputs} 222 L{levels} at {n}^3")
            alg = F.Alg(classical=True) if levels == 0 else F.Alg([s] * levels)
            F.dgemm_fmm(alg, n, n, n, 1.0, A, B, 0.0, C, leaf=F.LEAF_INT8X4)
            F.sync()
            err = F.fro_diff(C, ref) / (nA * nB)
            say(f"  normalised error {err:.3e} (bound {TARGET:g})")
            assert err <= TARGET
    F.check(F.lib().fmmhip_workspace_release(), "workspace_release")
