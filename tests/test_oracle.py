"""Pins the CPU oracle (oracle/fmm_oracle.c) before anything is checked
against it.  Every check here is independent of the oracle's own formulas:
exact integer identities (Brent equations), library routines (numpy.matmul),
brute force, the textbook Strassen formulas, closed forms of the geometry,
and the published SplitMix64 algorithm.  Mistakes are planted in the oracle's
data on the CPU to show each check can fail.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import ALG_NAMES, alg_path

HEADERS = {"222": (2, 2, 2, 7), "232": (2, 3, 2, 11), "322": (3, 2, 2, 11),
           "333": (3, 3, 3, 23), "444": (4, 4, 4, 49)}
# nonzeros of U, V, W per file (counted by hand from the reference's
# coefficient files, independently of the parser under test)
NNZ = {"222": (12, 12, 12), "232": (16, 16, 16), "322": (16, 17, 17),
       "333": (59, 53, 53), "444": (194, 194, 194)}


@pytest.fixture(scope="module")
def algs():
    return {n: O.parse(alg_path(n)) for n in ALG_NAMES}


def brent_numpy(U, V, W, mt, kt, nt):
    """Brent residual computed with numpy.einsum (independent of the C check)."""
    T = np.einsum("ar,br,cr->abc", U, V, W)
    want = np.zeros_like(T)
    for i in range(mt):
        for p in range(kt):
            for j in range(nt):
                want[i * kt + p, p * nt + j, i * nt + j] = 1
    return int(np.count_nonzero(T - want))


@pytest.mark.parametrize("name", ALG_NAMES)
def test_parse_header_and_nnz(algs, name):
    a = algs[name]
    assert (a.mt, a.kt, a.nt, a.R) == HEADERS[name]
    U, V, W = a.tables()
    assert (np.count_nonzero(U), np.count_nonzero(V), np.count_nonzero(W)) == NNZ[name]


@pytest.mark.parametrize("name", ALG_NAMES)
def test_brent_equations_hold(algs, name):
    a = algs[name]
    assert O.brent_violations(a) == 0
    U, V, W = a.tables()
    assert brent_numpy(U, V, W, a.mt, a.kt, a.nt) == 0


def test_brent_check_catches_planted_mistakes(algs):
    # a sign flip in one coefficient
    a = O.parse(alg_path("222"))
    a.U[0] = -a.U[0]
    assert O.brent_violations(a) > 0
    # the reference's header misreading: 232 read as m,n,k (bli_read_fmm.c:33)
    b = O.parse(alg_path("232"))
    b.kt, b.nt = b.nt, b.kt
    assert O.brent_violations(b) > 0


@pytest.mark.parametrize("pair", [("222", "222"), ("222", "232"), ("232", "322"), ("333", "222"), ("322", "232")])
def test_composition(algs, pair):
    a, b = algs[pair[0]], algs[pair[1]]
    c = O.compose(a, b)
    assert (c.mt, c.kt, c.nt, c.R) == (a.mt * b.mt, a.kt * b.kt, a.nt * b.nt, a.R * b.R)
    assert O.brent_violations(c) == 0
    # the composed tables equal the block-Kronecker product, built with numpy
    for X1, X2, Xc, (r1, c1), (r2, c2) in zip(
            a.tables(), b.tables(), c.tables(),
            [(a.mt, a.kt), (a.kt, a.nt), (a.mt, a.nt)], [(b.mt, b.kt), (b.kt, b.nt), (b.mt, b.nt)]):
        T1 = X1.reshape(r1, c1, a.R)
        T2 = X2.reshape(r2, c2, b.R)
        K = np.einsum("ijr,klq->ikjlrq", T1, T2).reshape(r1 * r2 * c1 * c2, a.R * b.R)
        assert np.array_equal(K, Xc)


@pytest.mark.parametrize("whole,tilde", [(13, 2), (7, 3), (2048, 2), (13824, 9), (5, 4), (3, 4), (0, 2), (1, 3)])
def test_part_geometry_closed_form(whole, tilde):
    sz, off = O.part_geometry(whole, tilde)
    assert sum(sz) == whole
    q, r = divmod(whole, tilde)
    assert sz == [q + 1] * r + [q] * (tilde - r)
    assert off == [sum(sz[:t]) for t in range(tilde)]


def test_part_geometry_worked_examples():
    # init_part_offsets (bli_strassen_ab.c:86-121) on the test's 3 x 13 x 7
    assert O.part_geometry(13, 2) == ([7, 6], [0, 7])
    assert O.part_geometry(3, 2) == ([2, 1], [0, 2])
    assert O.part_geometry(7, 3) == ([3, 2, 2], [0, 3, 5])


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (3, 13, 7), (64, 48, 80), (130, 257, 129)])
def test_classical_matches_numpy_and_brute_force(m, n, k):
    A = O.uniform(m, k, 1, 0)
    B = O.uniform(k, n, 1, 1)
    C0 = O.uniform(m, n, 1, 2)
    C = O.dgemm(1.5, A, B, -0.5, C0.copy(order="F"))
    ref = -0.5 * C0 + 1.5 * (A @ B)
    assert O.normalized_error(C, ref, A, B) < 1e-15
    # brute force, exact on integers
    Ai, Bi, Ci = O.integers(m, k, 2, 0, 3), O.integers(k, n, 2, 1, 3), O.integers(m, n, 2, 2, 3)
    got = O.dgemm(1.0, Ai, Bi, 1.0, Ci.copy(order="F"))
    brute = O.dgemm_naive(Ai, Bi, Ci.copy(order="F"))
    assert np.array_equal(got, brute)
    assert np.array_equal(got, Ci + Ai @ Bi)


def test_classical_beta_zero_does_not_read_c():
    A, B = O.uniform(5, 4, 3, 0), O.uniform(4, 6, 3, 1)
    C = np.full((5, 6), np.nan, order="F")
    O.dgemm(1.0, A, B, 0.0, C)
    assert np.all(np.isfinite(C))
    assert np.allclose(C, A @ B, rtol=0, atol=1e-14)


SHAPES = [(3, 13, 7), (31, 33, 35), (64, 64, 64), (97, 101, 103), (1, 5, 3), (2, 1, 1), (17, 1, 19), (4, 4, 0)]


@pytest.mark.parametrize("name", ALG_NAMES)
@pytest.mark.parametrize("levels", [1, 2])
def test_fmm_exact_on_integers(algs, name, levels):
    """Small-integer inputs make every algorithm exact: FMM == classical bit
    for bit.  Any wrong sign, index, dropped term or transposed operand in the
    FMM restatement breaks the equality."""
    if name == "444" and levels == 2:
        shapes = [(20, 18, 17)]
    else:
        shapes = SHAPES
    a = algs[name]
    g = a if levels == 1 else O.compose(a, a)
    for (m, n, k) in shapes:
        for alpha, beta in [(1.0, 1.0), (2.0, -3.0), (0.5, 0.0)]:
            A, B = O.integers(m, k, 7, 0, 4), O.integers(k, n, 7, 1, 4)
            C = O.integers(m, n, 7, 2, 4)
            want = O.dgemm(alpha, A, B, beta, C.copy(order="F"))
            got = O.fmm_dgemm(g, alpha, A, B, beta, C.copy(order="F"))
            assert np.array_equal(got, want), (name, levels, m, n, k, alpha, beta)


def test_fmm_restatement_catches_planted_mistakes(algs):
    A, B = O.integers(12, 12, 9, 0, 3), O.integers(12, 12, 9, 1, 3)
    C = np.zeros((12, 12), order="F")
    want = A @ B
    for plant in range(3):
        a = O.parse(alg_path("222"))
        if plant == 0:
            a.W[4] = -a.W[4]                     # wrong sign (W[0][4] = -1)
        elif plant == 1:
            a.V[2 * a.R + 3], a.V[1 * a.R + 3] = a.V[1 * a.R + 3], a.V[2 * a.R + 3]   # swapped parts (transpose)
        else:
            a.U[0] = 0                           # dropped term
        got = O.fmm_dgemm(a, 1.0, A, B, 0.0, C.copy(order="F"))
        assert not np.array_equal(got, want), plant


def test_textbook_strassen(algs):
    """222.txt is Strassen's 1969 algorithm: compare with the textbook
    M1..M7 formulas on 2x2 blocks (A11 A12; A21 A22)."""
    n = 16
    A, B = O.integers(n, n, 11, 0, 5), O.integers(n, n, 11, 1, 5)
    h = n // 2
    A11, A12, A21, A22 = A[:h, :h], A[:h, h:], A[h:, :h], A[h:, h:]
    B11, B12, B21, B22 = B[:h, :h], B[:h, h:], B[h:, :h], B[h:, h:]
    M1 = (A11 + A22) @ (B11 + B22)
    M2 = (A21 + A22) @ B11
    M3 = A11 @ (B12 - B22)
    M4 = A22 @ (B21 - B11)
    M5 = (A11 + A12) @ B22
    M6 = (A21 - A11) @ (B11 + B12)
    M7 = (A12 - A22) @ (B21 + B22)
    C = np.block([[M1 + M4 - M5 + M7, M3 + M5], [M2 + M4, M1 - M2 + M3 + M6]])
    got = O.fmm_dgemm(algs["222"], 1.0, A, B, 0.0, np.zeros((n, n), order="F"))
    assert np.array_equal(got, C)
    U, V, W = algs["222"].tables()
    assert list(U[:, 0]) == [1, 0, 0, 1] and list(V[:, 0]) == [1, 0, 0, 1]   # M1


@pytest.mark.parametrize("name", ALG_NAMES)
def test_fmm_random_error_small(algs, name):
    a = algs[name]
    m = n = k = 144
    A, B, C = O.uniform(m, k, 0, 0), O.uniform(k, n, 0, 1), O.uniform(m, n, 0, 2)
    ref = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    e1 = O.normalized_error(O.fmm_dgemm(a, 1.0, A, B, 1.0, C.copy(order="F")), ref, A, B)
    e2 = O.normalized_error(O.fmm_dgemm(O.compose(a, a), 1.0, A, B, 1.0, C.copy(order="F")), ref, A, B)
    assert e1 < 1e-15 and e2 < 1e-14


def _splitmix_stream(state, count):
    """Vigna's SplitMix64 generator, Python integers (textbook form)."""
    out, M = [], (1 << 64) - 1
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & M
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def _finalize(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_splitmix_known_answer():
    # published first outputs of SplitMix64 seeded with 1234567
    assert _splitmix_stream(1234567, 2) == [6457827717110365317, 3203168211198807973]


@pytest.mark.parametrize("seed,stream", [(0, 0), (0, 1), (42, 7)])
def test_uniform_fill_is_splitmix(seed, stream):
    m, n = 7, 5
    X = O.uniform(m, n, seed, stream)
    M = (1 << 64) - 1
    base = _finalize((seed ^ ((stream * 0xD1B54A32D192ED03) & M)) & M)
    zs = _splitmix_stream(base, m * n)
    want = np.array([2.0 * ((z >> 11) * 2.0 ** -53) - 1.0 for z in zs]).reshape(n, m).T
    assert np.array_equal(X, want)
    I = O.integers(m, n, seed, stream, 3)
    assert np.array_equal(I, np.array([float((z >> 11) % 7 - 3) for z in zs]).reshape(n, m).T)


def test_uniform_fill_statistics():
    X = O.uniform(512, 512, 0, 0)
    assert X.min() >= -1.0 and X.max() < 1.0
    assert abs(X.mean()) < 0.01 and abs(X.var() - 1.0 / 3.0) < 0.01
    assert not np.array_equal(X, O.uniform(512, 512, 0, 1))
    assert np.array_equal(X, O.uniform(512, 512, 0, 0))
