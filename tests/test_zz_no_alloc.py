"""The no-allocation entry point fmmhip_dgemm_fmm_noalloc through the C ABI.

Bar: its result equals fmmhip_dgemm_fmm's in the pre-summed mode without
split (pack 1, split 1: the same kernels and sums) bit for bit, on two calls
in a row, the second with a smaller input (so the second finds the first's
lists and leaves in the shared scratch); on integer inputs both equal the
oracle's FMM bit for bit.  Device free memory is unchanged across its calls;
problems beyond the scratch and general-strided operands are refused with
their error codes.
"""
import ctypes
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu

TOO_LARGE = 7
UNSUPPORTED = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(alg, m, n, k, alpha, A, B, beta, C):
    return F.lib().fmmhip_dgemm_fmm_noalloc(alg.h, m, n, k, alpha, A.ptr, 1, A.ld, B.ptr, 1, B.ld, beta, C.ptr, 1,
                                            C.ld, None)


def free_bytes():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    F.check(F.lib().fmmhip_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return f.value


@pytest.mark.parametrize("levels", [1, 2])
def test_noalloc_equals_presummed_entry(levels):
    """Two calls in a row (2048^3 then 1000 x 900 x 1100 for one level;
    1536^3 then 700 x 650 x 600 for two levels), alpha 0.7, beta -1.3,
    uniform inputs: bit for bit the existing entry's result.  Device free
    memory is the same before and after the two calls; it is read after one
    warm-up call of the entry on a small problem, since the HIP runtime sets
    up its copy path lazily and may take device memory for it on first use
    (the leaves of the first call alone would take over 100 MB)."""
    F.check(F.lib().fmmhip_init(), "init")
    shapes = [(2048, 2048, 2048), (1000, 900, 1100)] if levels == 1 else [(1536, 1536, 1536), (700, 650, 600)]
    alg = F.Alg([F.strassen()] * levels)
    ref = F.Alg([F.strassen()] * levels)
    ref.set_pack(1)
    ref.set_split(1)
    cases = []
    for i, (m, n, k) in enumerate(shapes):
        A = F.DeviceMatrix(m, k).fill_uniform(40 + i, 0)
        B = F.DeviceMatrix(k, n).fill_uniform(40 + i, 1)
        Cn = F.DeviceMatrix(m, n).fill_uniform(40 + i, 2)
        Cr = F.DeviceMatrix(m, n).fill_uniform(40 + i, 2)
        F.dgemm_fmm(ref, m, n, k, 0.7, A, B, -1.3, Cr)   # also loads the kernels before the memory check
        cases.append((m, n, k, A, B, Cn, Cr))
    F.sync()
    w0 = free_bytes()
    wA, wB, wC = F.DeviceMatrix(64, 64).fill_uniform(1, 0), F.DeviceMatrix(64, 64).fill_uniform(1, 1), \
        F.DeviceMatrix(64, 64)
    F.sync()
    w1 = free_bytes()
    F.check(noalloc(alg, 64, 64, 64, 1.0, wA, wB, 0.0, wC), "warm-up")
    F.sync()
    before = free_bytes()
    say(f"  warm-up call: device free memory {w1} -> {before} ({w1 - before} bytes; operands {w0 - w1})")
    for (m, n, k, A, B, Cn, Cr) in cases:
        say(f"noalloc L{levels} {m}x{n}x{k}")
        F.check(noalloc(alg, m, n, k, 0.7, A, B, -1.3, Cn), "dgemm_fmm_noalloc")
    F.sync()
    after = free_bytes()
    say(f"  device free memory before/after the no-allocation calls: {before} / {after}")
    assert after == before
    for (m, n, k, A, B, Cn, Cr) in cases:
        assert np.array_equal(Cn.to_numpy(), Cr.to_numpy()), (m, n, k)


def test_noalloc_integers_against_the_oracle():
    """Integer inputs, ragged shapes, 222 and 333 at one level and 222 at two:
    bit for bit the oracle's FMM, two calls in a row each."""
    F.check(F.lib().fmmhip_init(), "init")
    for name, levels in (("222", 1), ("333", 1), ("222", 2)):
        f = F.strassen() if name == "222" else F.new_fmm(alg_path(name))
        alg = F.Alg([f] * levels)
        og = O.parse(alg_path(name))
        og = og if levels == 1 else O.compose(og, og)
        for (m, n, k) in ((301, 280, 290), (97, 101, 103)):
            say(f"noalloc int {name} L{levels} {m}x{n}x{k}")
            A, B, C = O.integers(m, k, 41, 0, 3), O.integers(k, n, 41, 1, 3), O.integers(m, n, 41, 2, 3)
            want = O.fmm_dgemm(og, 2.0, A, B, 1.0, C.copy(order="F"))
            dA, dB, dC = F.DeviceMatrix(m, k), F.DeviceMatrix(k, n), F.DeviceMatrix(m, n)
            dA.upload(A)
            dB.upload(B)
            dC.upload(C)
            F.check(noalloc(alg, m, n, k, 2.0, dA, dB, 1.0, dC), "dgemm_fmm_noalloc")
            F.sync()
            assert np.array_equal(dC.to_numpy(), want), (name, levels, m, n, k)


def test_noalloc_refusals():
    """Beyond the scratch: FMMHIP_ERR_TOO_LARGE (C untouched); an operand
    with neither stride 1: FMMHIP_ERR_UNSUPPORTED; k = 0 scales C by beta."""
    F.check(F.lib().fmmhip_init(), "init")
    alg = F.Alg([F.strassen()])
    m = n = k = 4096                      # 7 x 2 x 2048^2 doubles: 448 MiB of leaves
    A = F.DeviceMatrix(m, k).fill_uniform(42, 0)
    B = F.DeviceMatrix(k, n).fill_uniform(42, 1)
    C = F.DeviceMatrix(m, n).fill_uniform(42, 2)
    keep = C.to_numpy()
    assert noalloc(alg, m, n, k, 1.0, A, B, 1.0, C) == TOO_LARGE
    F.sync()
    assert np.array_equal(C.to_numpy(), keep)
    rc = F.lib().fmmhip_dgemm_fmm_noalloc(alg.h, 64, 64, 64, 1.0, A.ptr, 2, 2 * m, B.ptr, 1, k, 1.0, C.ptr, 1, m,
                                          None)
    assert rc == UNSUPPORTED
    s = F.DeviceMatrix(50, 40).fill_uniform(42, 3)
    want = 0.5 * s.to_numpy()
    assert F.lib().fmmhip_dgemm_fmm_noalloc(alg.h, 50, 40, 0, 1.0, A.ptr, 1, 50, B.ptr, 1, 1, 0.5, s.ptr, 1, 50,
                                            None) == 0
    F.sync()
    assert np.array_equal(s.to_numpy(), want)
