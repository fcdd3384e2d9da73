"""GPU parity at the benchmark sizes, where size-dependent code paths run:
the product grouping of the products-materialised mode (13824^3 222 L2:
729 tiles in groups of 7), the two-stage operand summation and the
two-stage W stage (333 and 444 at two levels), the 864-wide 444 L2 leaves,
the rank-1152 update and the 16384^3 headline in the automatic mode.

Size-independent property: with small-integer inputs every algorithm is
exact in FP64 (all sums and products stay far below 2^53), so an FMM result
must equal the classical kernel's bit for bit (the classical kernel itself
is pinned against the oracle at the sizes the oracle finishes in seconds).
The plain configuration (222 L1, 2048^3, seed-0 SplitMix64 inputs) is
checked against the oracle's own FMM and classical product.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import alg_path

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def device():
    rc = F.lib().fmmhip_init()
    if rc != 0:
        pytest.fail(f"no usable gfx950 device: {F.STATUS.get(rc)}")
    yield
    F.check(F.lib().fmmhip_workspace_release(), "workspace_release")


_cache = {}


def operands(m, n, k):
    """Integer A, B, C (amplitude 2) and the classical result, kept per shape."""
    key = (m, n, k)
    if key not in _cache:
        _cache.clear()
        A = F.DeviceMatrix(m, k).fill_int(21, 0, 2)
        B = F.DeviceMatrix(k, n).fill_int(21, 1, 2)
        C0 = F.DeviceMatrix(m, n).fill_int(21, 2, 2)
        Cref = F.DeviceMatrix(m, n).fill_int(21, 2, 2)
        F.dgemm(m, n, k, 1.0, A, B, 1.0, Cref)
        F.sync()
        _cache[key] = (A, B, C0, Cref)
    return _cache[key]


def fmm_alg(name, levels):
    f = F.strassen() if name == "222" else F.new_fmm(alg_path(name))
    return F.Alg([f] * levels)


@pytest.mark.parametrize("m,n,k,name,levels,pack", [
    (13824, 13824, 13824, "222", 2, 1),
    (13824, 13824, 13824, "222", 2, 2),
    (13824, 13824, 13824, "333", 2, 1),
    (13824, 13824, 13824, "333", 2, 2),
    (13824, 13824, 13824, "444", 2, 1),
    (13824, 13824, 13824, "444", 2, 2),
    (13824, 13824, 1152, "444", 2, -1),
    (13824, 13824, 1152, "444", 2, 2),
    (16384, 16384, 16384, "222", 2, -1),
])
def test_fmm_equals_classical_at_full_size(m, n, k, name, levels, pack):
    say(f"full-size integer property: {name} L{levels} pack {pack} at {m}x{n}x{k}")
    A, B, C0, Cref = operands(m, n, k)
    alg = fmm_alg(name, levels)
    alg.set_pack(pack)
    expect = alg.pack_for(m, n, k)
    C = F.DeviceMatrix(m, n)
    F.check(F.lib().fmmhip_memcpy_d2d(C.ptr, C0.ptr, 8 * m * n, None), "copy C")
    F.dgemm_fmm(alg, m, n, k, 1.0, A, B, 1.0, C)
    F.sync()
    took = alg.last_pack()
    d = F.fro_diff(C, Cref)
    say(f"  mode {took} (pack_for {expect}), ||C_fmm - C_classical||_F = {d}")
    assert took == expect
    assert d == 0.0
    assert F.fro(Cref) > 0
    C.free()
    alg.close()


def test_plain_configuration_against_the_oracle():
    """The plain configuration: one-level Strassen, FP64 column-major,
    C += A*B at m = n = k = 2048, A, B, C uniform[-1,1] from seed 0
    (SplitMix64, the device fill is bit-identical to the oracle's), against
    the oracle's FMM (the reference's algorithm) and its classical product.
    Bars: normalised error ||C - C_ref||_F / (||A||_F ||B||_F) <= 1e-15
    against both (FP64 rounding of a 1024-deep leaf sum; measured values are
    printed)."""
    n = 2048
    A, B, C = O.uniform(n, n, 0, 0), O.uniform(n, n, 0, 1), O.uniform(n, n, 0, 2)
    dA = F.DeviceMatrix(n, n).fill_uniform(0, 0)
    dB = F.DeviceMatrix(n, n).fill_uniform(0, 1)
    dC = F.DeviceMatrix(n, n).fill_uniform(0, 2)
    assert np.array_equal(dA.to_numpy(), A) and np.array_equal(dC.to_numpy(), C)
    F.dgemm_fmm(F.Alg([F.strassen()]), n, n, n, 1.0, dA, dB, 1.0, dC)
    got = dC.to_numpy()
    ref_fmm = O.fmm_dgemm(O.parse(alg_path("222")), 1.0, A, B, 1.0, C.copy(order="F"))
    ref_cls = O.dgemm(1.0, A, B, 1.0, C.copy(order="F"))
    e_fmm = O.normalized_error(got, ref_fmm, A, B)
    e_cls = O.normalized_error(got, ref_cls, A, B)
    e_orc = O.normalized_error(ref_fmm, ref_cls, A, B)
    say(f"plain config 222 L1 2048^3 seed 0: vs oracle FMM {e_fmm:.3e}, vs oracle classical {e_cls:.3e} "
        f"(oracle FMM vs its classical {e_orc:.3e})")
    assert e_fmm <= 1e-15
    assert e_cls <= 1e-15


@pytest.mark.parametrize("name", ["333", "444"])
def test_automatic_mode_equals_the_reported_explicit_mode(name):
    """333 and 444 at two levels: the automatic mode's result equals, bit for
    bit on uniform inputs, the result of the explicit mode fmmhip_alg_pack_for
    reports (the automatic choice depends on the shape and the device only,
    never on free memory)."""
    m, n, k = 1728, 1600, 1500
    say(f"auto vs explicit pack: {name} L2 {m}x{n}x{k}")
    A = F.DeviceMatrix(m, k).fill_uniform(3, 0)
    B = F.DeviceMatrix(k, n).fill_uniform(3, 1)
    alg = fmm_alg(name, 2)
    mode = alg.pack_for(m, n, k)
    C1 = F.DeviceMatrix(m, n).fill_uniform(3, 2)
    F.dgemm_fmm(alg, m, n, k, 0.7, A, B, -1.3, C1)
    F.sync()
    assert alg.last_pack() == mode
    alg.set_pack(mode)
    C2 = F.DeviceMatrix(m, n).fill_uniform(3, 2)
    F.dgemm_fmm(alg, m, n, k, 0.7, A, B, -1.3, C2)
    F.sync()
    say(f"  mode {mode}")
    assert np.array_equal(C1.to_numpy(), C2.to_numpy())
    alg.close()
