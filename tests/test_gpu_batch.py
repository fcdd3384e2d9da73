"""GPU parity of the batching form (fmmhip_batch_*): many uneven FP64 and
FP32 GEMMs, each with its own algorithm (classical, 1- or 2-level Strassen,
forced or chosen by the tuner), run as one captured HIP graph.

Bars (stated per test):
  * integer-valued inputs: every candidate is exact, so each problem's C
    equals the oracle's classical product bit for bit, in both datatypes;
  * uniform[-1,1] inputs: normalised error against the oracle below 5e-15
    (FP64) and 1e-6 (FP32, against the product of the FP32-rounded inputs);
  * a problem's result in the batch equals the same problem run alone
    through the single-call entry with the same candidate, bit for bit;
  * every replay of the graph, with eager work between replays, reproduces
    the first replay bit for bit;
  * the tile (128 x 128 or 64 x 64, chosen per problem or forced) does not
    change any result: each element of C is accumulated in the same order on
    either, so forced-64, forced-128 and tuner-chosen batches agree bit for
    bit, and a single call on 64 x 64 tiles equals one on 128 x 128 tiles.
"""
import sys

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O

pytestmark = pytest.mark.gpu

# (dtype, m, n, k, cand); cand -1 = tuner's choice
MIX = [
    ("f8", 300, 280, 260, 2), ("f4", 300, 280, 260, 2), ("f8", 257, 255, 129, 1), ("f4", 257, 255, 129, 1),
    ("f8", 64, 64, 64, 0), ("f4", 65, 63, 62, 0), ("f8", 1, 5, 3, 2), ("f4", 17, 1, 19, 1),
    ("f8", 513, 100, 77, -1), ("f4", 100, 513, 77, -1), ("f8", 130, 140, 600, 2), ("f4", 130, 140, 600, 2),
    ("f8", 31, 33, 35, 1), ("f4", 31, 33, 35, 0), ("f8", 384, 256, 128, 0), ("f4", 256, 384, 128, 1),
    ("f8", 200, 64, 333, -1), ("f4", 96, 96, 1, 2), ("f8", 97, 101, 103, 2), ("f4", 7, 9, 11, -1),
]
ALPHA_BETA = [(1.0, 1.0), (2.0, -1.0), (1.0, 0.0), (-1.0, 2.0)]


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


def inputs(dt, m, n, k, i, integer):
    if integer:
        amp = 4 if dt == "f8" else 2
        A, B, C = O.integers(m, k, 11, 3 * i, amp), O.integers(k, n, 11, 3 * i + 1, amp), \
            O.integers(m, n, 11, 3 * i + 2, amp)
    else:
        A, B, C = O.uniform(m, k, 12, 3 * i), O.uniform(k, n, 12, 3 * i + 1), O.uniform(m, n, 12, 3 * i + 2)
    if dt == "f4":
        A, B, C = (np.asfortranarray(X.astype(np.float32)) for X in (A, B, C))
    return A, B, C


def build(integer):
    """Device operands and the batch over MIX; returns (batch, problems)."""
    b = F.Batch()
    probs = []
    for i, (dt, m, n, k, cand) in enumerate(MIX):
        A, B, C = inputs(dt, m, n, k, i, integer)
        alpha, beta = ALPHA_BETA[i % len(ALPHA_BETA)]
        dA, dB, dC = dev(A), dev(B), dev(C)
        b.add(m, n, k, alpha, dA, dB, beta, dC, cand=cand)
        probs.append(dict(dt=dt, m=m, n=n, k=k, cand=cand, alpha=alpha, beta=beta, A=A, B=B, C=C,
                          dA=dA, dB=dB, dC=dC))
    return b, probs


def expected(p):
    A, B, C = (np.asfortranarray(X.astype(np.float64)) for X in (p["A"], p["B"], p["C"]))
    return O.dgemm(p["alpha"], A, B, p["beta"], C.copy(order="F"))


def alg_for(cand):
    """The single-call algorithm the batch runs for `cand` (pre-summed, no
    product split)."""
    a = F.Alg(classical=True) if cand == 0 else F.Alg([F.strassen()] * cand)
    if cand > 0:
        a.set_pack(1)
    a.set_split(1)
    return a


def test_batch_exact_integers_mixed_dtypes():
    b, probs = build(integer=True)
    b.plan()
    info = b.info()
    say("batch int", info)
    assert info["problems"] == len(MIX)
    assert info["items_f64"] > 0 and info["items_f32"] > 0
    b.run()
    F.sync()
    for i, p in enumerate(probs):
        got = p["dC"].to_numpy()
        want = expected(p)
        if p["dt"] == "f4":
            want = want.astype(np.float32)
        chosen = b.choice(i)
        assert chosen in (0, 1, 2)
        if p["cand"] >= 0:
            assert chosen == p["cand"]
        assert np.array_equal(got, want), (i, p["dt"], p["m"], p["n"], p["k"], chosen)


def test_batch_random_error_and_alone_equals_mix():
    b, probs = build(integer=False)
    b.plan()
    b.run()
    F.sync()
    for i, p in enumerate(probs):
        got = p["dC"].to_numpy()
        A64, B64 = p["A"].astype(np.float64), p["B"].astype(np.float64)
        err = O.normalized_error(got.astype(np.float64), expected(p), A64, B64)
        tol = 5e-15 if p["dt"] == "f8" else 1e-6
        assert err < tol, (i, p["dt"], err)
        # the same problem alone, same candidate, fresh C
        cand = b.choice(i)
        m, n, k = p["m"], p["n"], p["k"]
        dC1 = dev(p["C"])
        alg = alg_for(cand)
        if p["dt"] == "f8":
            F.dgemm_fmm(alg, m, n, k, p["alpha"], p["dA"], p["dB"], p["beta"], dC1)
        else:
            F.sgemm_fmm(alg, m, n, k, p["alpha"], p["dA"], p["dB"], p["beta"], dC1)
        F.sync()
        assert np.array_equal(dC1.to_numpy(), got), (i, p["dt"], m, n, k, cand)


def run_batch(integer, tile=None):
    b, probs = build(integer)
    if tile is not None:
        for i in range(len(probs)):
            b.set_tile(i, tile)
    b.plan()
    b.run()
    F.sync()
    outs = [p["dC"].to_numpy() for p in probs]
    tiles = [b.tile(i) for i in range(len(probs))]
    cands = [b.choice(i) for i in range(len(probs))]
    b.close()
    return outs, tiles, cands, probs


def test_batch_tile_choice_does_not_change_results():
    auto, t_auto, c_auto, probs = run_batch(False)
    say("tiles chosen:", t_auto)
    assert set(t_auto) == {64, 128}            # the mix holds problems for either tile
    for tile in (64, 128):
        outs, tiles, cands, _ = run_batch(False, tile)
        say("forced", tile, tiles)
        assert tiles == [tile] * len(MIX) and cands == c_auto
        for i, (x, y) in enumerate(zip(auto, outs)):
            assert np.array_equal(x, y), (tile, i, MIX[i])
    # integers: exact on 64 x 64 tiles too
    outs, tiles, cands, probs = run_batch(True, 64)
    for i, p in enumerate(probs):
        want = expected(p)
        if p["dt"] == "f4":
            want = want.astype(np.float32)
        assert np.array_equal(outs[i], want), (i, MIX[i], cands[i])


# (dtype, m, n, k, cand, pack): ragged edges, leaves that are odd multiples
# of 64, every pack mode of the FP64 path
SINGLE = [
    ("f8", 320, 192, 448, 0, -1), ("f8", 1000, 700, 900, 1, 1), ("f8", 1000, 700, 900, 2, 1),
    ("f8", 1000, 700, 900, 1, 0), ("f8", 1000, 700, 900, 2, 2), ("f8", 1, 5, 3, 0, -1),
    ("f8", 513, 100, 77, 1, 1), ("f4", 320, 192, 448, 0, -1), ("f4", 1000, 700, 900, 2, 1),
    ("f4", 97, 101, 103, 1, 0),
]


@pytest.mark.parametrize("case", SINGLE, ids=lambda c: "-".join(map(str, c)))
def test_single_call_64_tiles_equal_128_tiles(case):
    dt, m, n, k, cand, pack = case
    i = SINGLE.index(case)
    A, B, C = inputs(dt, m, n, k, 40 + i, False)
    alpha, beta = ALPHA_BETA[i % len(ALPHA_BETA)]
    dA, dB = dev(A), dev(B)
    outs = []
    for shape in (0, 2):
        alg = alg_for(cand)
        if cand > 0 and pack >= 0:
            alg.set_pack(pack)
        alg.set_shape(shape)
        dC = dev(C)
        fn = F.dgemm_fmm if dt == "f8" else F.sgemm_fmm
        fn(alg, m, n, k, alpha, dA, dB, beta, dC)
        F.sync()
        outs.append(dC.to_numpy())
        alg.close()
    assert np.array_equal(outs[0], outs[1]), case
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    want = O.dgemm(alpha, np.asfortranarray(A64), np.asfortranarray(B64), beta,
                   np.asfortranarray(C.astype(np.float64)))
    err = O.normalized_error(outs[1].astype(np.float64), want, A64, B64)
    assert err < (5e-15 if dt == "f8" else 1e-6), (case, err)


# (m, n, k, cand, pack): FP64 shapes whose leaves the fast kernel takes
FAST = [(1024, 1024, 1024, 0, -1), (2048, 1024, 1536, 1, 1), (2048, 2048, 2048, 2, 1), (2048, 2048, 2048, 2, 2),
        (1024, 1024, 512, 0, -1)]


@pytest.mark.parametrize("case", FAST, ids=lambda c: "-".join(map(str, c)))
def test_fast_kernel_equals_shape0(case):
    """The automatic shape takes the fast FP64 kernel (shape 3) on these
    shapes, and its C equals the 512-thread kernel's (shape 0) bit for bit,
    with the product split of the materialised mode too; integer inputs
    equal the oracle."""
    m, n, k, cand, pack = case
    i = FAST.index(case)
    for integer in (False, True):
        A, B, C = inputs("f8", m, n, k, 60 + i, integer)
        alpha, beta = (0.7, -1.3) if not integer else (2.0, -1.0)
        dA, dB = dev(A), dev(B)
        outs, shapes = [], []
        for shape in (-1, 0):
            alg = alg_for(cand)
            if cand > 0:
                alg.set_pack(pack)
            alg.set_shape(shape)
            if pack == 2:
                alg.set_split(-1)
            dC = dev(C)
            F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC)
            F.sync()
            outs.append(dC.to_numpy())
            shapes.append(alg.last_shape())
            alg.close()
        assert shapes == [3, 0], shapes
        assert np.array_equal(outs[0], outs[1]), (case, integer)
        if integer:
            want = O.dgemm(alpha, np.asfortranarray(A), np.asfortranarray(B), beta, np.asfortranarray(C.copy()))
            assert np.array_equal(outs[0], want), case


# (dtype or leaf, m, n, k, cand, pack): FP32 operands and the FP32 leaf on
# leaves the fast kernel takes, with 1, 2, 3 and 8 of its 32-wide k-steps
# (its k-step order peels the last two steps)
FAST32 = [("f4", 256, 256, 32, 0, -1), ("f4", 256, 384, 64, 0, -1), ("f4", 384, 256, 96, 0, -1),
          ("f4", 512, 512, 1024, 2, 1), ("f4", 512, 512, 384, 2, 2), ("f8/f32", 256, 256, 32, 0, -1),
          ("f8/f32", 256, 256, 64, 0, -1), ("f8/f32", 512, 512, 256, 2, 1), ("f8/f32", 512, 512, 128, 2, 2)]


@pytest.mark.parametrize("case", FAST32, ids=lambda c: "-".join(map(str, c)).replace("/", "_"))
def test_fast_kernel_fp32_equals_shape0(case):
    """FP32 operands (sgemm) and FP64 operands with the FP32 leaf: the
    automatic shape takes the fast kernel, whose C equals shape 0's bit for
    bit; integer inputs equal the oracle."""
    kind, m, n, k, cand, pack = case
    i = FAST32.index(case)
    dt = "f4" if kind == "f4" else "f8"
    for integer in (False, True):
        A, B, C = inputs(dt, m, n, k, 80 + i, integer)
        alpha, beta = (0.7, -1.3) if not integer else (2.0, -1.0)
        dA, dB = dev(A), dev(B)
        outs, shapes = [], []
        for shape in (-1, 0):
            alg = alg_for(cand)
            if cand > 0:
                alg.set_pack(pack)
            alg.set_shape(shape)
            dC = dev(C)
            if kind == "f4":
                F.sgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC)
            else:
                F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC, leaf=F.LEAF_F32)
            F.sync()
            outs.append(dC.to_numpy())
            shapes.append(alg.last_shape())
            alg.close()
        assert shapes == [3, 0], (case, shapes)
        assert np.array_equal(outs[0], outs[1]), (case, integer)
        if integer:
            A64, B64, C64 = (np.asfortranarray(X.astype(np.float64)) for X in (A, B, C))
            want = O.dgemm(alpha, A64, B64, beta, C64.copy(order="F"))
            assert np.array_equal(outs[0].astype(np.float64), want), case


# (dtype or leaf, m, n, k, levels): products materialised with leaf m or n
# not a multiple of 128, or leaf k not a multiple of 16 (500, 250: padded to
# 512, 256 with zeros the summation writes), so the leaves and slots are
# padded to whole tiles and k-steps; the few tiles per product also make the
# launch deal whole product groups to the XCDs
PADDED = [("f8", 600, 520, 512, 1), ("f8", 1000, 1000, 1024, 2), ("f8/f32", 600, 520, 512, 1),
          ("f4", 700, 900, 1024, 2), ("f8", 1400, 300, 768, 1), ("f8", 1000, 1000, 1000, 1),
          ("f8", 1000, 900, 1000, 2), ("f8/f32", 1000, 700, 1000, 1), ("f4", 1000, 1000, 1000, 1)]


@pytest.mark.parametrize("case", PADDED, ids=lambda c: "-".join(map(str, c)).replace("/", "_"))
def test_padded_leaves_on_fast_kernel(case):
    """The materialised mode over tile-padded leaves takes the fast kernel
    (automatic shape) and gives the same C, bit for bit, as the general
    kernel over the same leaves and as the pre-summed mode (unpadded kernel
    extents, W epilogue fused); integer inputs equal the oracle."""
    kind, m, n, k, levels = case
    i = PADDED.index(case)
    dt = "f4" if kind == "f4" else "f8"
    for integer in (False, True):
        A, B, C = inputs(dt, m, n, k, 140 + i, integer)
        alpha, beta = (0.7, -1.3) if not integer else (2.0, -1.0)
        dA, dB = dev(A), dev(B)
        outs, shapes = [], []
        for pack, shape, split in ((2, -1, -1), (2, 0, -1), (1, 0, 1)):
            alg = F.Alg([F.strassen()] * levels)
            alg.set_pack(pack)
            alg.set_shape(shape)
            alg.set_split(split)
            dC = dev(C)
            if kind == "f4":
                F.sgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC)
            elif kind == "f8":
                F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC)
            else:
                F.dgemm_fmm(alg, m, n, k, alpha, dA, dB, beta, dC, leaf=F.LEAF_F32)
            F.sync()
            outs.append(dC.to_numpy())
            shapes.append(alg.last_shape())
            alg.close()
        assert shapes[:2] == [3, 0], (case, shapes)
        assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2]), (case, integer)
        if integer:
            A64, B64, C64 = (np.asfortranarray(X.astype(np.float64)) for X in (A, B, C))
            want = O.dgemm(alpha, A64, B64, beta, C64.copy(order="F"))
            assert np.array_equal(outs[0].astype(np.float64), want), case


# problems on 128 x 128 tiles whose leaves have interior tiles (run by the
# fast kernel) and edge tiles (run by the general kernel): (dtype, m, n, k, cand)
SPLIT = [("f8", 320, 448, 256, 0), ("f4", 448, 320, 192, 0), ("f8", 640, 576, 512, 1), ("f8", 1088, 960, 768, 2),
         ("f4", 704, 832, 256, 1), ("f8", 392, 264, 128, 0)]


def test_batch_interior_tiles_on_fast_kernel_equal_single_calls():
    """Each problem's interior tiles go to the fast kernel and its edge tiles
    to the general one; every C equals the same call alone bit for bit and,
    on integer inputs, the oracle."""
    for integer in (False, True):
        b = F.Batch()
        probs = []
        for i, (dt, m, n, k, cand) in enumerate(SPLIT):
            A, B, C = inputs(dt, m, n, k, 100 + i, integer)
            alpha, beta = (1.0, 1.0) if integer else ALPHA_BETA[i % len(ALPHA_BETA)]
            dA, dB, dC = dev(A), dev(B), dev(C)
            b.add(m, n, k, alpha, dA, dB, beta, dC, cand=cand)
            b.set_tile(i, 128)
            probs.append((dt, m, n, k, cand, alpha, beta, A, B, C, dA, dB, dC))
        b.plan()
        b.run()
        F.sync()
        for i, (dt, m, n, k, cand, alpha, beta, A, B, C, dA, dB, dC) in enumerate(probs):
            got = dC.to_numpy()
            dC1 = dev(C)
            alg = alg_for(cand)
            alg.set_shape(0)
            (F.dgemm_fmm if dt == "f8" else F.sgemm_fmm)(alg, m, n, k, alpha, dA, dB, beta, dC1)
            F.sync()
            alg.close()
            assert np.array_equal(dC1.to_numpy(), got), (i, dt, m, n, k, cand, integer)
            if integer:
                A64, B64, C64 = (np.asfortranarray(X.astype(np.float64)) for X in (A, B, C))
                want = O.dgemm(alpha, A64, B64, beta, C64.copy(order="F"))
                assert np.array_equal(got.astype(np.float64), want), (i, dt, m, n, k, cand)
        info = b.info()
        fast64, fast32 = b.fast_items()
        b.close()
        say(f"split batch (integer={integer}): items f64 {info['items_f64']} ({fast64} fast), "
            f"f32 {info['items_f32']} ({fast32} fast)")
        # both kernels ran tiles of both datatypes
        assert 0 < fast64 < info["items_f64"] and 0 < fast32 < info["items_f32"]


def test_batch_graph_replays_with_eager_work_between():
    b, probs = build(integer=False)
    b.plan()
    first = None
    side_A = F.DeviceMatrix(256, 256).fill_uniform(5, 0)
    side_C = F.DeviceMatrix(256, 256).fill_uniform(5, 1)
    for rep in range(3):
        for p in probs:                       # reset every C eagerly
            p["dC"].upload(p["C"])
        F.dgemm(256, 256, 256, 1.0, side_A, side_A, 1.0, side_C)
        b.run()
        F.sync()
        outs = [p["dC"].to_numpy() for p in probs]
        say("replay", rep)
        if first is None:
            first = outs
        else:
            for i, (x, y) in enumerate(zip(first, outs)):
                assert np.array_equal(x, y), (rep, i)


def test_batch_tuner_candidates_and_replan():
    """Large problems let the tuner pick Strassen; adding a problem drops the
    plan (run refuses until re-planned)."""
    b = F.Batch()
    n = 2048
    A = F.DeviceMatrix(n, n).fill_int(2, 0, 2)
    Bm = F.DeviceMatrix(n, n).fill_int(2, 1, 2)
    C1 = F.DeviceMatrix(n, n).fill_int(2, 2, 2)
    C2 = F.DeviceMatrix(n, n).fill_int(2, 2, 2)
    b.add(n, n, n, 1.0, A, Bm, 1.0, C1, cand=-1)
    b.plan()
    say("tuner choice at 2048^3 FP64:", b.choice(0))
    b.run()
    F.dgemm(n, n, n, 1.0, A, Bm, 1.0, C2)
    F.sync()
    assert F.fro_diff(C1, C2) == 0.0
    s = F.DeviceMatrix(8, 8).fill_int(2, 3, 2)
    b.add(8, 8, 8, 1.0, s, s, 0.0, F.DeviceMatrix(8, 8), cand=0)
    with pytest.raises(F.FmmError):
        b.run()
    b.plan()
    b.run()
    F.sync()
    b.close()


def test_tuner_calibration_on_device(tmp_path):
    """Calibration times every candidate on the device and refits the cost
    model; the fitted table round-trips through a file, and the refitted
    model still picks an FMM candidate for FP64 16384^3 and classical for
    small problems (the measured ordering of the candidates)."""
    before = tmp_path / "before.txt"
    F.tuner_save(str(before))
    try:
        F.tuner_calibrate()
        after = tmp_path / "after.txt"
        F.tuner_save(str(after))
        text = after.read_text()
        say(text)
        effs = [float(l.split()[3]) for l in text.splitlines() if l.startswith("cand ")]
        assert len(effs) == 6 and all(0.2 < e <= 0.98 for e in effs)
        # the 64 x 64 tile's measured rate relative to the 128 x 128 one
        teff = [float(x) for l in text.splitlines() if l.startswith("tile_eff ") for x in l.split()[1:]]
        assert teff[0] == 1.0 and 0.6 < teff[1] < 1.1, teff
        assert F.tuner_choose(True, 256, 256, 256)[0] == 0
        assert F.tuner_choose(True, 16384, 16384, 16384)[0] in (1, 2)
        F.tuner_load(str(after))
        assert (tmp_path / "after.txt").read_text() == text
    finally:
        F.tuner_load(str(before))
