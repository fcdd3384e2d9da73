"""CPU tests of the library's host side through its C ABI: the library loads
and exports what include/fmm_blis_amd.h declares, the coefficient reader
(new_fmm, frame/base/bli_read_fmm.c:24-80) agrees with the oracle on every
file and rejects malformed ones, STRASSEN_FMM equals 222.txt, composition
and the Brent check, and the error behaviour of the object entries.  No
device computation happens here."""
import ctypes
import os
import re
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import fmm_blis_amd as F
import oracle as O
from conftest import ALG_NAMES, ROOT, alg_path

HAS_GPU = os.path.exists("/dev/kfd")


def header_functions():
    text = open(os.path.join(ROOT, "include", "fmm_blis_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"typedef[^;]*\(\s*\*[^;]*;", "", text)   # function-pointer types
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text))
    keywords = {"if", "sizeof", "defined"}
    return sorted(n for n in names if n not in keywords and not n.isupper())


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    declared = header_functions() + ["STRASSEN_FMM", "FMM_BLIS_PACK_UKR", "FMM_BLIS_GEMM_UKR"]
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    L = F.lib()
    for n in F.EXPORTED:
        getattr(L, n)


@pytest.mark.parametrize("name", ALG_NAMES)
def test_new_fmm_matches_oracle(name):
    f = F.new_fmm(alg_path(name))
    o = O.parse(alg_path(name))
    try:
        assert (f.m_tilde, f.k_tilde, f.n_tilde, f.R) == (o.mt, o.kt, o.nt, o.R)
        for got, want in zip(f.tables(), o.tables()):
            assert np.array_equal(np.array(got), want)
    finally:
        F.free_fmm(f)


def test_strassen_builtin_equals_222():
    s = F.strassen()
    o = O.parse(alg_path("222"))
    assert (s.m_tilde, s.k_tilde, s.n_tilde, s.R) == (2, 2, 2, 7)
    for got, want in zip(s.tables(), o.tables()):
        assert np.array_equal(np.array(got), want)


BAD_FILES = {
    "bad_header": "2 2 x 7\n",
    "short_header": "2 2 2\n1 0\n",
    "short_row": "1 1 1 2\n1\n#\n1 1\n#\n1 1\n",
    "too_few_rows": "1 1 1 1\n1\n#\n1\n",
    "too_many_rows": "1 1 1 1\n1\n1\n1\n1\n",
    "token": "1 1 1 1\n1\n#\nq\n#\n1\n",
    "zero_dim": "0 1 1 1\n",
}


@pytest.mark.parametrize("case", sorted(BAD_FILES))
def test_new_fmm_rejects_malformed(tmp_path, case):
    p = tmp_path / f"{case}.txt"
    p.write_text(BAD_FILES[case])
    f = F.new_fmm(str(p))
    assert f.R == 0 and not f.U and not f.V and not f.W
    F.free_fmm(f)


def test_new_fmm_missing_file(tmp_path):
    f = F.new_fmm(str(tmp_path / "nope.txt"))
    assert f.R == 0 and not f.U


def test_new_fmm_long_lines_and_comments(tmp_path):
    # a 1x1x1 R=200 algorithm whose rows are far longer than 128 characters
    # (the reference's LINE_N), with comments between the blocks
    R = 200
    u = [1] + [0] * (R - 1)
    rows = [" ".join(map(str, u))] * 3
    p = tmp_path / "long.txt"
    p.write_text(f"1 1 1 {R}\n# comment\n{rows[0]}\n#\n{rows[1]}\n# another\n\n{rows[2]}\n")
    f = F.new_fmm(str(p))
    assert f.R == R and f.tables()[0][0][:2] == [1, 0]
    a = F.Alg([f])
    assert a.info()["R"] == 1          # all-zero products are dropped
    assert a.brent_violations() == 0
    F.free_fmm(f)


@pytest.mark.parametrize("names", [("222",), ("232",), ("322",), ("333",), ("444",), ("222", "222"),
                                   ("222", "333"), ("232", "322"), ("444", "444")])
def test_alg_compose_matches_oracle(names):
    fs = [F.new_fmm(alg_path(n)) for n in names]
    a = F.Alg(fs)
    info = a.info()
    g = O.parse(alg_path(names[0]))
    for n in names[1:]:
        g = O.compose(g, O.parse(alg_path(n)))
    U, V, W = g.tables()
    assert (info["mt"], info["kt"], info["nt"], info["R"]) == (g.mt, g.kt, g.nt, g.R)
    assert (info["nnz_u"], info["nnz_v"], info["nnz_w"]) == tuple(int(np.count_nonzero(x)) for x in (U, V, W))
    if names != ("444", "444"):
        assert a.brent_violations() == 0
    a.close()
    for f in fs:
        F.free_fmm(f)


def test_alg_set_split_values():
    """Product split per tile: -1 (per call), 1, 2, 4 or 8; others rejected."""
    a = F.Alg([F.strassen()] * 2)
    for g in (-1, 1, 2, 4, 8):
        a.set_split(g)
    for g in (0, 3, 16, -2):
        with pytest.raises(F.FmmError):
            a.set_split(g)
    a.close()


def test_alg_set_pack_values():
    """Pack mode: -1 (per call), 0 fused, 1 pre-summed, 2 products
    materialised; others rejected."""
    a = F.Alg([F.strassen()] * 2)
    for p in (-1, 0, 1, 2):
        a.set_pack(p)
    for p in (3, -2, 100):
        with pytest.raises(F.FmmError):
            a.set_pack(p)
    a.close()


def test_alg_create_rejects_invalid_table(tmp_path):
    text = open(alg_path("222")).read().replace("1 0 1 0 1 -1 0", "1 0 1 0 1 1 0", 1)
    p = tmp_path / "bad222.txt"
    p.write_text(text)
    f = F.new_fmm(str(p))
    assert f.R == 7
    with pytest.raises(F.FmmError):
        F.Alg([f])
    F.free_fmm(f)


@pytest.mark.skipif(HAS_GPU, reason="checks the behaviour without a device")
def test_dropin_without_device_records_error():
    L = F.lib()
    assert L.bli_plugin_register_fmm_blis() == F.BLIS_FAILURE
    L.fmm_blis_amd_set_abort_on_error(0)
    try:
        A = np.zeros((3, 7), order="F"); B = np.zeros((7, 13), order="F"); C = np.zeros((3, 13), order="F")
        oa, ob, oc = F.obj_attach(A, 3, 7, 1, 3), F.obj_attach(B, 7, 13, 1, 7), F.obj_attach(C, 3, 13, 1, 3)
        one, s1 = F.obj_scalar(1.0)
        L.bli_strassen_ab(one, oa, ob, one, oc)
        assert L.fmm_blis_amd_last_error() == 5   # FMMHIP_ERR_NODEVICE
    finally:
        L.fmm_blis_amd_set_abort_on_error(1)


@pytest.mark.skipif(HAS_GPU, reason="checks the behaviour without a device")
def test_dropin_aborts_by_default_like_blis():
    code = textwrap.dedent(f"""
        import sys, numpy as np
        sys.path.insert(0, {os.path.join(ROOT, 'fmm-blis_amd')!r})
        import fmm_blis_amd as F
        A = np.zeros((3, 7), order="F"); B = np.zeros((7, 13), order="F"); C = np.zeros((3, 13), order="F")
        one, s = F.obj_scalar(1.0)
        F.lib().bli_strassen_ab(one, F.obj_attach(A, 3, 7, 1, 3), F.obj_attach(B, 7, 13, 1, 7), one,
                                F.obj_attach(C, 3, 13, 1, 3))
        print("returned")
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "bli_strassen_ab_ex" in r.stderr


def test_obj_attach_layout():
    A = np.arange(12.0).reshape(3, 4, order="F")
    o = F.obj_attach(A, 3, 4, 1, 3)
    assert (o.dim[0], o.dim[1], o.rs, o.cs, o.elem_size, o.info & 7) == (3, 4, 1, 3, 8, F.BLIS_DOUBLE)
    import ctypes
    assert ctypes.sizeof(F.ObjT) == 160      # BLIS obj_t on LP64


def test_c_header_layout_matches_binding(tmp_path):
    """The C declarations of obj_t and fmm_t (compiled with gcc) have the
    offsets the Python binding uses."""
    src = tmp_path / "lay.c"
    src.write_text(textwrap.dedent("""
        #include <stdio.h>
        #include <stddef.h>
        #include "fmm_blis_amd.h"
        int main(void) {
          printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(obj_t), offsetof(obj_t, buffer),
                 offsetof(obj_t, rs), offsetof(obj_t, scalar), offsetof(obj_t, n_panel),
                 sizeof(fmm_t), offsetof(fmm_t, k_tilde), offsetof(fmm_t, U));
          return 0;
        }"""))
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    want = [ctypes.sizeof(F.ObjT), F.ObjT.buffer.offset, F.ObjT.rs.offset, F.ObjT.scalar.offset,
            F.ObjT.n_panel.offset, ctypes.sizeof(F.FmmT), F.FmmT.k_tilde.offset, F.FmmT.U.offset]
    assert got == want


def test_constdata_layout_matches_the_library_read():
    """BLIS_CONSTANT scalars (BLIS_ONE, ...) carry a constdata_t
    (bli_type_defs.h:1200-1208: float s, double d, scomplex c, dcomplex z,
    gint_t i); the library reads the double value at byte 8, which is where
    the platform C ABI puts d after a 4-byte float."""
    C = F.ConstData
    assert (C.s.offset, C.d.offset, C.c.offset, C.z.offset, C.i.offset) == (0, 8, 16, 24, 40)
    assert ctypes.sizeof(C) == 48


def test_tuner_choices_follow_the_cost_model():
    """The tuner's choices (host logic only): classical for small or thin
    problems, two-level Strassen for FP64 16384^3; a level needs leaves of
    at least two tiles and k of at least 8 BK-steps per level."""
    assert F.tuner_choose(True, 256, 256, 256)[0] == 0
    assert F.tuner_choose(True, 16384, 16384, 16384)[0] == 2
    assert F.tuner_choose(True, 16384, 16384, 64)[0] == 0           # rank-64 update
    assert F.tuner_choose(True, 16384, 16384, 300)[0] in (0, 1)     # k too short for two levels
    c, est = F.tuner_choose(False, 8192, 8192, 8192)
    assert c in (0, 1, 2) and est > 0
    # the chosen candidate's estimate is reported, and an FMM candidate is
    # chosen only when it beats every simpler one by the margin: 3% where its
    # leaves are whole 128 x 128 tiles and whole k-steps, 5% elsewhere
    for dbl in (True, False):
        for (m, n, k) in [(4096, 4096, 4096), (16384, 16384, 16384), (2048, 8192, 1024), (7104, 6336, 3392),
                          (12288, 12288, 12288), (6336, 4800, 7424)]:
            for lone in (True, False):
                c, t, est = F.tuner_choose_tile(dbl, m, n, k, lone=lone)
                assert abs(F.tuner_estimate(dbl, c, m, n, k, lone=lone) - est) < 1e-9
                d = 2 ** c
                whole = t == 128 and m % (128 * d) == 0 and n % (128 * d) == 0 and k % ((16 if dbl else 32) * d) == 0
                for simpler in range(c):
                    assert est < (0.97 if whole else 0.95) * F.tuner_estimate(dbl, simpler, m, n, k, lone=lone) + 1e-12
    # FMM ahead of the simpler candidates by more than the margin on whole
    # tiles (4 x 12288^3 in a batch ran 170 ms on two levels, 203
    # classical), not where edge tiles run the general kernel
    assert F.tuner_choose_tile(True, 12288, 12288, 12288, lone=False)[:2] == (2, 128)
    assert F.tuner_choose(True, 7104, 6336, 3392, lone=False)[0] == 0
    # more work never costs less
    assert F.tuner_estimate(True, 0, 8192, 8192, 8192) > F.tuner_estimate(True, 0, 4096, 4096, 4096)
    with pytest.raises(F.FmmError):
        F.tuner_choose(True, 0, 5, 5)


def test_tuner_tile_choice_follows_padding():
    """The tile the tuner picks (host logic only): 128 x 128 where the leaf
    dimensions are multiples of 128 (the larger tile runs faster on whole
    tiles), 64 x 64 where the 128 x 128 grid pads by more than the smaller
    tile's rate loss (a leaf edge of 192 pads to 256, a third of the work
    wasted), and 128 x 128 for a lone small classical problem (the
    one-problem path splits its k instead)."""
    assert F.tuner_choose_tile(True, 16384, 16384, 16384)[:2] == (2, 128)
    assert F.tuner_choose_tile(True, 4096, 4096, 4096, lone=False)[1] == 128
    assert F.tuner_choose_tile(True, 192, 192, 4096, lone=False)[:2] == (0, 64)
    assert F.tuner_choose_tile(False, 576, 320, 2048, lone=False)[1] == 64
    assert F.tuner_choose_tile(True, 192, 192, 4096, lone=True)[1] == 128
    # 1280 = 10 x 128; its one-level leaf 640 = 5 x 128, its two-level leaf 320 = 5 x 64
    c, t, est = F.tuner_choose_tile(True, 1280, 1280, 16384, lone=False)
    assert t == (64 if c == 2 else 128)
    # the reported estimate is the chosen candidate's on its faster tile
    for (m, n, k) in [(192, 192, 4096), (4096, 4096, 4096), (3000, 1000, 700)]:
        c, t, est = F.tuner_choose_tile(True, m, n, k, lone=False)
        assert abs(F.tuner_estimate(True, c, m, n, k, lone=False) - est) < 1e-9
        assert F.tuner_choose(True, m, n, k, lone=False) == (c, est)


def test_tuner_table_save_load_round_trip(tmp_path):
    p = tmp_path / "tuner.txt"
    F.tuner_save(str(p))
    before = [F.tuner_estimate(True, c, 8192, 8192, 8192) for c in range(3)]
    text = p.read_text()
    assert text.count("cand ") == 6 and "bw " in text and "tile_eff 1.000000 0.880000" in text
    # the 64 x 64 tile's rate enters the tile choice: at half the rate its
    # saved padding (256^2 against 192^2) no longer pays
    assert F.tuner_choose_tile(True, 192, 192, 4096, lone=False)[1] == 64
    slow = tmp_path / "slow64.txt"
    slow.write_text(text.replace("tile_eff 1.000000 0.880000", "tile_eff 1.000000 0.500000"))
    F.tuner_load(str(slow))
    try:
        assert F.tuner_choose_tile(True, 192, 192, 4096, lone=False)[1] == 128
    finally:
        F.tuner_load(str(p))
    # a table with another FP64 classical efficiency changes that estimate only
    lines = [ln if not ln.startswith("cand 0 f64") else "cand 0 f64 0.390000 4.000000" for ln in text.splitlines()]
    q = tmp_path / "t2.txt"
    q.write_text("\n".join(lines) + "\n")
    F.tuner_load(str(q))
    try:
        after = [F.tuner_estimate(True, c, 8192, 8192, 8192) for c in range(3)]
        assert after[0] > 1.9 * before[0] and after[1:] == before[1:]
    finally:
        F.tuner_load(str(p))
    assert [F.tuner_estimate(True, c, 8192, 8192, 8192) for c in range(3)] == before
    bad = tmp_path / "bad.txt"
    bad.write_text("nothing useful\n")
    with pytest.raises(F.FmmError):
        F.tuner_load(str(bad))
    with pytest.raises(F.FmmError):
        F.tuner_load(str(tmp_path / "missing.txt"))


@pytest.mark.skipif(HAS_GPU, reason="checks the behaviour without a device")
def test_batch_create_without_device_fails():
    h = ctypes.c_void_p()
    assert F.lib().fmmhip_batch_create(ctypes.byref(h)) == 5       # FMMHIP_ERR_NODEVICE


@pytest.mark.skipif(HAS_GPU, reason="checks the behaviour without a device")
def test_server_create_without_device_fails():
    h = ctypes.c_void_p()
    assert F.lib().fmmhip_server_create(ctypes.byref(h), 1 << 30, 4) == 5       # FMMHIP_ERR_NODEVICE


def test_reference_style_c_harness_compiles_and_links(tmp_path):
    """The C counterpart of the reference's harness (test_strassen_oapi.c)
    compiles against the header and links against the library."""
    exe = tmp_path / "strassen_oapi"
    subprocess.run(["gcc", "-std=c11", "-O2", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "strassen_oapi.c"), "-o", str(exe),
                    "-L", os.path.dirname(F.LIB_PATH), "-lfmm_blis_amd", "-lm",
                    f"-Wl,-rpath,{os.path.dirname(F.LIB_PATH)}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
