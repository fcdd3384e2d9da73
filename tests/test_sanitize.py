"""Host C of the library under AddressSanitizer and UndefinedBehaviorSanitizer:
the coefficient-file parser, level composition, part geometry and
descriptor lists (fmm-blis_amd/csrc/fmm_spec.c) built with gcc
-fsanitize=address,undefined and driven by tests/c/spec_sanitize.c over the
algorithm files, malformed and oversized files, one- and two-level and mixed
compositions and uneven geometries.  Any sanitizer report fails the run
(-fno-sanitize-recover=all).  No GPU is involved."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def asan_available(tmp):
    src = os.path.join(tmp, "probe.c")
    with open(src, "w") as f:
        f.write("int main(void) { return 0; }\n")
    r = subprocess.run(["gcc", "-fsanitize=address,undefined", src, "-o", os.path.join(tmp, "probe")],
                       capture_output=True)
    return r.returncode == 0


def test_spec_under_sanitizers(tmp_path):
    if shutil.which("gcc") is None or not asan_available(str(tmp_path)):
        pytest.skip("gcc with libasan/libubsan is not available")
    exe = tmp_path / "spec_sanitize"
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-fno-sanitize-recover=all", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "fmm-blis_amd", "csrc"), os.path.join(ROOT, "tests", "c", "spec_sanitize.c"),
           os.path.join(ROOT, "fmm-blis_amd", "csrc", "fmm_spec.c"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden", "algorithms"), str(scratch)],
                       capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "spec_sanitize: ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
