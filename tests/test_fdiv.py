"""The stencils' division by f32(2h) (csrc/fdiv.hpp) against IEEE division.

The interior kernels divide central differences by a launch constant with
Markstein's fused multiply-add sequence from the host's RN(1/b); hillshade is
bit-exact only if every such quotient equals the IEEE one.  tests/cpp/
fdiv_check.cpp compiles the same header on the host (g++, no contraction) and
compares bits over random and exhaustive dividends, the quotients of elevation
differences by 2h for pixel sizes in use, and zeros, subnormals, infinities
and NaN; the unguarded form curvature uses must agree wherever its range
holds (zeros up to their sign).  To show the harness discriminates, the uncorrected product
a * RN(1/b) planted in a copy of the header must fail it.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "fdiv_check.cpp")


def _build(tmp_path, extra=()):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "fdiv_check")
    cmd = [gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", *extra,
           "-o", exe, SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_div_rn_equals_ieee_division(tmp_path):
    exe = _build(tmp_path)
    res = subprocess.run([exe, "20000000", "2"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "mismatches 0" in res.stdout


def test_the_check_rejects_the_uncorrected_product(tmp_path):
    # the same harness with div_rn replaced by a * y: the check must fail
    hdr = tmp_path / "fujishadergpu_amd" / "csrc"
    hdr.mkdir(parents=True)
    src = open(os.path.join(HERE, "..", "fujishadergpu_amd", "csrc", "fdiv.hpp")).read()
    body_start = src.index("inline float div_rn(")
    body_end = src.index("}\n", src.index("return a / b;")) + 2
    planted = src[:body_start] + (
        "inline float div_rn(float a, float b, float y) { (void)b; return a * y; }\n"
    ) + src[body_end:]
    (hdr / "fdiv.hpp").write_text(planted)
    cpp = tmp_path / "tests" / "cpp"
    cpp.mkdir(parents=True)
    shutil.copy(SRC, cpp / "fdiv_check.cpp")
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "planted")
    res = subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas",
                          "-o", exe, str(cpp / "fdiv_check.cpp")], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    res = subprocess.run([exe, "200000", "0"], capture_output=True, text=True, timeout=300)
    assert res.returncode != 0 and "mismatches 0" not in res.stdout
