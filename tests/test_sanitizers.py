"""CPU: the oracle and the library's host-side parameter code built with
AddressSanitizer and UndefinedBehaviorSanitizer (gcc/g++, host code only) and
run on small cases.  Any out-of-bounds access, use after free, leak or
undefined operation aborts the run with a report."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]

PARAMS_MAIN = r"""
#include <stdio.h>
#include <initializer_list>
#include "bec_api.h"
int main() {
    bec_model_params m;
    bec_calc_params c;
    bec_2d_params p;
    if (bec_default_model_params(&m) != BEC_OK) return 1;
    for (int nz : {64, 128, 256}) {
        m.nvz = nz;
        if (bec_fill_derived_params(&m, &c) != BEC_OK) return 2;
    }
    m.nvx = m.nvy = 64;
    if (bec_fill_2d_params(&m, &p) != BEC_OK) return 3;
    m.nvx = 3;   // rejected, not read out of bounds
    if (bec_fill_derived_params(&m, &c) == BEC_OK) return 4;
    if (bec_fill_derived_params(nullptr, &c) == BEC_OK || bec_fill_2d_params(&m, nullptr) == BEC_OK) return 5;
    printf("params ok: mu %.6f, 2D g11 %.6f\n", c.mu, p.g11);
    return 0;
}
"""


def _build_and_run(cmd, exe):
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "runtime error" not in p.stderr and "ERROR: AddressSanitizer" not in p.stderr, p.stderr
    return p.stdout


@pytest.fixture(scope="module")
def tmp():
    d = tempfile.mkdtemp()
    yield d
    shutil.rmtree(d, ignore_errors=True)


def test_oracle_under_sanitizers(tmp):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    exe = os.path.join(tmp, "orc_san")
    out = _build_and_run(["gcc", "-std=c11"] + SAN + ["-I", os.path.join(ROOT, "oracle"),
                          os.path.join(ROOT, "oracle", "bec_oracle.c"),
                          os.path.join(ROOT, "oracle", "sanitize_main.c"), "-lm", "-o", exe], exe)
    assert "sanitized run" in out


def test_host_parameter_code_under_sanitizers(tmp):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src, exe = os.path.join(tmp, "params_main.cpp"), os.path.join(tmp, "params_san")
    with open(src, "w") as f:
        f.write(PARAMS_MAIN)
    out = _build_and_run(["g++", "-std=c++17"] + SAN + ["-I", os.path.join(ROOT, "include"), src,
                          os.path.join(ROOT, "bec_amd", "csrc", "params.cpp"), "-o", exe], exe)
    assert "params ok" in out
