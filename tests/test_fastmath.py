"""CPU: the device sin/cos of the nonlinear step (bec_amd/csrc/fastmath.hpp) built for
the host -- the same IEEE single-precision operations -- against double precision."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <stdio.h>
#include <stdlib.h>
#include "fastmath.hpp"
int main(int argc, char **argv) {
    float x; float s, c;
    const bool by_pi = argc > 1;
    const float S = argc > 2 ? (float)atof(argv[2]) : 1.0f;
    while (fread(&x, sizeof(float), 1, stdin) == 1) {
        if (argc > 2) bec::fast_sincos_pi_scaled(x, S, &s, &c);
        else if (by_pi) bec::fast_sincos_pi(x, &s, &c);
        else bec::fast_sincosf(x, &s, &c);
        fwrite(&s, sizeof(float), 1, stdout);
        fwrite(&c, sizeof(float), 1, stdout);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "h.cpp")
    exe = os.path.join(d, "h")
    with open(src, "w") as f:
        f.write(HARNESS)
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "bec_amd", "csrc"), src, "-o", exe],
                   check=True)
    return exe


def _run(exe, x, by_pi=False, scale=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    args = ["pi"] if by_pi or scale is not None else []
    if scale is not None:
        args.append(repr(float(scale)))
    out = subprocess.run([exe] + args, input=x.tobytes(), stdout=subprocess.PIPE, check=True).stdout
    sc = np.frombuffer(out, dtype=np.float32).reshape(-1, 2)
    return sc[:, 0], sc[:, 1]


@pytest.mark.parametrize("by_pi", [False, True], ids=["halfpi_reduction", "pi_reduction"])
def test_fast_sincos_accuracy(harness, by_pi):
    r = np.random.default_rng(0)
    x = np.concatenate([np.linspace(-8, 8, 200001), r.uniform(-200, 200, 200000), r.uniform(-1e-3, 1e-3, 1000),
                        np.array([0.0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, 1e4])]).astype(np.float32)
    s, c = _run(harness, x, by_pi)
    xd = x.astype(np.float64)
    es = np.abs(s - np.sin(xd))
    ec = np.abs(c - np.cos(xd))
    ulp = np.spacing(np.float32(1.0)) / 2   # 6e-8
    assert es.max() < 2.5 * ulp and ec.max() < 2.5 * ulp, (es.max(), ec.max())
    # tiny arguments: relative accuracy of sin
    small = np.abs(xd) < 1e-3
    rel = es[small] / np.maximum(np.abs(np.sin(xd[small])), 1e-30)
    assert rel.max() < 2 * ulp
    # the identity sin^2 + cos^2 = 1 holds to rounding
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 4 * ulp


@pytest.mark.parametrize("scale", [2.0 ** -8, 2.0 ** -10])
def test_scaled_sincos_is_the_scaled_result_bit_for_bit(harness, scale):
    """fast_sincos_pi_scaled(x, S) with S a power of two (the 1/N^2 the xy pass folds in)
    returns exactly S times fast_sincos_pi(x): scaling by a power of two commutes with the
    rounding of every multiply-add."""
    r = np.random.default_rng(1)
    x = np.concatenate([np.linspace(-30, 30, 100001), r.uniform(-1e4, 1e4, 50000),
                        r.uniform(-1e-6, 1e-6, 1000), np.array([0.0, np.pi / 2, -np.pi])]).astype(np.float32)
    s1, c1 = _run(harness, x, by_pi=True)
    s2, c2 = _run(harness, x, scale=scale)
    assert np.array_equal((s1 * np.float32(scale)).view(np.uint32), s2.view(np.uint32))
    assert np.array_equal((c1 * np.float32(scale)).view(np.uint32), c2.view(np.uint32))
