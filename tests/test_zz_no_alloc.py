"""GPU: bec_projections_static (no device allocation; fixed scratch array)
against bec_projections on the same state, two calls in a row, the second on a
smaller lattice, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc

pytestmark = pytest.mark.gpu


def _state(bec, nxy, nz, ne, seed):
    c = gpu_calc(bec, **config1_model(bec, N=30000, nvx=nxy, nvy=nxy, nvz=nz, ne=ne))
    z, y, x = np.meshgrid(np.arange(nz) - nz // 2, np.arange(nxy) - nxy // 2, np.arange(nxy) - nxy // 2,
                          indexing="ij")
    psi = 8.0 * np.exp(-(x ** 2 + y ** 2) / 8.0 - z ** 2 / (nz * 6.0)) * np.exp(0.01j * z)
    return c, psi, seed


def _both(bec, s, alpha, c):
    L = bec.lib()
    outs = []
    for name in ("bec_projections", "bec_projections_static"):
        arrs = [np.full((c.nvy, c.nvx), np.nan), np.full((c.nvy, c.nvx), np.nan),
                np.full((c.nvy, c.nvz), np.nan), np.full((c.nvy, c.nvz), np.nan),
                np.full(c.nvz, np.nan), np.full(c.nvz, np.nan)]
        st = getattr(L, name)(s._ctx, 0, C.c_double(alpha), *[a.ctypes.data_as(C.c_void_p) for a in arrs])
        assert st == 0, (name, st)
        outs.append(arrs)
    return outs


def test_static_projections_equal_allocating_ones_twice(bec):
    for (nxy, nz, ne, seed, alpha) in ((32, 256, 4, 11, 0.3), (16, 64, 2, 12, 1.1)):
        c, psi, seed = _state(bec, nxy, nz, ne, seed)
        with bec.Solver(c, precision=bec.FP32) as s:
            s.init_evolution(psi, seed, equilibrate=True)
            s.propagate(3)
            ref, got = _both(bec, s, alpha, c)
        for r, g in zip(ref, got):
            assert np.all(np.isfinite(g))
            assert np.array_equal(r, g)
        # the projections are sums of the same per-cell values
        assert ref[0].sum() == pytest.approx(ref[4].sum(), rel=1e-12)
        assert ref[2].sum() == pytest.approx(ref[4].sum(), rel=1e-12)


def test_static_projections_reject_null_context(bec):
    assert bec.lib().bec_projections_static(None, 0, C.c_double(0.0), None, None, None, None, None, None) == 1
