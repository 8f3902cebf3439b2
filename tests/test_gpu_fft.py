"""GPU: the batched 3D FFT sub-API (batchfftPlan3d/Execute, batchfft.h:83-108) against
numpy's FFT and the oracle FFT, and state conversions against the oracle."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(128, 16, 16, 3), (64, 16, 16, 2), (256, 32, 32, 1), (128, 32, 32, 2)]


def _data(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec,tol", [(0, 2e-6), (1, 1e-13)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(bec, shape, prec, tol, sign):
    nz, ny, nx, batch = shape
    x = _data((batch, nz, ny, nx), 1)
    plan = bec.FFTPlan(nz, ny, nx, batch, prec)
    y = plan.execute(x, sign)
    plan.close()
    if sign < 0:
        ref = np.fft.fftn(x, axes=(1, 2, 3))
    else:
        ref = np.fft.ifftn(x, axes=(1, 2, 3)) * (nz * ny * nx)
    assert rel_l2(y, ref) < tol


def test_fft_matches_oracle_fft(bec, oracle):
    x = _data((2, 128, 16, 16), 7)
    plan = bec.FFTPlan(128, 16, 16, 2, 1)
    y = plan.execute(x, -1)
    plan.close()
    assert rel_l2(y, oracle.fft3d(x, -1)) < 1e-13


def test_fft_rejects_bad_shapes(bec):
    with pytest.raises(bec.BecError):
        bec.FFTPlan(100, 16, 16, 1)
    with pytest.raises(bec.BecError):
        bec.FFTPlan(128, 16, 32, 1)


@pytest.mark.parametrize("prec,tol", [(0, 2e-6), (1, 1e-13)])
def test_state_roundtrip_and_kspace(bec, oracle, prec, tol):
    from tests.helpers import gpu_calc, oracle_calc
    kw = dict(N=30000, ne=3)
    c = gpu_calc(bec, **kw)
    oc = oracle_calc(oracle, **kw)
    a = _data((3, 128, 16, 16), 3)
    b = _data((3, 128, 16, 16), 4)
    with bec.Solver(c, precision=prec) as s:
        s.set_state(a, b, bec.SPACE_X)
        ax, bx = s.get_state(bec.SPACE_X)
        assert rel_l2(ax, a) < tol and rel_l2(bx, b) < tol
        ak, bk = s.get_state(bec.SPACE_K)
        assert rel_l2(ak, oracle.to_kspace(oc, a)) < tol
        assert rel_l2(bk, oracle.to_kspace(oc, b)) < tol
        s.set_state(ak, bk, bec.SPACE_K)
        ax2, _ = s.get_state(bec.SPACE_X)
        assert rel_l2(ax2, a) < 2 * tol
