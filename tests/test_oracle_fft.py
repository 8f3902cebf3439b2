"""CPU: pins the oracle FFT (our own radix-2 transform) with a brute-force DFT,
numpy's FFT and exact identities; the reference relies on cuFFT/FFTW whose sign
pairing (FORWARD = exp(-i) for k->x, INVERSE = exp(+i) for x->k,
programs/bec_wigner_3d.cuda/wigner.cu:106-107,121,129-130) is checked here."""
import numpy as np
import pytest


def _rand(shape, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


@pytest.mark.parametrize("shape", [(4, 4, 4), (8, 4, 16), (2, 8, 2), (16, 2, 4)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_bruteforce_dft(oracle, shape, sign):
    x = _rand(shape, 1)
    assert np.allclose(oracle.fft3d(x, sign), oracle.dft3d_bruteforce(x, sign), rtol=0, atol=1e-11)


def test_bruteforce_dft_matches_closed_form(oracle):
    # a single plane wave exp(+2 pi i m.x/n) transforms (sign -1) to n^3 at index m
    nz, ny, nx = 4, 8, 4
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = (1, 3, 2)
    f = np.exp(2j * np.pi * (m[0] * z / nz + m[1] * y / ny + m[2] * x / nx))
    F = oracle.dft3d_bruteforce(f, -1)
    expect = np.zeros_like(F)
    expect[m] = nz * ny * nx
    assert np.allclose(F, expect, atol=1e-10)


@pytest.mark.parametrize("shape", [(128, 16, 16), (64, 32, 32), (256, 16, 16)])
def test_fft_matches_numpy(oracle, shape):
    x = _rand((2,) + shape, 2)
    assert np.allclose(oracle.fft3d(x, -1), np.fft.fftn(x, axes=(1, 2, 3)), atol=1e-9)
    n = np.prod(shape)
    assert np.allclose(oracle.fft3d(x, +1), np.fft.ifftn(x, axes=(1, 2, 3)) * n, atol=1e-9)


def test_fft_identities(oracle):
    x = _rand((32, 8, 16), 3)
    n = x.size
    X = oracle.fft3d(x, -1)
    assert abs(np.sum(np.abs(X) ** 2) - n * np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(X) ** 2)   # Parseval
    assert np.allclose(oracle.fft3d(X, +1) / n, x, atol=1e-12)                                      # round trip
    d = np.zeros((8, 4, 4), complex)
    d[0, 0, 0] = 1.0
    assert np.allclose(oracle.fft3d(d, -1), np.ones_like(d))                                        # delta -> constant


def test_axis_masks_compose(oracle):
    x = _rand((2, 16, 8, 8), 4)
    xy = oracle.fft_axes(x, -1, 3)
    full = oracle.fft_axes(xy, -1, 4)
    assert np.allclose(full, oracle.fft3d(x, -1), atol=1e-10)


def test_kspace_convention(oracle):
    """to_kspace = INVERSE + 1/cells, to_xspace = FORWARD: a round trip is the identity."""
    c = oracle.derive(oracle.model(N=30000, nvz=64, ne=1))
    x = _rand((64, 16, 16), 5)
    k = oracle.to_kspace(c, x)
    assert np.allclose(k, np.fft.ifftn(x), atol=1e-14)
    assert np.allclose(oracle.to_xspace(c, k), x, atol=1e-12)
