"""CPU: pins the oracle's calculateAverages (orc_averages, wigner.cu:305-337) with
closed forms: synthetic ensembles whose quadratures take the values 0..ne-1
(variance (ne^2 - 1)/12) and +-1/2 (variance 1/4), identical trajectories
(variance 0), and the vacuum noise of initializeEnsembles (variance 1/(4 dV)
per quadrature, times (ne - 1)/ne for the population variance)."""
import numpy as np
import pytest

from tests.helpers import config1_model, oracle_calc


def _calc(oracle, ne):
    return oracle_calc(oracle, **config1_model(oracle, nvz=64, ne=ne))


def test_synthetic_ensemble_closed_form(oracle):
    ne = 6
    c = _calc(oracle, ne)
    j = np.arange(ne, dtype=float)[:, None]
    a = np.broadcast_to(j + 2j * j, (ne, c.cells)).astype(complex)
    b = np.broadcast_to(0.5 * (-1.0) ** j + 0j, (ne, c.cells)).astype(complex)
    out = oracle.averages(c, a, b)
    dV = c.dx * c.dy * c.dz
    sa = c.cells * np.sum(5 * j ** 2)
    sb = c.cells * ne * 0.25
    norm = c.N / ((sa + sb) / ne)
    assert out["N"] == pytest.approx((sa + sb) * dV / ne, rel=1e-12)
    assert out["norm"] == pytest.approx(norm, rel=1e-12)
    assert out["var_re_a"] == pytest.approx(norm * (ne ** 2 - 1) / 12, rel=1e-12)
    assert out["var_im_a"] == pytest.approx(norm * 4 * (ne ** 2 - 1) / 12, rel=1e-12)
    assert out["var_re_b"] == pytest.approx(norm * 0.25, rel=1e-12)
    assert out["var_im_b"] == 0.0


def test_identical_trajectories_have_no_spread(oracle):
    ne = 3
    c = _calc(oracle, ne)
    rng = np.random.default_rng(0)
    one = rng.standard_normal(c.cells) + 1j * rng.standard_normal(c.cells)
    out = oracle.averages(c, np.tile(one, (ne, 1)), np.tile(2 * one, (ne, 1)))
    for k in ("var_re_a", "var_im_a", "var_re_b", "var_im_b"):
        assert abs(out[k]) < 1e-25 * out["norm"]


def test_vacuum_noise_variance(oracle):
    ne = 16
    c = _calc(oracle, ne)
    dV = c.dx * c.dy * c.dz
    gs = np.zeros(c.cells, complex)
    a, b = oracle.init_ensembles(c, gs, 0x5EED)
    out = oracle.averages(c, a, b)
    expect = 0.25 / dV * (ne - 1) / ne * out["norm"]
    for k in ("var_re_a", "var_im_a", "var_re_b", "var_im_b"):
        assert out[k] == pytest.approx(expect, rel=0.02), k
