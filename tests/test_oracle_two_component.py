"""CPU: pins the oracle's two-component ground state (the coupled imaginary-time
GPEs of docs/wigner_simulation.report/report.tex:200-211).

- H_a a and H_b b are the gradients of the total energy functional, checked by
  central differences (a coupling written with the wrong density, or a missing
  factor, fails this).
- With g12 = 0 each component solves the one-component lattice eigenproblem,
  checked with the independent one-component residual.
- With g11 = g22 = g12 the components coincide and their sum solves the
  one-component problem with N atoms.
"""
import numpy as np
import pytest

from tests.helpers import config1_model


def _calc(oracle, **kw):
    base = config1_model(oracle, ne=1)
    base.update(kw)
    return oracle.derive(oracle.model(**base))


def _smooth_fields(c, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(-1, 1, c.nvz)[:, None, None]
    y = np.linspace(-1, 1, c.nvy)[None, :, None]
    x = np.linspace(-1, 1, c.nvx)[None, None, :]
    env = np.exp(-(x ** 2 + y ** 2) * 2 - z ** 2 * 3)
    a = env * (3 + rng.normal(size=(c.nvz, c.nvy, c.nvx)) * 0.3)
    b = env * (2 + rng.normal(size=(c.nvz, c.nvy, c.nvx)) * 0.3)
    return a.astype(complex), b.astype(complex)


def test_h2_is_the_energy_gradient(oracle):
    c = _calc(oracle, nvz=64)
    a, b = _smooth_fields(c, 1)
    da, db = _smooth_fields(c, 2)
    ha, hb = oracle.apply_h2(c, a, b)

    def F(a_, b_):   # total energy / dV (energy2 is per particle over sum |psi|^2)
        return oracle.energy2(c, a_, b_) * float(np.sum(np.abs(a_) ** 2 + np.abs(b_) ** 2))

    eps = 1e-4
    for d, h, which in ((da, ha, 0), (db, hb, 1)):
        if which == 0:
            fd = (F(a + eps * d, b) - F(a - eps * d, b)) / (2 * eps)
        else:
            fd = (F(a, b + eps * d) - F(a, b - eps * d)) / (2 * eps)
        exact = 2 * float(np.real(np.vdot(d, h)))
        assert fd == pytest.approx(exact, rel=1e-7), (which, fd, exact)


def test_decoupled_components_solve_one_component_problems(oracle):
    c = _calc(oracle, nvz=64)
    c.g12 = 0.0
    a, b, rep = oracle.ground_state2(c, frac_a=0.4, tol=1e-9, max_iter=5000)
    assert rep.converged and rep.residual < 1e-9
    assert rep.N_a == pytest.approx(0.4 * c.N, rel=1e-12)
    assert rep.N_b == pytest.approx(0.6 * c.N, rel=1e-12)
    ra, mua = oracle.residual(c, a)
    cb = _calc(oracle, nvz=64)
    cb.g11 = c.g22
    rb, mub = oracle.residual(cb, b)
    assert ra < 2e-9 and rb < 2e-9
    assert mua == pytest.approx(rep.mu_a, rel=1e-12)
    assert mub == pytest.approx(rep.mu_b, rel=1e-12)
    # the component with more atoms and the same trap sits higher in mu
    assert rep.mu_b > rep.mu_a


def test_symmetric_couplings_give_equal_components(oracle):
    c = _calc(oracle, nvz=64)
    c.g22 = c.g12 = c.g11
    a, b, rep = oracle.ground_state2(c, tol=1e-9, max_iter=5000)
    assert rep.converged
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-7
    r, mu = oracle.residual(c, np.sqrt(2.0) * a)
    assert r < 1e-8
    assert mu == pytest.approx(rep.mu_a, rel=1e-9)


def test_immiscible_pair_separates(oracle):
    """a12 = 100 a0 > sqrt(a11 a22): the smaller-a component forms the core."""
    c = _calc(oracle, a12=100.0, nvz=64)
    a, b, rep = oracle.ground_state2(c, tol=1e-7, max_iter=5000)
    assert rep.converged
    na, nb = np.abs(a) ** 2, np.abs(b) ** 2
    mid = slice(c.nvz // 2 - 8, c.nvz // 2 + 8)
    assert nb[mid].sum() / nb.sum() > na[mid].sum() / na.sum() + 0.2
    # mirror symmetry of the trap survives
    assert np.abs(na - na[::-1]).max() < 1e-6 * na.max()
