"""CPU: pins the oracle's restatement of the 2D program (programs/bec_wigner_2d.cpp,
orc2d_run / orc2d_residual) with closed forms.

- Its Hamiltonian -lap + r^2/4 on the reference's lattice (dk = pi/xmax,
  dx = 2 xmax/(n-1)) has the exact ground state exp(-r^2/(4 s)), s = n/(n-1),
  with eigenvalue s; the DFT-consistent dk fails.
- Its imaginary-time loop (fixed step count, itmax = 3 midpoint) converges to
  that state as O(dt^2).
- With g = 0, no noise and detuning E, the Ramsey sequence (U(pi/2), steps, U^+
  at measurement, wigner.cpp:341-353,393-425) gives N1 = N0 cos^2(E t / 2),
  N2 = N0 sin^2(E t / 2) and visibility 1 at t = (k + 1) dtW after the pulse:
  the two components see the same linear Hamiltonian, so only their relative
  phase moves.  A sign slip in either pulse or a lost unitarity fails it.
- With vacuum noise and g = 0 (and no detuning) U^+ undoes U, so at the
  measurement b1 is the initial noise of b evolved unitarily.  The 2D program
  draws noise of variance Vc^2/dV per quadrature and subtracts (Va + Vb)/4 per
  cell (quirk Q12: the two do not cancel), so the ensemble means are
  E[N2] = n (2 Vb^2 - (Va + Vb)/4) and E[N1] = N0 + n (2 Va^2 - (Va + Vb)/4)
  over n cells, checked within 4 standard errors.
"""
import numpy as np
import pytest


def _grid(p):
    x = -p.xmax + 2 * p.xmax / (p.nvx - 1) * np.arange(p.nvx)
    y = -p.ymax + 2 * p.ymax / (p.nvy - 1) * np.arange(p.nvy)
    Y, X = np.meshgrid(y, x, indexing="ij")
    return X, Y


def test_lattice_ground_state_closed_form(oracle):
    p = oracle.params2d(nvx=64, nvy=64, xmax=14.0, ymax=14.0, g11=0.0, g12=0.0, g22=0.0)
    s = 64 / 63
    X, Y = _grid(p)
    psi = np.exp(-(X ** 2 + Y ** 2) / (4 * s)).astype(complex)
    r, mu = oracle.residual2d(p, psi)
    assert r < 1e-9
    assert mu == pytest.approx(s, rel=1e-10)
    # the Gaussian of the continuum (s = 1) is not the lattice eigenstate
    r1, _ = oracle.residual2d(p, np.exp(-(X ** 2 + Y ** 2) / 4).astype(complex))
    assert r1 > 1e-3


def test_imaginary_time_loop_converges_quadratically(oracle):
    s = 64 / 63
    res = []
    for dt in (0.04, 0.02):
        p = oracle.params2d(nvx=64, nvy=64, xmax=14.0, ymax=14.0, g11=0.0, g12=0.0, g22=0.0, mu=s,
                            dtG=dt, ntG=int(60 / dt), ne=1, n_meas=1, ntW=1, Va=0.0, Vb=0.0)
        out = oracle.run2d(p)
        res.append(out["report"].gs_residual)
    assert res[0] < 5e-3
    assert 3.5 < res[0] / res[1] < 4.5


@pytest.mark.parametrize("E", [0.0, 1.7, -3.1])
def test_ramsey_fringe_closed_form(oracle, E):
    nm, dtW = 12, 0.05
    p = oracle.params2d(nvx=32, nvy=32, xmax=10.0, ymax=10.0, g11=0.0, g12=0.0, g22=0.0, E=E, mu=1.0,
                        dtG=0.05, ntG=200, ne=1, n_meas=nm, ntW=1, dtW=dtW, Va=0.0, Vb=0.0)
    out = oracle.run2d(p)
    N0 = out["m"][:, 0].sum()
    t = (np.arange(nm) + 1) * dtW
    assert np.allclose(out["m"][0], N0 * np.cos(E * t / 2) ** 2, rtol=1e-9, atol=1e-9 * N0)
    assert np.allclose(out["m"][1], N0 * np.sin(E * t / 2) ** 2, rtol=1e-9, atol=1e-9 * N0)
    assert np.allclose(out["vis"], 1.0, atol=1e-10)
    # unitarity of the real-time steps: the total is constant
    assert np.allclose(out["m"].sum(axis=0), N0, rtol=1e-11)


@pytest.mark.parametrize("Va,Vb", [(1.0, 1.0), (0.2, 0.5)])
def test_noise_and_vacuum_subtraction_of_the_ensemble(oracle, Va, Vb):
    n = 32 * 32
    kw = dict(nvx=32, nvy=32, xmax=10.0, ymax=10.0, g11=0.0, g12=0.0, g22=0.0, E=0.0, mu=1.0,
              dtG=0.05, ntG=200, n_meas=1, ntW=1, dtW=0.05)
    N0 = oracle.run2d(oracle.params2d(ne=1, Va=0.0, Vb=0.0, **kw))["m"][0, 0]
    out = oracle.run2d(oracle.params2d(ne=256, Va=Va, Vb=Vb, **kw))
    (N1, N2), (e1, e2) = out["m"][:, 0], out["var"][:, 0]   # means and their standard errors
    exp1 = N0 + n * (2 * Va ** 2 - (Va + Vb) / 4)
    exp2 = n * (2 * Vb ** 2 - (Va + Vb) / 4)
    assert e1 > 0 and e2 > 0
    assert abs(N1 - exp1) < 4 * e1, (N1, exp1, e1)
    assert abs(N2 - exp2) < 4 * e2, (N2, exp2, e2)


def test_ground_state_is_stationary_in_real_time(oracle):
    # the 2D counterpart: with g = 0 the lattice ground state keeps its density
    # (|a|^2 = |psi_0|^2 / 2 after U(pi/2)) through the real-time steps, which a
    # flipped kinetic or trap sign would not
    s = 32 / 31
    p = oracle.params2d(nvx=32, nvy=32, xmax=10.0, ymax=10.0, g11=0.0, g12=0.0, g22=0.0, E=0.0, mu=s, dtG=0.02,
                        ntG=2000, ne=1, n_meas=1, ntW=60, dtW=0.01, Va=0.0, Vb=0.0)
    out = oracle.run2d(p)
    d0 = np.abs(out["gs"]) ** 2
    assert np.abs(out["avn"] - d0 / 2).max() < 1e-4 * d0.max()
