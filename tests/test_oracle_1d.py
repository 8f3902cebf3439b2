"""CPU: pins the oracle's restatement of the 1D program
(programs/bec_wigner_2d.cpp/wigner_1d.cpp, orc1d_run / orc1d_residual / orc_dft1d)
with closed forms.

- Its transform is FFTW's unnormalised DFT of any length (the program's lattice
  has 41 points, a prime): checked against numpy.fft and the defining sum.
- Its Hamiltonian -d^2/dz^2 + z^2/4 on the program's lattice (dk = pi/zmax,
  dz = 2 zmax/(nv - 1)) has the exact ground state exp(-z^2/(4 s)), s = nv/(nv-1),
  with eigenvalue s/2.
- Its imaginary-time loop (fixed step count, itmax = 3 midpoint) converges to
  that state as O(dt^2).
- With g = 0, no noise and detuning E, the Ramsey sequence (U(pi/2), steps, U^+
  at measurement) gives N1 = N0 cos^2(E t/2), N2 = N0 sin^2(E t/2) and
  visibility 1 at t = (k + 1) dt.
- With vacuum noise, g = 0 and no detuning, U^+ undoes U, so the ensemble
  means are E[N1] = N0 + nv (Va^2/2 - (Va + Vb)/4) and E[N2] = nv (Vb^2/2 -
  (Va + Vb)/4): the program draws 1/(2 dz) per mode and subtracts (Va + Vb)/4
  per cell.  Checked within 4 standard errors.
"""
import numpy as np
import pytest


@pytest.mark.parametrize("n", [1, 5, 41, 64, 113])
def test_dft_any_length(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.dft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    assert np.allclose(oracle.dft1d(x, +1), n * np.fft.ifft(x), rtol=0, atol=1e-12 * n)
    if n <= 41:
        j = np.arange(n)
        brute = np.array([np.sum(x * np.exp(-2j * np.pi * j * k / n)) for k in range(n)])
        assert np.allclose(oracle.dft1d(x, -1), brute, rtol=0, atol=1e-12 * n)


def test_lattice_size_follows_the_program(oracle):
    assert oracle.nv1d(oracle.params1d()) == 41
    assert oracle.nv1d(oracle.params1d(zmax=10.0, dz=0.3)) == 67
    assert oracle.nv1d(oracle.params1d(zmax=14.0, dz=0.25)) == 113


def _z(p, nv):
    return -p.zmax + p.dz * np.arange(nv)


def test_lattice_ground_state_closed_form(oracle):
    p = oracle.params1d(zmax=14.0, dz=0.25, g11=0.0, g12=0.0, g22=0.0)
    nv = oracle.nv1d(p)
    s = nv / (nv - 1)
    z = _z(p, nv)
    r, mu = oracle.residual1d(p, np.exp(-z ** 2 / (4 * s)).astype(complex))
    assert r < 1e-9
    assert mu == pytest.approx(s / 2, rel=1e-10)
    r1, _ = oracle.residual1d(p, np.exp(-z ** 2 / 4).astype(complex))
    assert r1 > 1e-3


def test_imaginary_time_loop_converges_quadratically(oracle):
    res, ntot = [], None
    for dt in (0.04, 0.02):
        p = oracle.params1d(zmax=14.0, dz=0.25, g11=0.0, g12=0.0, g22=0.0, mu=0.5 * 113 / 112, dt=dt,
                            ntG=int(40 / dt), ntW=1, ne=1, Va=0.0, Vb=0.0)
        out = oracle.run1d(p)
        res.append(out["report"].gs_residual)
        ntot = out["ntot"]
    assert res[0] < 5e-3
    assert 3.5 < res[0] / res[1] < 4.5
    # the N series starts from the uniform n0 = 1 state: N = nv dz
    assert ntot[0] == pytest.approx(113 * 0.25, rel=1e-12)


@pytest.mark.parametrize("E", [0.0, 1.7, -3.1])
def test_ramsey_fringe_closed_form(oracle, E):
    nt, dt = 12, 0.05
    p = oracle.params1d(zmax=10.0, dz=0.5, g11=0.0, g12=0.0, g22=0.0, E=E, mu=0.5, dt=dt, ntG=200, ntW=nt, ne=2,
                        Va=0.0, Vb=0.0)
    out = oracle.run1d(p)
    N0 = out["m"][:, 0].sum()
    # ntW - 1 equilibration steps leave the single-component state unchanged in number
    assert N0 == pytest.approx(out["report"].N0, rel=1e-9)
    t = (np.arange(nt) + 1) * dt
    assert np.allclose(out["m"][0], N0 * np.cos(E * t / 2) ** 2, rtol=1e-9, atol=1e-9 * N0)
    assert np.allclose(out["m"][1], N0 * np.sin(E * t / 2) ** 2, rtol=1e-9, atol=1e-9 * N0)
    assert np.allclose(out["vis"], 1.0, atol=1e-10)
    assert np.allclose(out["m"].sum(axis=0), N0, rtol=1e-11)


@pytest.mark.parametrize("Va,Vb", [(1.0, 1.0), (0.5, 1.0)])
def test_noise_and_vacuum_subtraction_of_the_ensemble(oracle, Va, Vb):
    kw = dict(zmax=10.0, dz=0.5, g11=0.0, g12=0.0, g22=0.0, E=0.0, mu=0.5, dt=0.05, ntG=200, ntW=1)
    nv = 41
    N0 = oracle.run1d(oracle.params1d(ne=2, Va=0.0, Vb=0.0, **kw))["m"][0, 0]
    out = oracle.run1d(oracle.params1d(ne=400, Va=Va, Vb=Vb, **kw))
    (N1, N2), (e1, e2) = out["m"][:, 0], out["var"][:, 0]
    exp1 = N0 + nv * (Va ** 2 / 2 - (Va + Vb) / 4)
    exp2 = nv * (Vb ** 2 / 2 - (Va + Vb) / 4)
    assert e1 > 0 and e2 > 0
    assert abs(N1 - exp1) < 4 * e1, (N1, exp1, e1)
    assert abs(N2 - exp2) < 4 * e2, (N2, exp2, e2)


def test_single_trajectory_error_is_nan_as_written(oracle):
    # sqrt(var / (ne - 1)) with ne = 1 (wigner_1d.cpp:405-410) is 0/0
    out = oracle.run1d(oracle.params1d(ntW=3, ne=1, tmaxGP=0.1))
    assert np.all(np.isnan(out["var"]))
    assert np.all(np.isfinite(out["m"]))


def test_ground_state_is_stationary_in_real_time(oracle):
    # e^{+i k^2 dt/2} and e^{+i (V - mu) dt} (Q1) evolve with the same H as the
    # imaginary-time loop relaxes with, so with g = 0 the ground state after
    # U(pi/2) keeps its density: |a|^2 = |psi_0|^2 / 2 after 199 steps.  A
    # flipped kinetic sign moves the density at O(1).
    p = oracle.params1d(zmax=14.0, dz=0.25, g11=0.0, g12=0.0, g22=0.0, E=0.0, mu=0.5 * 113 / 112, dt=0.01,
                        ntG=4000, ntW=100, ne=2, Va=0.0, Vb=0.0)
    out = oracle.run1d(p)
    d0 = np.abs(out["gs"]) ** 2
    assert np.abs(out["avn"] - d0 / 2).max() < 1e-4 * d0.max()
