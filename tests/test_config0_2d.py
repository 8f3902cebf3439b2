"""CPU: the 2D reduction of the Rb-87 cloud (bec_fill_2d_params) and the oracle's
renormalised gradient-flow ground state of the 2D program (gs_rule 2).

The 2D program (programs/bec_wigner_2d.cpp/wigner.cpp:106-194) works with the
Hamiltonian -lap + r^2/4 + g |psi|^2 in dimensionless constants.  The
configuration "64x64, Rb-87 a11 = 100.4 a0, a22 = 95.0 a0, a12 = 97.66 a0,
N = 30k, ground state to residual 1e-8, pi/2 pulse, 2000 steps of 10 us" is
mapped by taking the trap's radial plane (97.6 Hz) as the program's plane and
freezing the axial direction (11.96 Hz) in its oscillator ground state:
length unit sqrt(hbar / (2 m w)), g = 4 sqrt(2 pi) a / a_z.  The mapping is
checked here against the SI constants computed independently; the flow
against closed forms (the lattice oscillator state when g -> 0, the
Thomas-Fermi limit, N held exactly)."""
import numpy as np
import pytest

import bec_amd as B

HBAR, A0, M = 1.054571628e-34, 5.2917720859e-11, 1.443160648e-25


def _model(**kw):
    d = dict(N=30000, nvx=64, nvy=64, nvz=2, ne=1, Va=0.0, Vb=0.0, detuning=0.0, dt_evo=1e-5, dt_steady=1e-5)
    d.update(kw)
    return B.default_model(**d)


def _oracle_params(oracle, p, **kw):
    """The same parameters in the oracle's structure."""
    names = [f for f, _ in oracle.Params2D._fields_]
    d = {f: getattr(p, f) for f in names}
    d.update(kw)
    return oracle.Params2D(**d)


def test_mapping_constants():
    p = B.params2d_from_model(_model())
    w = 2 * np.pi * 97.6
    a_z = np.sqrt(HBAR / (M * 2 * np.pi * 11.96))
    l = np.sqrt(HBAR / (2 * M * w))
    # g = g2D / (hbar w l^2), g2D = 4 pi hbar^2 a / m / (sqrt(2 pi) a_z)
    for a, g in ((100.4, p.g11), (97.66, p.g12), (95.0, p.g22)):
        g2d = 4 * np.pi * HBAR ** 2 * a * A0 / M / (np.sqrt(2 * np.pi) * a_z)
        assert g == pytest.approx(g2d / (HBAR * w * l * l), rel=1e-12)
    assert p.g11 == pytest.approx(0.0170828, rel=1e-5)
    # Thomas-Fermi: N = int (mu - r^2/4)/g d^2r = 2 pi mu^2 / g
    assert 2 * np.pi * p.mu ** 2 / p.g11 == pytest.approx(30000, rel=1e-12)
    assert p.xmax == p.ymax == pytest.approx(1.2 * 2 * np.sqrt(p.mu), rel=1e-14)
    assert p.dtW == pytest.approx(w * 1e-5, rel=1e-14) and p.dtG == pytest.approx(w * 1e-5, rel=1e-14)
    assert (p.nvx, p.nvy, p.gs_rule, p.n_atoms, p.gs_tol, p.ntW, p.n_meas) == (64, 64, 2, 30000, 1e-8, 1, 1)
    # the kinetic phase per step stays far from the split step's unstable band at pi
    k = np.pi / p.xmax * 32
    assert 2 * k * k * p.dtW < 3.0


def test_flow_reaches_the_lattice_oscillator_state(oracle):
    """g -> 0: the flow converges to the lattice eigenstate exp(-r^2/(4 s)),
    s = n/(n-1), with N atoms and residual 1e-8."""
    # mu is the preconditioner shift here (and sets the Thomas-Fermi start)
    p = oracle.params2d(nvx=64, nvy=64, xmax=10.0, ymax=10.0, g11=1e-9, g12=0.0, g22=0.0, mu=20.0, ne=1,
                        n_meas=1, ntW=1, Va=0.0, Vb=0.0, gs_rule=2, gs_tol=1e-8, gs_max_iter=5000, n_atoms=500.0)
    out = oracle.run2d(p)
    rep = out["report"]
    assert rep.gs_residual < 1e-8 and rep.gs_iterations < 5000
    assert rep.N0 == pytest.approx(500.0, rel=1e-12)
    s = 64 / 63
    x = -10.0 + 20.0 / 63 * np.arange(64)
    Y, X = np.meshgrid(x, x, indexing="ij")
    ref = np.exp(-(X ** 2 + Y ** 2) / (4 * s))
    ref *= np.sqrt(500.0 / ((ref ** 2).sum() * (20.0 / 63) ** 2))
    gs = out["gs"]
    assert np.abs(gs - ref).max() / ref.max() < 1e-7


def test_flow_config0_holds_n_and_approaches_thomas_fermi(oracle):
    """The mapped Rb-87 configuration: N = 30000 exactly, residual 1e-8, the
    chemical potential within 5% above the Thomas-Fermi value (kinetic energy),
    the oracle's own residual of the returned state agrees."""
    p = _oracle_params(oracle, B.params2d_from_model(_model()), n_meas=1)
    out = oracle.run2d(p)
    rep = out["report"]
    assert rep.gs_residual < 1e-8
    assert rep.N0 == pytest.approx(30000, rel=1e-10)
    r, mu = oracle.residual2d(p, out["gs"])
    assert r < 1e-8
    assert 1.0 < mu / p.mu < 1.05
    # no noise, U^+ at measurement: right after the first step nearly all atoms are back in |1>
    assert out["m"][0, 0] == pytest.approx(30000, rel=1e-3)
    assert out["vis"][0] == pytest.approx(1.0, abs=1e-6)
