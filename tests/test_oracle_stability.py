"""CPU: linear stability of the real-time split step, and where it fails.

For a perturbation u cos(k.x) of a uniform background of density n0, one step
of the split scheme (kinetic rotation by theta = (k^2/2) dt, then the nonlinear
phase exp(i g |psi|^2 dt)) maps (Re u, Im u) by
    M = [[1, 0], [2 a, 1]] [[cos theta, -sin theta], [sin theta, cos theta]],  a = g n0 dt,
with det M = 1 and trace 2 cos theta - 2 a sin theta.  The step amplifies the
mode when |cos theta - a sin theta| > 1, i.e. for theta in (m pi - 2 atan a, m pi).
A lattice is stable at dt when its largest kinetic phase per step stays below
pi - 2 atan(a) at the peak density.

This decides the 128x32x32 lattice of the precision study: at the reference's
default 40 us step (programs/bec_wigner_3d.cuda/wigner.cpp:95) its largest
kinetic phase is 5.2 rad, inside the first unstable band, and a 1e-7
perturbation of the FP64 oracle's state grows by orders of magnitude in a few
hundred steps; at the 20 us the reference's report uses on that lattice
(docs/wigner_simulation.report/report.tex:804-810) the phase is 2.6 rad and the
perturbation stays put.  The 128x16x16 lattice of the headline run is stable at
40 us (1.7 rad)."""
import numpy as np
import pytest

from tests.helpers import oracle_calc


def _step_matrix(theta, a):
    rot = np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])
    return np.array([[1.0, 0.0], [2.0 * a, 1.0]]) @ rot


def _theta_max(c, dt):
    kx = np.pi / c.xmax * (c.nvx // 2)
    ky = np.pi / c.ymax * (c.nvy // 2)
    kz = np.pi / c.zmax * (c.nvz // 2)
    return 0.5 * (kx * kx + ky * ky + kz * kz) * dt


def _band_edge(a):
    return np.pi - 2.0 * np.arctan(a)


def test_linear_band_closed_form():
    """The spectral radius of M exceeds 1 exactly inside (m pi - 2 atan a, m pi)."""
    for a in (0.02, 0.115, 0.3):
        th = np.linspace(0.0, 3 * np.pi, 30001)
        rho = np.array([max(abs(np.linalg.eigvals(_step_matrix(t, a)))) for t in th])
        unstable = rho > 1.0 + 1e-9
        inside = np.zeros_like(th, dtype=bool)
        for m in (1, 2, 3):
            inside |= (th > m * np.pi - 2 * np.arctan(a) + 1e-6) & (th < m * np.pi - 1e-6)
        edge = np.zeros_like(th, dtype=bool)
        for m in (1, 2, 3):
            for e in (m * np.pi - 2 * np.arctan(a), m * np.pi):
                edge |= np.abs(th - e) < 1e-3
        assert np.array_equal(unstable[~edge], inside[~edge])
        # the largest growth per step is sqrt(1 + a^2) + a, at theta = pi - atan a
        assert rho.max() == pytest.approx(np.sqrt(1 + a * a) + a, rel=1e-6)


@pytest.mark.parametrize("nxy,nz,N,dt,stable", [(16, 128, 30000, 4e-5, True), (32, 128, 30000, 4e-5, False),
                                                (32, 128, 30000, 2e-5, True), (32, 256, 70000, 4e-5, False),
                                                (32, 256, 70000, 2e-5, True)])
def test_lattice_stability_criterion(oracle, nxy, nz, N, dt, stable):
    c = oracle_calc(oracle, N=N, nvx=nxy, nvy=nxy, nvz=nz, ne=1, dt_evo=dt)
    a = c.mu * c.dt_evo   # peak g n dt of the Thomas-Fermi cloud
    assert (_theta_max(c, c.dt_evo) < _band_edge(a)) == stable


def _growth(oracle, c, gs, pert, steps):
    out = []
    for g0 in (gs, pert):
        a, b = oracle.init_evolution(c, g0, 1, equilibrate=False)
        a, b = oracle.propagate(c, a, b, c.dt_evo, steps)
        out.append(np.concatenate([a.ravel(), b.ravel()]))
    return np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0])


def test_oracle_perturbation_growth_128x32x32(oracle):
    """The FP64 oracle itself: at 40 us a 1e-7 perturbation of the ground state
    grows by more than 100x within 250 steps; at 20 us it stays within 2x."""
    kw = dict(N=30000, nvx=32, nvy=32, nvz=128, ne=1, Va=0.0, Vb=0.0, gamma111=0.0, gamma12=0.0, gamma22=0.0)
    c40 = oracle_calc(oracle, dt_evo=4e-5, **kw)
    c20 = oracle_calc(oracle, dt_evo=2e-5, **kw)
    gs, rep = oracle.ground_state(c40, rule=oracle.RULE_REFERENCE, tol=1e-6, max_iter=3000)
    rng = np.random.default_rng(3)
    pert = gs * (1 + 1e-7 * (rng.standard_normal(gs.shape) + 1j * rng.standard_normal(gs.shape)))
    d0 = np.linalg.norm(pert - gs) / np.linalg.norm(gs)
    g40 = _growth(oracle, c40, gs, pert, 250) / d0
    g20 = _growth(oracle, c20, gs, pert, 250) / d0
    print(f"perturbation growth over 250 steps: 40 us {g40:.3g}, 20 us {g20:.3g}")
    assert g40 > 100.0
    assert g20 < 2.0
