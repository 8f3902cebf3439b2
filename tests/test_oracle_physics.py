"""CPU: pins the oracle's split-step physics with closed forms and invariants
(programs/bec_wigner_3d.cuda/wigner_kernel.cu, wigner.cu).  Each test is chosen
so that a plausible slip -- a dropped term, a flipped sign, a wrong k index or
dk, a transposed pulse -- fails it."""
import math

import numpy as np
import pytest

from tests.helpers import config1_model


def _calc(oracle, **kw):
    return oracle.derive(oracle.model(**config1_model(oracle, **kw)))


def _wide_harmonic(oracle, g=0.0):
    """A 64x32x32 lattice with a box wide and fine enough for spectral accuracy."""
    c = _calc(oracle, nvx=32, nvy=32, nvz=64, ne=1)
    c.xmax = c.ymax = 8.0
    c.zmax = 23.5
    c.dx, c.dy, c.dz = 2 * c.xmax / 31, 2 * c.ymax / 31, 2 * c.zmax / 63
    c.dkx, c.dky, c.dkz = math.pi / c.xmax, math.pi / c.ymax, math.pi / c.zmax
    c.g11 = g
    return c


def _grids(c):
    x = -c.xmax + c.dx * np.arange(c.nvx)
    y = -c.ymax + c.dy * np.arange(c.nvy)
    z = -c.zmax + c.dz * np.arange(c.nvz)
    return np.meshgrid(z, y, x, indexing="ij")


def test_harmonic_ground_state_closed_form(oracle):
    """With g = 0 the lattice Hamiltonian (dk = pi/xmax, dx = 2 xmax/(n-1): quirk Q2)
    is an oscillator of frequency s = n/(n-1) in x' = x/s, so the ground state is
    exp(-x^2/(2 s_x) - y^2/(2 s_y) - z^2/(2 lambda s_z)) with
    mu = (s_x + s_y + s_z/lambda)/2."""
    c = _wide_harmonic(oracle)
    sx, sy, sz = c.nvx / (c.nvx - 1), c.nvy / (c.nvy - 1), c.nvz / (c.nvz - 1)
    Z, Y, X = _grids(c)
    psi = np.exp(-X ** 2 / (2 * sx) - Y ** 2 / (2 * sy) - Z ** 2 / (2 * c.lambda_ * sz)).astype(complex)
    r, mu = oracle.residual(c, psi)
    assert r < 1e-8                       # box truncation and aliasing of the Gaussian
    assert mu == pytest.approx((sx + sy + sz / c.lambda_) / 2, rel=1e-10)
    # a wrong dk (2 pi / (n dx), the DFT-consistent one) gives a visibly different state
    c2 = _wide_harmonic(oracle)
    c2.dkx = 2 * math.pi / (c2.nvx * c2.dx)
    assert oracle.residual(c2, psi)[0] > 1e-3


def test_split_step_ground_state_converges_to_closed_form(oracle):
    """Imaginary-time split step (wigner.cu:84-175) with g = 0: the fixed point
    approaches the exact ground state as O(dt^2) in the state (residual) and, mu
    being variational, as O(dt^4) in mu."""
    c = _wide_harmonic(oracle)
    sx, sz = c.nvx / (c.nvx - 1), c.nvz / (c.nvz - 1)
    exact = (2 * sx + sz / c.lambda_) / 2
    Z, Y, X = _grids(c)
    guess = np.exp(-(X ** 2 + Y ** 2) / (2 * sx) - Z ** 2 / (2 * c.lambda_ * sz)).astype(complex)
    errs, res = [], []
    for dt in (0.1, 0.05):
        c.dt_steady = dt
        c.N = 1000
        psi, rep = oracle.ground_state(c, guess=guess, rule=oracle.RULE_REFERENCE, tol=1e-11, max_iter=3000)
        assert rep.converged
        errs.append(abs(rep.mu - exact))
        res.append(rep.residual)
    assert errs[0] < 1e-5 and res[0] < 1e-2
    assert 3.5 < res[0] / res[1] < 4.5       # state: second order in dt
    assert 12 < errs[0] / errs[1] < 20       # mu: fourth order


def test_reference_rule_ground_state_config1(oracle):
    """The reference stopping rule converges and the state holds N atoms; its
    'energy' (quirk Q3: x-space psi times k-space psi, no conjugate) differs from
    <psi|H|psi>."""
    c = _calc(oracle, ne=1)
    psi, rep = oracle.ground_state(c, rule=oracle.RULE_REFERENCE, tol=1e-6)
    assert rep.converged and 100 < rep.iterations < 2000
    assert rep.N == pytest.approx(30000, rel=1e-10)
    assert rep.mu == pytest.approx(c.mu, rel=0.1)          # near the Thomas-Fermi value
    assert abs(rep.E_ref - rep.E) > 1e-3 * abs(rep.E)
    assert rep.residual < 2e-3


def test_gs_iterations_equal_reference_loop(oracle):
    """orc_ground_state and the loop body restated step by step agree."""
    c = _calc(oracle, ne=1)
    psi, rep = oracle.ground_state(c, rule=oracle.RULE_REFERENCE, tol=1e-6)
    tf = oracle.tf_state(c)
    tf = tf * np.sqrt(c.N / (np.sum(np.abs(tf) ** 2) * c.dx * c.dy * c.dz))
    a, E = oracle.gs_iterate(c, oracle.to_kspace(c, tf), rep.iterations)
    assert np.allclose(oracle.to_xspace(c, a), psi, atol=1e-12 * np.abs(psi).max())


def test_kinetic_phase_sign_and_nyquist(oracle):
    """propagateKSpaceRealTime multiplies by exp(+i k^2/2 dt/2) (quirk Q1) with the
    Nyquist index taken positive (getK, wigner_kernel.cu:15-21)."""
    c = _calc(oracle, ne=1, nvz=64)
    dt = 0.37
    for idx in [(0, 0, 0), (5, 3, 2), (32, 8, 8), (63, 15, 9)]:
        a = np.zeros((64, 16, 16), complex)
        a[idx] = 1.0
        a2, _ = oracle.kinetic_rt(c, a, a.copy(), dt)
        k = [(i - n if 2 * i > n else i) * d for i, n, d in zip(idx, (64, 16, 16), (c.dkz, c.dky, c.dkx))]
        e = sum(v * v for v in k) / 2
        assert a2[idx] == pytest.approx(np.exp(1j * e * dt / 2), abs=1e-14)


def test_free_evolution_is_exact(oracle):
    """With V ~ 0 and g = l = 0 kinetic factors commute: n steps of dt equal one step of n dt."""
    c = _calc(oracle, ne=1, nvz=64)
    c.xmax = c.ymax = 1e-9
    c.dx = c.dy = 2e-9 / 15
    c.zmax, c.dz = 1e-9, 2e-9 / 63
    c.lambda_ = 1.0
    c.g11 = c.g12 = c.g22 = 0.0
    c.l111 = c.l12 = c.l22 = 0.0
    c.detuning = 0.0
    c.dkx = c.dky = c.dkz = 0.3
    r = np.random.default_rng(0)
    a = r.standard_normal((64, 16, 16)) + 1j * r.standard_normal((64, 16, 16))
    b = r.standard_normal((64, 16, 16)) + 1j * r.standard_normal((64, 16, 16))
    a1, b1 = oracle.propagate(c, a, b, 0.01, 10)
    a2, b2 = oracle.propagate(c, a, b, 0.1, 1)
    assert np.allclose(a1, a2, atol=1e-12) and np.allclose(b1, b2, atol=1e-12)


def _cells(c, seed):
    r = np.random.default_rng(seed)
    sh = (1, c.nvz, c.nvy, c.nvx)
    return (r.standard_normal(sh) + 1j * r.standard_normal(sh)) * 5, (r.standard_normal(sh) + 1j * r.standard_normal(sh)) * 5


def test_nonlinear_step_invariants(oracle):
    """propagateXSpaceTwoComponent (wigner_kernel.cu:184-229)."""
    c = _calc(oracle, ne=1, nvz=64, detuning=-41.0)
    a, b = _cells(c, 1)
    dt = 0.05
    na, nb = np.abs(a) ** 2, np.abs(b) ** 2
    # losses: |a|^2 decays by exp(-dt (l111 na^2 + l12 nb)), |b|^2 by exp(-dt (l22 nb + l12 na))
    a1, b1 = oracle.nonlinear_rt(c, a, b, dt)
    assert np.allclose(np.abs(a1) ** 2, na * np.exp(-dt * (c.l111 * na ** 2 + c.l12 * nb)), rtol=1e-12)
    assert np.allclose(np.abs(b1) ** 2, nb * np.exp(-dt * (c.l22 * nb + c.l12 * na)), rtol=1e-12)
    # without losses the step is a pure phase; the phase of a is +dt (V + g11 na + g12 nb) (quirk Q1)
    c.l111 = c.l12 = c.l22 = 0.0
    a1, b1 = oracle.nonlinear_rt(c, a, b, dt)
    assert np.allclose(np.abs(a1), np.abs(a), rtol=1e-13) and np.allclose(np.abs(b1), np.abs(b), rtol=1e-13)
    V = oracle.potential(c)[None]
    ph = np.angle(a1 / a)
    want = np.angle(np.exp(1j * dt * (V + c.g11 * na + c.g12 * nb)))
    assert np.allclose(np.exp(1j * ph), np.exp(1j * want), atol=1e-11)
    # detuning shifts only b's phase, by -dt * detuning
    c2 = _calc(oracle, ne=1, nvz=64, detuning=0.0)
    c2.l111 = c2.l12 = c2.l22 = 0.0
    a2, b2 = oracle.nonlinear_rt(c2, a, b, dt)
    assert np.allclose(a2, a1, atol=1e-12)
    assert np.allclose(b1 / b2, np.exp(-1j * dt * c.detuning), atol=1e-11)
    # with |psi| conserved the midpoint iteration (itmax = 2) changes nothing
    c.itmax = 2
    a3, b3 = oracle.nonlinear_rt(c, a, b, dt)
    assert np.allclose(a3, a1, atol=1e-11) and np.allclose(b3, b1, atol=1e-11)
    # itmax = 0 leaves da = 0: zero output (quirk Q10)
    c.itmax = 0
    a4, b4 = oracle.nonlinear_rt(c, a, b, dt)
    assert np.all(a4 == 0) and np.all(b4 == 0)


def test_pulse_algebra(oracle):
    """applyHalfPiPulse twice is applyPiPulse; both are unitary per cell
    (wigner_kernel.cu:320-341)."""
    r = np.random.default_rng(3)
    a = r.standard_normal(1000) + 1j * r.standard_normal(1000)
    b = r.standard_normal(1000) + 1j * r.standard_normal(1000)
    h1 = oracle.half_pi_pulse(*oracle.half_pi_pulse(a, b))
    p1 = oracle.pi_pulse(a, b)
    assert np.allclose(h1[0], p1[0]) and np.allclose(h1[1], p1[1])
    ha, hb = oracle.half_pi_pulse(a, b)
    assert np.allclose(np.abs(ha) ** 2 + np.abs(hb) ** 2, np.abs(a) ** 2 + np.abs(b) ** 2)


def test_virtual_pulse_sends_all_atoms_to_b(oracle):
    """At t = 0 after the pi/2 pulse of initEvolution and halfPiRotate(alpha = 0)
    every atom is in b: component ratio -1, visibility 1 (wigner.cu:257-277)."""
    c = _calc(oracle, ne=1, Va=0.0, Vb=0.0)
    psi = oracle.tf_state(c)
    a, b = oracle.half_pi_pulse(psi[None], np.zeros_like(psi)[None])
    n1, n2 = oracle.rotate_sums(c, a, b, 0.0)
    assert abs(n1) < 1e-12 * abs(n2)
    assert oracle.visibility_grid(c, a, b) == pytest.approx(1.0, abs=1e-12)


def test_vacuum_noise_statistics(oracle):
    """initializeEnsembles (wigner_kernel.cu:295-317): noise of variance 1/(2 dV) per
    mode, independent quadratures and components; vacuum subtraction removes it
    on average (density(), wigner_kernel.cu:283-286)."""
    c = _calc(oracle, ne=4)
    zero = np.zeros((c.nvz, c.nvy, c.nvx), complex)
    a, b = oracle.init_ensembles(c, zero, 0x5EED)
    dV = c.dx * c.dy * c.dz
    for f in (a, b):
        assert abs(np.mean(f)) < 0.02 / math.sqrt(dV)
        assert np.var(f.real) == pytest.approx(0.25 / dV, rel=0.02)
        assert np.var(f.imag) == pytest.approx(0.25 / dV, rel=0.02)
        assert abs(np.corrcoef(f.real.ravel(), f.imag.ravel())[0, 1]) < 0.01
    assert abs(np.corrcoef(a.real.ravel(), b.real.ravel())[0, 1]) < 0.01
    n1, n2 = oracle.rotate_sums(c, a, b, 0.0)
    vac = c.ne * c.cells * c.V / (2 * dV)
    assert abs(n1) < 0.02 * vac and abs(n2) < 0.02 * vac
    # a different seed gives different noise, the same seed the same
    a2, _ = oracle.init_ensembles(c, zero, 0x5EED)
    a3, _ = oracle.init_ensembles(c, zero, 0x5EEE)
    assert np.array_equal(a, a2) and not np.allclose(a, a3)


def test_equilibration_runs_at_least_one_step(oracle):
    """for(t = 0; t <= t_equilib; t += dt_evo) in float (wigner.cu:222-223, quirk Q5)."""
    c = _calc(oracle, ne=1)
    assert oracle.equilibration_steps(c) == 1
    c.t_equilib = 10 * c.dt_evo
    assert oracle.equilibration_steps(c) in (10, 11)


def test_observables_from_sums_match_per_cell(oracle):
    """Na, Nb after any virtual pulse follow from four per-trajectory sums; this is
    how the GPU path evaluates the 13-angle visibility of getVisibility."""
    c = _calc(oracle, ne=2)
    r = np.random.default_rng(5)
    sh = (2, c.nvz, c.nvy, c.nvx)
    a = r.standard_normal(sh) + 1j * r.standard_normal(sh)
    b = r.standard_normal(sh) + 1j * r.standard_normal(sh)
    S = oracle.traj_sums(c, a, b).sum(axis=0)
    vac = 2 * c.cells * c.V / (2 * c.dx * c.dy * c.dz)
    for alpha in (0.0, 0.5, 2.5, 6.0):
        n1, n2 = oracle.rotate_sums(c, a, b, alpha)
        cr = math.sin(alpha) * S[2] + math.cos(alpha) * S[3]
        assert n1 == pytest.approx(0.5 * (S[0] + S[1]) + cr - vac, rel=1e-12)
        assert n2 == pytest.approx(0.5 * (S[0] + S[1]) - cr - vac, rel=1e-12)


def _step(oracle, c, a_k, b_k, dt, kin_sign):
    """One real-time step K/2 N K/2 assembled from the oracle's pieces; kin_sign = -1
    plants a flipped kinetic phase (the mistake the stationarity test must catch)."""
    a_k, b_k = oracle.kinetic_rt(c, a_k, b_k, kin_sign * dt)   # the half step exp(+i (k^2/2) dt/2)
    a_x, b_x = oracle.to_xspace(c, a_k), oracle.to_xspace(c, b_k)
    a_x, b_x = oracle.nonlinear_rt(c, a_x, b_x, dt)
    a_k, b_k = oracle.to_kspace(c, a_x), oracle.to_kspace(c, b_x)
    return oracle.kinetic_rt(c, a_k, b_k, kin_sign * dt)


def test_ground_state_is_stationary_in_real_time_3d(oracle):
    """The interacting 3D ground state (|1> only, g11 at the 30k-atom configuration,
    no pulse, no losses, no noise) keeps its density under the real-time step:
    e^{+i (k^2/2) dt/2} and e^{+i (V + g n) dt} (quirk Q1) are the same Hamiltonian
    the imaginary-time solvers relax with.  The oracle's propagate keeps |a|^2
    to 2e-4 of its peak over 200 steps (8 ms, most of a radial period); with the
    kinetic sign flipped the density moves by more than 10% of its peak."""
    c = oracle.derive(oracle.model(N=30000, nvx=16, nvy=16, nvz=128, ne=1, a12=0.0, Va=0.0, Vb=0.0,
                                   gamma111=0.0, gamma12=0.0, gamma22=0.0, detuning=0.0))
    # a12 = 0 decouples the two-component solve: its |1> is the one-component ground state of N/2 atoms
    a, b, rep = oracle.ground_state2(c, frac_a=0.5, tol=1e-9, max_iter=5000)
    assert rep.converged
    assert oracle.residual(c, a)[0] < 1.5e-9
    d0 = np.abs(a) ** 2
    zero = np.zeros_like(a)
    a_k, b_k = oracle.to_kspace(c, a), oracle.to_kspace(c, zero)
    ok_k = oracle.propagate(c, a_k, b_k, c.dt_evo, 200)[0]
    moved_ok = np.abs(np.abs(oracle.to_xspace(c, ok_k)) ** 2 - d0).max() / d0.max()
    fa, fb = a_k, b_k
    for _ in range(200):
        fa, fb = _step(oracle, c, fa, fb, c.dt_evo, -1)
    moved_flip = np.abs(np.abs(oracle.to_xspace(c, fa)) ** 2 - d0).max() / d0.max()
    # the assembled step with the right sign is the oracle's propagate
    ra, rb = a_k, b_k
    for _ in range(3):
        ra, rb = _step(oracle, c, ra, rb, c.dt_evo, +1)
    pa = oracle.propagate(c, a_k, b_k, c.dt_evo, 3)[0]
    print(f"density change over 200 steps: {moved_ok:.2e} (right sign), {moved_flip:.2e} (flipped kinetic sign)")
    assert np.abs(ra - pa).max() < 1e-12 * np.abs(pa).max()
    assert moved_ok < 2e-4
    assert moved_flip > 0.1
