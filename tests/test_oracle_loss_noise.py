"""CPU: pins the oracle's loss-induced Wigner noise (orc_nonlinear_rt_noise) against
exact quantum dynamics.

The reference derives the diffusion of its loss terms
(docs/wigner_simulation.report/report.tex:624-656, docs/icap2010.poster/poster.tex:245-290)
but never integrates it.  The local step with noise acts on every cell as an
independent mode, with amplitude alpha = psi sqrt(dV).  So its ensemble
statistics can be checked against the exact master equation of one or two
modes, solved in Fock space:

- two-body loss in b, rate kappa = l22 / (2 dV): dP_n/dt = kappa [(n+2)(n+1) P_{n+2} - n(n-1) P_n];
- three-body loss in a, kappa3 = l111 / (2 dV^2): events at rate 2 kappa3/3 n(n-1)(n-2);
- a-b loss, kappa = l12 / (2 dV): one atom of each at rate 2 kappa n_a n_b.

The initial state is coherent, sampled with Wigner noise <|delta alpha|^2> = 1/2.
Number moments follow from the symmetric ordering: <n> = <|alpha|^2> - 1/2,
Var n = Var|alpha|^2 - 1/4, Cov(n_a, n_b) = Cov(|alpha|^2, |beta|^2).

With the noise, the number variance follows the exact one.  Without it, the
variance collapses to a fraction of the exact value, so the magnitude of the
diffusion is pinned.  The a-b channel removes atoms in pairs, so Var(n_a - n_b)
is conserved; that pins the cross term D12.  The leading-order drift that the
kernels integrate (the reference's) biases the mean by O(1/n).  Tolerances are
written per check.
"""
import numpy as np
import pytest


def _calc(oracle, **loss):
    c = oracle.derive(oracle.model(nvx=16, nvy=16, nvz=64, ne=4))
    c.g11 = c.g12 = c.g22 = 0.0
    c.detuning = 0.0
    c.itmax = 1
    c.l111, c.l12, c.l22 = loss.get("l111", 0.0), loss.get("l12", 0.0), loss.get("l22", 0.0)
    return c


def _dV(c):
    return c.dx * c.dy * c.dz


def _coherent(rng, n0, shape):
    return np.sqrt(n0) + 0.5 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def _rk4(P, rhs, dt, steps):
    for _ in range(steps):
        k1 = rhs(P)
        k2 = rhs(P + dt / 2 * k1)
        k3 = rhs(P + dt / 2 * k2)
        k4 = rhs(P + dt * k3)
        P = P + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return P


def _poisson(n0, nmax):
    from scipy.stats import poisson
    return poisson.pmf(np.arange(nmax + 1), n0)


def _evolve(oracle, c, a, b, dt, steps, noise, seed=0x10550):
    ax, bx = a.copy(), b.copy()
    for s in range(steps):
        if noise:
            ax, bx = oracle.nonlinear_rt_noise(c, ax, bx, dt, seed, s)
        else:
            ax, bx = oracle.nonlinear_rt(c, ax, bx, dt)
    return ax, bx


def _moments(w):
    # <n>, Var n of one mode from the Wigner samples of |alpha|^2
    return w.mean() - 0.5, w.var() - 0.25


def test_two_body_loss_one_mode(oracle):
    n0, kappa, dt, steps = 200.0, 1.0 / 2000.0, 0.0125, 400
    c = _calc(oracle, l22=1.0)
    dV = _dV(c)
    c.l22 = 2 * kappa * dV
    M = c.cells * c.ne
    rng = np.random.default_rng(1)
    b = _coherent(rng, n0, M) / np.sqrt(dV)
    a = np.zeros(M, complex)
    n = np.arange(int(n0 + 12 * np.sqrt(n0)) + 1)

    def rhs(P):
        d = -kappa * n * (n - 1) * P
        d[:-2] += kappa * n[2:] * (n[2:] - 1) * P[2:]
        return d
    P = _rk4(_poisson(n0, n[-1]), rhs, dt, steps)
    m_ex = (n * P).sum()
    v_ex = (n * n * P).sum() - m_ex ** 2
    _, bx = _evolve(oracle, c, a, b, dt, steps, True)
    m, v = _moments(np.abs(bx) ** 2 * dV)
    print("two-body: exact <n> %.3f var %.3f; noise <n> %.3f var %.3f" % (m_ex, v_ex, m, v))
    assert abs(m - m_ex) < 0.01 * m_ex          # O(1/n) drift bias: 0.4% here
    assert abs(v - v_ex) < 0.05 * v_ex          # Monte Carlo error ~0.6%
    _, bx0 = _evolve(oracle, c, a, b, dt, steps, False)
    _, v0 = _moments(np.abs(bx0) ** 2 * dV)
    assert v0 < 0.3 * v_ex                      # without the noise: ~0.17 of the exact variance


def test_three_body_loss_one_mode(oracle):
    n0, dt, steps = 100.0, 0.0125, 400
    kappa3 = 1.0 / (2 * 5.0 * n0 ** 2)          # the mean roughly halves over the run
    c = _calc(oracle, l111=1.0)
    dV = _dV(c)
    c.l111 = 2 * kappa3 * dV ** 2
    M = c.cells * c.ne
    rng = np.random.default_rng(2)
    a = _coherent(rng, n0, M) / np.sqrt(dV)
    b = np.zeros(M, complex)
    n = np.arange(int(n0 + 12 * np.sqrt(n0)) + 1)
    r = 2 * kappa3 / 3

    def rhs(P):
        d = -r * n * (n - 1) * (n - 2) * P
        d[:-3] += r * n[3:] * (n[3:] - 1) * (n[3:] - 2) * P[3:]
        return d
    P = _rk4(_poisson(n0, n[-1]), rhs, dt, steps)
    m_ex = (n * P).sum()
    v_ex = (n * n * P).sum() - m_ex ** 2
    ax, _ = _evolve(oracle, c, a, b, dt, steps, True)
    m, v = _moments(np.abs(ax) ** 2 * dV)
    print("three-body: exact <n> %.3f var %.3f; noise <n> %.3f var %.3f" % (m_ex, v_ex, m, v))
    assert abs(m - m_ex) < 0.02 * m_ex
    assert abs(v - v_ex) < 0.06 * v_ex
    ax0, _ = _evolve(oracle, c, a, b, dt, steps, False)
    _, v0 = _moments(np.abs(ax0) ** 2 * dV)
    assert v0 < 0.5 * v_ex


def test_pair_loss_two_modes(oracle):
    n0, kappa, dt, steps = 40.0, 1.0 / (2 * 5.0 * 40.0), 0.0125, 400
    c = _calc(oracle, l12=1.0)
    dV = _dV(c)
    c.l12 = 2 * kappa * dV
    M = c.cells * c.ne
    rng = np.random.default_rng(3)
    a = _coherent(rng, n0, M) / np.sqrt(dV)
    # a relative phase: the exact dynamics does not depend on it, but the cross
    # term D12 = l12 a conj(b) / 2 must carry it (conj(a) b would not)
    b = np.exp(1j) * _coherent(rng, n0, M) / np.sqrt(dV)
    nm = int(n0 + 12 * np.sqrt(n0)) + 1
    na, nb = np.meshgrid(np.arange(nm), np.arange(nm), indexing="ij")
    r = 2 * kappa

    def rhs(P):
        d = -r * na * nb * P
        d[:-1, :-1] += r * na[1:, 1:] * nb[1:, 1:] * P[1:, 1:]
        return d
    p0 = _poisson(n0, nm - 1)
    P = _rk4(np.outer(p0, p0), rhs, dt, steps)
    ma, mb = (na * P).sum(), (nb * P).sum()
    va = (na * na * P).sum() - ma ** 2
    cov = (na * nb * P).sum() - ma * mb
    vd = ((na - nb) ** 2 * P).sum() - (ma - mb) ** 2
    ax, bx = _evolve(oracle, c, a, b, dt, steps, True)
    wa, wb = np.abs(ax) ** 2 * dV, np.abs(bx) ** 2 * dV
    m, v = _moments(wa)
    cw = np.cov(wa, wb)[0, 1]
    vdw = np.var(wa - wb) - 0.5
    print("pair: exact <na> %.3f var %.3f cov %.3f var(na-nb) %.3f; noise %.3f %.3f %.3f %.3f"
          % (ma, va, cov, vd, m, v, cw, vdw))
    assert abs(vd - 2 * n0) < 1e-6 * n0          # exactly conserved by the master equation
    assert abs(m - ma) < 0.02 * ma
    assert abs(v - va) < 0.08 * va
    assert abs(cw - cov) < 0.08 * abs(va)
    assert abs(vdw - vd) < 0.05 * vd              # the cross term keeps the pairs correlated


def test_noise_streams_are_independent_per_step_and_group(oracle):
    z0 = np.array([oracle.loss_noise_cell(5, 0, i, 0) for i in range(4000)])
    z1 = np.array([oracle.loss_noise_cell(5, 0, i, 1) for i in range(4000)])
    z2 = np.array([oracle.loss_noise_cell(5, 0, i, 0, group=1) for i in range(4000)])
    for z in (z0, z1, z2):
        assert abs(z.mean()) < 0.05 and abs(z.var() - 1) < 0.05
    assert abs(np.corrcoef(z0.ravel(), z1.ravel())[0, 1]) < 0.03
    assert abs(np.corrcoef(z0.ravel(), z2.ravel())[0, 1]) < 0.03
    # the initial-noise stream (word 3 = 0) is a different stream
    zi = np.array([oracle.noise_cell(5, 0, i, 0) for i in range(4000)])
    assert abs(np.corrcoef(z0.ravel(), zi.ravel())[0, 1]) < 0.03
