"""GPU: uneven grids in one context (the batched uneven-ensembles configuration:
trajectories at 128x16x16 with N = 30k beside trajectories at 256x32x32 with
N = 70k, config-1 physics) against the FP64 oracle, and the full-size run held
to properties that do not depend on its size.

Both groups step at 20 us: at 40 us the 256x32x32 lattice's largest kinetic
phase per step is 4.9 rad, inside the split step's first unstable band (see
tests/test_oracle_stability.py); 20 us keeps it at 2.4 rad.  The throughput of
the kernels does not depend on dt."""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc, oracle_calc, rel_l2

pytestmark = pytest.mark.gpu

DT = 2e-5


def _models(oracle, ne1, ne2, **kw):
    m1 = config1_model(oracle, ne=ne1, dt_evo=DT, **kw)
    m2 = config1_model(oracle, N=70000, nvx=32, nvy=32, nvz=256, ne=ne2, dt_evo=DT, **kw)
    return m1, m2


def _ground_state(bec, oracle, m):
    with bec.Solver(gpu_calc(bec, **dict(m, ne=1)), precision=1) as g:
        psi, rep, st = g.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                                precision=1, max_iter=20000)
    assert st == 0, rep
    return psi


@pytest.fixture(scope="module")
def states(bec, oracle):
    m1, m2 = _models(oracle, 1, 1)
    g1, g2 = _ground_state(bec, oracle, m1), _ground_state(bec, oracle, m2)
    assert oracle.residual(oracle_calc(oracle, **m1), g1)[0] < 1.05e-6
    assert oracle.residual(oracle_calc(oracle, **m2), g2)[0] < 1.05e-6
    return g1, g2


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 5e-5)])
def test_uneven_groups_match_oracle_per_trajectory(bec, oracle, states, prec, tol):
    """3 trajectories at 128x16x16 and 2 at 256x32x32 in one context, different
    seeds, 12 steps sampled every 4: every trajectory of each group against the
    oracle, and the per-trajectory sums of each sample."""
    g1, g2 = states
    m1, m2 = _models(oracle, 3, 2)
    c1, c2 = gpu_calc(bec, **m1), gpu_calc(bec, **m2)
    o = [oracle_calc(oracle, **m1), oracle_calc(oracle, **m2)]
    seeds = (0x5EED, 0xBEEF)
    with bec.Solver(None, precision=prec, groups=[c1, c2]) as s:
        s.init_evolution(g1, seeds[0], group=0, equilibrate=True)
        s.init_evolution(g2, seeds[1], group=1, equilibrate=True)
        samples = s.propagate(12, sample_every=4)
        got = [s.get_state(bec.SPACE_X, group=g) for g in (0, 1)]
        obs = [s.observables(g) for g in (0, 1)]
    assert samples.shape == (3, 5, 4)
    for g, (oc, psi) in enumerate(zip(o, (g1, g2))):
        oa, ob = oracle.init_evolution(oc, psi, seeds[g], equilibrate=True)
        sums = []
        for _ in range(3):
            oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, 4)
            sums.append(oracle.traj_sums(oc, oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)))
        ax, bx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
        ea = rel_l2(got[g][0], ax, axis=(1, 2, 3))
        eb = rel_l2(got[g][1], bx, axis=(1, 2, 3))
        assert max(ea.max(), eb.max()) < tol, (g, ea, eb)
        lo = 0 if g == 0 else 3
        for k in range(3):
            assert rel_l2(samples[k, lo:lo + oc.ne], sums[k]) < max(tol, 1e-12) * 10, (g, k)
        assert abs(obs[g]["visibility_grid"] - oracle.visibility_grid(oc, ax, bx)) < max(tol, 1e-10) * 10


def _run_full(bec, oracle, states, prec, losses, steps=400, every=50):
    g1, g2 = states
    kw = {} if losses else dict(gamma111=0.0, gamma12=0.0, gamma22=0.0)
    m1, m2 = _models(oracle, 256, 128, **kw)
    with bec.Solver(None, precision=prec, groups=[gpu_calc(bec, **m1), gpu_calc(bec, **m2)]) as s:
        s.init_evolution(g1, 0x5EED, group=0, equilibrate=True)
        s.init_evolution(g2, 0xBEEF, group=1, equilibrate=True)
        first = np.concatenate([s.trajectory_sums(0), s.trajectory_sums(1)])
        samples = s.propagate(steps, sample_every=every)
        obs = [s.observables(g) for g in (0, 1)]
    return first, samples, obs


@pytest.fixture(scope="module")
def full_runs(bec, oracle, states):
    return {(p, l): _run_full(bec, oracle, states, p, l) for p in (0, 1) for l in (False, True)}


@pytest.mark.parametrize("prec,bar", [(1, 1e-12), (0, 5e-5)])
def test_full_size_lossless_norm(full_runs, prec, bar):
    """256 + 128 trajectories, no losses: every step is unitary, so each
    trajectory's sum |a|^2 + |b|^2 keeps its initial value."""
    first, samples, _ = full_runs[(prec, False)]
    n0 = first[:, 0] + first[:, 1]
    for k in range(samples.shape[0]):
        n = samples[k, :, 0] + samples[k, :, 1]
        assert np.max(np.abs(n / n0 - 1.0)) < bar, k


@pytest.mark.parametrize("prec", [0, 1])
def test_full_size_lossy_atoms_fall(full_runs, prec):
    """With losses every nonlinear step removes atoms at every cell: each
    trajectory's norm falls from sample to sample."""
    first, samples, _ = full_runs[(prec, True)]
    prev = first[:, 0] + first[:, 1]
    for k in range(samples.shape[0]):
        n = samples[k, :, 0] + samples[k, :, 1]
        assert np.all(n < prev), k
        prev = n


def test_full_size_fp32_populations_match_fp64(full_runs):
    """FP32 state with FP64 sums against the all-FP64 run of the same seeds:
    the populations and visibility of each group agree to 1e-4 of N."""
    for losses in (False, True):
        o32 = full_runs[(0, losses)][2]
        o64 = full_runs[(1, losses)][2]
        for g in (0, 1):
            N = o64[g]["N1"] + o64[g]["N2"]
            for k in ("N1", "N2", "Na", "Nb"):
                assert abs(o32[g][k] - o64[g][k]) < 1e-4 * abs(N), (losses, g, k)
            assert abs(o32[g]["visibility"] - o64[g]["visibility"]) < 1e-4, (losses, g)
