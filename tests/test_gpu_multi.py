"""GPU: the uneven grids of one context stepped by one launch per pass
(bec_set_multi_launch; the table of tile ranges in multi.hip) against the FP64
oracle, and against the per-group launches on the same seeds.

The multi-group launch runs the per-group kernels' tile bodies with the
default radix split (z_r1) of each lattice, as the per-group default schedule
does, and the same xy tile; only the workgroup shape and the order in which
tiles are visited differ, and neither changes any value: the two launch paths
agree bit for bit.  Both groups step at 20 us (the 256x32x32 lattice is
unstable at 40 us; see tests/test_oracle_stability.py)."""
import numpy as np
import pytest

from tests.helpers import gpu_calc, oracle_calc, rel_l2
from tests.test_gpu_groups import DT, _ground_state, _models

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def states(bec, oracle):
    m1, m2 = _models(oracle, 1, 1)
    return _ground_state(bec, oracle, m1), _ground_state(bec, oracle, m2)


def _run(bec, oracle, states, ne1, ne2, steps, every, multi, graphs=True, losses=True, noise=False):
    g1, g2 = states
    kw = {} if losses else dict(gamma111=0.0, gamma12=0.0, gamma22=0.0)
    m1, m2 = _models(oracle, ne1, ne2, **kw)
    with bec.Solver(None, precision=0, groups=[gpu_calc(bec, **m1), gpu_calc(bec, **m2)]) as s:
        s.init_evolution(g1, 0x5EED, group=0, equilibrate=True)
        s.init_evolution(g2, 0xBEEF, group=1, equilibrate=True)
        s.set_graphs(graphs)
        if noise:
            s.set_loss_noise(True, seed=0x77)
        if multi:
            s.set_multi_launch(True)
        samples = s.propagate(steps, sample_every=every)
        hy = [s.get_state(bec.SPACE_HYBRID, group=g) for g in (0, 1)]
        xs = [s.get_state(bec.SPACE_X, group=g) for g in (0, 1)]
    return samples, hy, xs


def test_multi_launch_matches_oracle_per_trajectory(bec, oracle, states):
    """3 trajectories at 128x16x16 and 2 at 256x32x32, 12 steps sampled every
    4, one launch per pass: every trajectory against the FP64 oracle (the FP32
    bar of the per-group launches)."""
    m1, m2 = _models(oracle, 3, 2)
    o = [oracle_calc(oracle, **m1), oracle_calc(oracle, **m2)]
    samples, _, xs = _run(bec, oracle, states, 3, 2, 12, 4, multi=True)
    assert samples.shape == (3, 5, 4)
    for g, (oc, psi, seed) in enumerate(zip(o, states, (0x5EED, 0xBEEF))):
        oa, ob = oracle.init_evolution(oc, psi, seed, equilibrate=True)
        oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, 12)
        ax, bx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
        ea = rel_l2(xs[g][0], ax, axis=(1, 2, 3))
        eb = rel_l2(xs[g][1], bx, axis=(1, 2, 3))
        assert max(ea.max(), eb.max()) < 5e-5, (g, ea, eb)


@pytest.mark.parametrize("graphs", [True, False])
def test_multi_launch_bit_exact_against_per_group(bec, oracle, states, graphs):
    """The same 5 trajectories through both launch paths (graph-captured
    segments, and one launch per pass without graphs): identical hybrid
    states and samples."""
    s0, h0, _ = _run(bec, oracle, states, 3, 2, 20, 10, multi=False, graphs=graphs)
    s1, h1, _ = _run(bec, oracle, states, 3, 2, 20, 10, multi=True, graphs=graphs)
    for g in (0, 1):
        for c in (0, 1):
            assert np.array_equal(h0[g][c], h1[g][c]), (g, c, np.abs(h0[g][c] - h1[g][c]).max())
    assert np.array_equal(s0, s1)


def test_multi_launch_full_size_bit_exact(bec, oracle, states):
    """The full configuration (256 trajectories at 128x16x16 and 128 at
    256x32x32, whose tiles outnumber the resident workgroups many times over,
    so every workgroup walks tiles of both lattices), 8 steps with losses:
    identical to the per-group launches."""
    s0, h0, _ = _run(bec, oracle, states, 256, 128, 8, 4, multi=False)
    s1, h1, _ = _run(bec, oracle, states, 256, 128, 8, 4, multi=True)
    for g in (0, 1):
        for c in (0, 1):
            assert np.array_equal(h0[g][c], h1[g][c]), (g, c)
    assert np.array_equal(s0, s1)


def test_multi_launch_loss_noise_bit_exact(bec, oracle, states):
    """With the loss-induced noise on, each group keeps its own Philox stream
    and step counter in the shared launch: identical to the per-group
    launches."""
    s0, h0, _ = _run(bec, oracle, states, 3, 2, 10, 5, multi=False, noise=True)
    s1, h1, _ = _run(bec, oracle, states, 3, 2, 10, 5, multi=True, noise=True)
    for g in (0, 1):
        for c in (0, 1):
            assert np.array_equal(h0[g][c], h1[g][c]), (g, c)
    assert np.array_equal(s0, s1)


def test_multi_launch_rejects_fp64(bec, oracle):
    m1, m2 = _models(oracle, 1, 1)
    with bec.Solver(None, precision=1, groups=[gpu_calc(bec, **m1), gpu_calc(bec, **m2)]) as s:
        with pytest.raises(bec.BecError) as e:
            s.set_multi_launch(True)
        assert e.value.status == bec.api.ERR_UNSUPPORTED
