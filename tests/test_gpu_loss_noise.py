"""GPU: the loss-induced Wigner noise (bec_set_loss_noise) against the oracle's
restatement (orc_propagate_noise) on the same Philox streams, and the
properties of its step counter: replayed step graphs, a resumed run and
switching it off give the same bits as the plain launches."""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc, oracle_calc, rel_l2  # noqa: F401

pytestmark = pytest.mark.gpu

SEED, NSEED = 0x5EED, 0x1055
# the 128x16x16 configuration with its loss rates raised 100-fold, so that the
# noise moves a cell by a few per cent per step
LOSS = dict(gamma111=5.4e-28, gamma12=0.78e-11, gamma22=1.5e-12)


@pytest.fixture(scope="module")
def gs(oracle):
    c = oracle_calc(oracle, **config1_model(oracle, ne=3))
    psi, rep = oracle.ground_state(c, rule=oracle.RULE_REFERENCE, tol=1e-6)
    assert rep.converged
    return psi


def _pair(bec, oracle, **kw):
    m = config1_model(oracle, **dict(LOSS, **kw))
    return gpu_calc(bec, **m), oracle_calc(oracle, **m)


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 3e-5)])
def test_loss_noise_matches_oracle(bec, oracle, gs, prec, tol):
    c, oc = _pair(bec, oracle, ne=3)
    nsteps, first = 25, 7
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.set_loss_noise(True, seed=NSEED, first_step=first)
        s.propagate(nsteps)
        assert s.loss_noise_step() == first + nsteps
        ax, bx = s.get_state(bec.SPACE_X)
    oa0, ob0 = oracle.init_evolution(oc, gs, SEED, equilibrate=False)
    oa, ob = oracle.propagate_noise(oc, oa0, ob0, oc.dt_evo, nsteps, NSEED, first)
    oax, obx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
    ea = rel_l2(ax, oax, axis=(1, 2, 3))
    eb = rel_l2(bx, obx, axis=(1, 2, 3))
    print("prec", prec, "per-trajectory error a", ea, "b", eb, flush=True)
    assert ea.max() < tol, ea
    assert eb.max() < tol, eb
    # the noise is far above the tolerance: the noiseless step lands elsewhere
    pa, pb = oracle.propagate(oc, oa0, ob0, oc.dt_evo, nsteps)
    assert rel_l2(bx, oracle.to_xspace(oc, pb)) > 1000 * tol


def _run(bec, c, gs, nsteps, graphs=True, noise=True, first=0):
    with bec.Solver(c, precision=bec.FP32) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.set_graphs(graphs)
        if noise:
            s.set_loss_noise(True, seed=NSEED, first_step=first)
        s.propagate(nsteps)
        return s.get_state(bec.SPACE_K)


def test_step_graphs_replay_the_noise_stream(bec, oracle, gs):
    # 96 steps: three replays of the captured 32-step block carry the counter on
    c, _ = _pair(bec, oracle, ne=8)
    a1, b1 = _run(bec, c, gs, 96, graphs=True)
    a0, b0 = _run(bec, c, gs, 96, graphs=False)
    assert np.array_equal(a1, a0) and np.array_equal(b1, b0)


def test_resumed_run_continues_the_stream(bec, oracle, gs):
    # FP64: a checkpoint round trip (get_state / set_state transform the stored
    # representation) is exact to rounding, while resuming at a wrong step of
    # the stream would change every cell by the noise
    c, _ = _pair(bec, oracle, ne=8)
    with bec.Solver(c, precision=bec.FP64) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.set_loss_noise(True, seed=NSEED)
        s.propagate(15)
        ak, bk = s.get_state(bec.SPACE_K)
        step = s.loss_noise_step()
        s.propagate(25)
        a_full, b_full = s.get_state(bec.SPACE_K)
    assert step == 15

    def resume(first):
        with bec.Solver(c, precision=bec.FP64) as s:
            s.init_evolution(gs, SEED, equilibrate=False)
            s.set_state(ak, bk, bec.SPACE_K)
            s.set_loss_noise(True, seed=NSEED, first_step=first)
            s.propagate(25)
            return s.get_state(bec.SPACE_K)
    a2, b2 = resume(step)
    assert rel_l2(a2, a_full) < 1e-12 and rel_l2(b2, b_full) < 1e-12
    a3, b3 = resume(0)
    assert rel_l2(b3, b_full) > 1e-4


def test_switching_the_noise_off_restores_the_plain_step(bec, oracle, gs):
    c, _ = _pair(bec, oracle, ne=8)
    a_plain, b_plain = _run(bec, c, gs, 40, noise=False)
    with bec.Solver(c, precision=bec.FP32) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.set_loss_noise(True, seed=NSEED)
        s.set_loss_noise(False)
        s.propagate(40)
        a, b = s.get_state(bec.SPACE_K)
        with pytest.raises(bec.BecError):
            s.loss_noise_step()
    assert np.array_equal(a, a_plain) and np.array_equal(b, b_plain)
