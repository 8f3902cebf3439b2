"""GPU: the 128x16x16 Wigner ensemble at its full size (64 trajectories, 20000
steps of 40 us, losses on, Philox vacuum noise, a sample every 250 steps),
where the oracle would take hours, through properties that hold at any size:

- losses only remove atoms: the ensemble atom number falls at every sample,
  by far more than the FP32 drift of the lossless step;
- without losses the norm is conserved up to the FP32 drift: FP32 unit-modulus
  factors are not exactly of modulus one.  The in-register twiddles are chosen
  so that their modulus errors cancel over each transform (fft_reg.hpp), and the
  table entries are the float pairs nearest modulus one; the drift left is about
  2.5e-9 per step (5.1e-5 over the run; 1.6e-7 per step with nearest-rounded
  twiddles); the FP64 path conserves the norm to 1e-16 per step;
- the FP32 run (FP64 sums) and an all-FP64 GPU run give the same ensemble
  populations along the whole run and the same visibility (to 1e-4) while the
  trajectories stay close; later the noisy trajectories separate, and the
  ensemble coherence agrees within its statistical error.
"""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc

pytestmark = pytest.mark.gpu

SEED = 0x5EED
STEPS, EVERY, NE = 20000, 250, 64


@pytest.fixture(scope="module")
def gs(bec):
    c = gpu_calc(bec, **config1_model(bec, ne=1))
    with bec.Solver(c, precision=bec.FP64) as s:
        psi, rep, _ = s.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                               precision=bec.FP64)
    assert rep["residual"] < 1e-6
    return psi


def _run(bec, gs, prec, norm=False, **kw):
    c = gpu_calc(bec, **config1_model(bec, ne=NE, **{k: v for k, v in kw.items() if k != "with_spread"}))
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=True)
        samples = s.propagate(STEPS, sample_every=EVERY)      # [80][64][4]
    dV = c.dx * c.dy * c.dz
    if norm:                                                    # the full norm, vacuum included
        return samples[:, :, 0:2].sum(axis=2).mean(axis=1) * dV
    vac = c.cells * c.V                                         # vacuum atoms of both components
    pops = samples[:, :, 0:2].mean(axis=1) * dV - vac / 2       # per component, minus its vacuum
    cross = np.hypot(samples[:, :, 2].mean(axis=1), samples[:, :, 3].mean(axis=1)) * dV
    if kw.pop("with_spread", False):
        # standard error of the ensemble-mean cross term <a|b> (both quadratures)
        se = np.hypot(samples[:, :, 2].std(axis=1, ddof=1), samples[:, :, 3].std(axis=1, ddof=1)) * dV / np.sqrt(NE)
        return pops, cross, se
    return pops, cross


def test_losses_only_remove_atoms(bec, gs):
    pops, _ = _run(bec, gs, bec.FP32)
    N = pops.sum(axis=1)
    steps = np.diff(N)
    print("N from %.1f to %.1f; smallest loss per interval %.3g" % (N[0], N[-1], -steps.max()), flush=True)
    assert np.all(steps < -1e-4 * N[:-1]), steps.max()


@pytest.mark.parametrize("prec,per_step", [(0, 1e-8), (1, 1e-15)])
def test_lossless_run_conserves_the_norm(bec, gs, prec, per_step):
    T = _run(bec, gs, prec, norm=True, gamma111=0.0, gamma12=0.0, gamma22=0.0)
    drift = np.abs(T - T[0]).max() / T[0]
    print("lossless, precision %d: largest relative change of the norm over %d steps %.3g" % (prec, STEPS, drift),
          flush=True)
    assert drift < STEPS * per_step


def test_fp32_matches_fp64_over_the_full_run(bec, gs):
    p32, c32, se32 = _run(bec, gs, bec.FP32, with_spread=True)
    p64, c64, se64 = _run(bec, gs, bec.FP64, with_spread=True)
    N = p64.sum(axis=1)
    ep = np.abs(p32 - p64).max() / N.max()
    # the visibility 2|<a|b>| / (Na + Nb) per sample
    v32, v64 = 2 * c32 / p32.sum(axis=1), 2 * c64 / p64.sum(axis=1)
    ev = np.abs(v32 - v64)
    z = np.abs(c32 - c64) / np.hypot(se32, se64)
    first = np.argmax(np.arange(len(ev)) * EVERY >= 4000)
    print("FP32 vs FP64 over %d steps: populations %.3g of N; visibility %.3g over the first 4000 steps, %.3g "
          "over the run; coherence difference %.2f combined standard errors at most (mean %.2f)"
          % (STEPS, ep, ev[:first + 1].max(), ev.max(), z.max(), z.mean()), flush=True)
    # populations: the FP32 norm drift (5.1e-5 of the norm over the lossless run) and
    # the separation of the noisy trajectories (measured 2e-5 to 1e-4 of N)
    assert ep < 1e-3
    # visibility: within 1e-4 while the trajectories stay close; over the whole noisy
    # run the FP32 and FP64 trajectories separate, and the ensemble coherence then
    # agrees within its statistical error
    assert ev[:first + 1].max() < 1e-4
    assert z.mean() < 1.0 and z.max() < 4.0
