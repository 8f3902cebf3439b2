"""GPU: the reference's one stored result curve, data/gpu_sp_30k.yaml (copied to
tests/golden/gpu_sp_30k.yaml): visibility against time of the 30k-atom Ramsey
run, "GPU, single precision, 16x16x128, 40mus timestep", "no noise", 61 samples
over 0-0.6 s.

The file does not record the loss rates.  It is reproduced with the rates that
are commented out in the reference's parameter block
(programs/bec_wigner_3d.cuda/wigner.cpp:74-76: gamma111 = 5.4e-30 cm^6/s,
gamma12 = 0.78e-13 cm^3/s, gamma22 = 1.194e-13 cm^3/s) and otherwise the
block's defaults (wigner.cpp:43-98) with N = 30000 and no vacuum noise:
RMS 0.008 over the 61 samples, the minimum 0.629 at 0.18 s (stored: 0.6355 at
0.18 s) and the revival 0.907 at 0.32 s (stored: 0.9087 at 0.32 s).  With the
block's active lossless setting the curve runs about 10% slower (RMS 0.10); the
detuning (0 or -41 Hz), the equilibration step, FP32 against FP64 and the
ground-state rule each move the RMS by less than 1e-4."""
import os

import numpy as np
import pytest
import yaml

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LOSSES = dict(gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.194e-13)


@pytest.fixture(scope="module")
def stored():
    d = yaml.safe_load(open(os.path.join(HERE, "golden", "gpu_sp_30k.yaml")))
    t = np.array(yaml.safe_load(d["xarray"]))
    v = np.array(yaml.safe_load(d["yarray"]))
    assert len(t) == len(v) == 61
    return t, v


def _curve(bec, **kw):
    m = dict(N=30000, nvx=16, nvy=16, nvz=128, ne=1, Va=0.0, Vb=0.0, detuning=0.0, gamma111=0.0, gamma12=0.0,
             gamma22=0.0, dt_evo=4e-5, dt_steady=1e-5, t_equilib=0.0, itmax=1)
    m.update(kw)
    c = bec.derive(bec.default_model(**m))
    with bec.Solver(c, precision=bec.FP64) as g:
        gs, rep, st = g.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-7,
                                               precision=1, max_iter=20000)
    assert st == 0
    with bec.Solver(c, precision=bec.FP32) as s:
        s.init_evolution(gs, 1, equilibrate=True)
        v0 = s.observables()["visibility"]
        sums = s.propagate(15000, sample_every=250)[:, 0, :]
    vis = 2 * np.hypot(sums[:, 2], sums[:, 3]) / (sums[:, 0] + sums[:, 1])
    return np.concatenate([[v0], vis])


def test_stored_visibility_curve_with_the_commented_loss_rates(bec, stored):
    t, vref = stored
    v = _curve(bec, **LOSSES)
    assert v.shape == vref.shape
    d = v - vref
    rms = float(np.sqrt(np.mean(d ** 2)))
    print(f"RMS {rms:.4f}, max |dev| {np.abs(d).max():.4f}")
    assert rms < 0.012
    assert np.abs(d).max() < 0.025
    # the first minimum and the revival: same sample, values close
    i_min, j_min = np.argmin(v[:26]), np.argmin(vref[:26])
    assert i_min == j_min == 18                      # 0.18 s
    assert abs(v[i_min] - vref[j_min]) < 0.01
    i_rev, j_rev = 25 + np.argmax(v[25:41]), 25 + np.argmax(vref[25:41])
    assert i_rev == j_rev == 32                      # 0.32 s
    assert abs(v[i_rev] - vref[j_rev]) < 0.005


def test_stored_curve_is_missed_without_losses(bec, stored):
    """The lossless block setting oscillates slower: the stored curve pins the losses."""
    _, vref = stored
    v = _curve(bec)
    assert float(np.sqrt(np.mean((v - vref) ** 2))) > 0.05
    assert 25 + np.argmax(v[25:41]) != 32
