"""GPU: the FP32/FP64 split (FP32 state, FFT and nonlinear step; FP64 sums) against an
all-FP64 GPU run and the FP64 CPU oracle, per observable, at 128x32x32."""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

SEED = 0x5EED


def _model(ne, nxy=32, nz=128, dt=4e-5):
    return dict(N=30000, nvx=nxy, nvy=nxy, nvz=nz, ne=ne, gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14,
                detuning=0.0, dt_evo=dt, dt_steady=1e-5, t_equilib=0.0, Va=1.0, Vb=1.0, itmax=1)


# (lattice, step, FP32 state bar, FP32 observable bar) after 200 steps.  The
# 128x32x32 lattice runs at 20 us: at 40 us its largest kinetic phase per step
# (5.2 rad) lies in the split step's unstable band and any perturbation, FP64
# rounding included, grows (tests/test_oracle_stability.py).
CASES = [(16, 4e-5, 5e-4, 1e-4), (32, 2e-5, 5e-4, 1e-4)]


@pytest.mark.parametrize("nxy,dt,tol_state32,tol_obs32", CASES)
def test_precision_split_vs_fp64_oracle(bec, oracle, nxy, dt, tol_state32, tol_obs32):
    c1 = bec.derive(bec.default_model(**_model(1, nxy=nxy, dt=dt)))
    with bec.Solver(c1, precision=bec.FP64) as g:
        gs, rep, st = g.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-8,
                                               precision=bec.FP64, max_iter=20000)
    assert st == 0
    ne, steps = 2, 200
    oc = oracle.derive(oracle.model(**_model(ne, nxy=nxy, dt=dt)))
    c = bec.derive(bec.default_model(**_model(ne, nxy=nxy, dt=dt)))
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=True)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, steps)
    ox, oy = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
    dV = oc.dx * oc.dy * oc.dz
    n1, n2 = oracle.rotate_sums(oc, ox, oy, 0.0)
    vg = oracle.visibility_grid(oc, ox, oy)
    S = oracle.traj_sums(oc, ox, oy).sum(axis=0)
    phase = np.arctan2(S[3], S[2])
    report = {}
    for prec, tol_state, tol_obs in ((bec.FP64, 1e-10, 1e-10), (bec.FP32, tol_state32, tol_obs32)):
        with bec.Solver(c, precision=prec) as s:
            s.init_evolution(gs, SEED, equilibrate=True)
            s.propagate(steps)
            ga, gb = s.get_state(bec.SPACE_X)
            o = s.observables()
        e = {"state": float(max(rel_l2(ga[i], ox[i]) for i in range(ne))),
             # atom-number errors as fractions of the total: after the pulse one
             # port can be nearly empty, where a relative error says nothing
             "Na": abs(o["Na"] - n1 * dV) / abs((n1 + n2) * dV), "Nb": abs(o["Nb"] - n2 * dV) / abs((n1 + n2) * dV),
             "visibility_grid": abs(o["visibility_grid"] - vg),
             "phase": abs(float(np.angle(np.exp(1j * (o["phase"] - phase)))))}
        report[prec] = e
        assert e["state"] < tol_state, e
        for k in ("Na", "Nb", "visibility_grid", "phase"):
            assert e[k] < tol_obs, (prec, k, e)
    print(f"precision split vs FP64 oracle (128x{nxy}x{nxy}, 2 traj, 200 steps):", report)


def test_fp32_visibility_error_over_long_run(bec):
    """Mean-field run (no noise) at 128x16x16: the FP32 split stays within 1e-4 of the
    all-FP64 GPU run in visibility over 4000 steps (measured 3.5e-5).  The total atom
    number drifts in FP32 because unit-modulus factors are not exactly of modulus
    one, although the twiddles and tables are chosen so that their errors cancel
    (measured 5.7e-5 after 4000 steps; 1.1e-4 with the twiddles balanced only in
    pairs and nearest-rounded tables); its bar is 2e-4."""
    from bec_amd import precision_study as P
    r = P.main(["--traj", "8", "--steps", "4000", "--sample-every", "500", "--nxy", "16", "--va", "0", "--vb", "0"])
    e = r["fp32_vs_fp64_gpu"]
    print("FP32 vs FP64 over 4000 steps: visibility %.3g (grid %.3g), N_total %.3g, phase %.3g"
          % (e["visibility"]["max_abs"], e["visibility_grid"]["max_abs"], e["N_total"]["max_rel"],
             e["phase"]["max_abs"]), flush=True)
    assert e["visibility_grid"]["max_abs"] <= 1e-4, e
    assert e["visibility"]["max_abs"] <= 1e-4, e
    assert e["N_total"]["max_rel"] <= 2e-4, e


def _config1_meanfield_20000(bec):
    from bec_amd import precision_study as P
    return P.main(["--traj", "8", "--steps", "20000", "--sample-every", "250", "--nxy", "16", "--va", "0",
                   "--vb", "0"])


@pytest.fixture(scope="module")
def long_run(bec):
    return _config1_meanfield_20000(bec)


@pytest.mark.xfail(strict=True, reason="FP32 misses the 1e-4 visibility target over 20000 steps "
                   "(measured 1.6e-4; the norm drift of the transforms is down to 2.5e-9 per step, and "
                   "what is left is the rounding of the whole step); the all-FP64 context meets it")
def test_fp32_visibility_within_1e4_over_20000_steps(long_run):
    """Config-1 physics without noise, 20000 steps (0.8 s): the 1e-4 visibility
    target against the all-FP64 GPU run (which equals the FP64 oracle to 1e-11
    step by step)."""
    e = long_run["fp32_vs_fp64_gpu"]
    assert e["visibility"]["max_abs"] <= 1e-4, e


def test_fp32_error_over_20000_steps_stays_at_its_measured_level(long_run):
    """The same run held to what FP32 reaches, so that a regression shows: the
    visibility within 2.5e-4 (measured 1.6e-4; on the angle grid 6e-4, measured
    4.0e-4: a drifted phase moves the grid's maximum to the next angle), the
    atom number within 3e-4 of N (measured 1.6e-4), the relative phase within
    4e-3 rad (measured 2.5e-3)."""
    e = long_run["fp32_vs_fp64_gpu"]
    print("FP32 vs FP64 over 20000 steps:", e, flush=True)
    assert e["visibility"]["max_abs"] <= 2.5e-4, e
    assert e["visibility_grid"]["max_abs"] <= 6e-4, e
    assert e["N_total"]["max_rel"] <= 3e-4, e
    assert e["phase"]["max_abs"] <= 4e-3, e
