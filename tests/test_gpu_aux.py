"""GPU: calculateEvolution on several groups, failure reporting and the bit-exact
checkpoint.

* bec_calculate_evolution (wigner.cu:340-411) is one propagate step plus the
  observables of every group: checked against propagate(1) + observables on an
  identical context, and against the FP64 oracle's step and observables.
* Failure reporting forced from the host: a NaN written into the state gives
  BEC_ERR_NONFINITE from the observables; a NaN guess gives it from the ground
  state; an iteration cap below what the tolerance needs gives
  BEC_ERR_NOT_CONVERGED with the reached residual reported (the reference's
  loop, wigner.cu:114,155, has no cap and spins).
* The hybrid-space checkpoint restores a run bit for bit, loss noise included
  (the host copies of wigner.cu:166,207 are the reference's seam)."""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc, oracle_calc, rel_l2

pytestmark = pytest.mark.gpu

SEED = 0x5EED


@pytest.fixture(scope="module")
def gs(oracle):
    c = oracle_calc(oracle, **config1_model(oracle, ne=1))
    psi, rep = oracle.ground_state(c, rule=oracle.RULE_REFERENCE, tol=1e-6)
    assert rep.converged
    return psi


@pytest.fixture(scope="module")
def gs64(bec, oracle):
    """Ground state of the 64x32x32 lattice (gradient flow on the GPU, FP64), checked by the oracle."""
    m = config1_model(oracle, nvx=32, nvy=32, nvz=64, ne=1)
    with bec.Solver(gpu_calc(bec, **m), precision=1) as g:
        psi, rep, st = g.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                                precision=1, max_iter=20000)
    assert st == 0
    assert oracle.residual(oracle_calc(oracle, **m), psi)[0] < 1.05e-6
    return psi


def _groups(bec, oracle, ne1=3, ne2=2):
    m1 = config1_model(oracle, ne=ne1, detuning=-41.0)
    m2 = config1_model(oracle, nvx=32, nvy=32, nvz=64, ne=ne2, detuning=-41.0)
    return (gpu_calc(bec, **m1), gpu_calc(bec, **m2)), (oracle_calc(oracle, **m1), oracle_calc(oracle, **m2))


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 3e-5)])
def test_calculate_evolution_two_groups(bec, oracle, gs, gs64, prec, tol):
    (c1, c2), (o1, o2) = _groups(bec, oracle)
    with bec.Solver(None, precision=prec, groups=[c1, c2]) as s, \
            bec.Solver(None, precision=prec, groups=[c1, c2]) as r:
        for ctx in (s, r):
            ctx.init_evolution(gs, SEED, group=0, equilibrate=True)
            ctx.init_evolution(gs64, SEED + 7, group=1, equilibrate=True)
        got = [s.calculate_evolution() for _ in range(3)][-1]
        r.propagate(1)
        r.propagate(1)
        r.propagate(1)
        ref = [r.observables(g) for g in (0, 1)]
        states = [s.get_state(bec.SPACE_X, group=g) for g in (0, 1)]
        t = s.t
    assert len(got) == 2
    for g in (0, 1):
        for k in ("Na", "Nb", "N_total", "visibility", "visibility_grid", "phase", "N1", "N2"):
            assert got[g][k] == pytest.approx(ref[g][k], rel=1e-12, abs=1e-12), (g, k)
    assert t == pytest.approx(3 * c1.dt_evo, rel=1e-14)
    # against the oracle: three reference steps, then the observables of the x-space state
    for g, (oc, psi, seed) in enumerate(((o1, gs, SEED), (o2, gs64, SEED + 7))):
        oa, ob = oracle.init_evolution(oc, psi, seed, equilibrate=True)
        for _ in range(3):
            oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, 1)
        ax, bx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
        assert max(rel_l2(states[g][0], ax), rel_l2(states[g][1], bx)) < tol
        n1, n2 = oracle.rotate_sums(oc, ax, bx, 0.0)
        dV = oc.dx * oc.dy * oc.dz
        assert got[g]["Na"] == pytest.approx(n1 * dV, rel=max(tol, 1e-10) * 10)
        assert got[g]["Nb"] == pytest.approx(n2 * dV, rel=max(tol, 1e-10) * 10)
        assert abs(got[g]["visibility_grid"] - oracle.visibility_grid(oc, ax, bx)) < max(tol, 1e-10) * 10


@pytest.mark.parametrize("prec", [0, 1])
def test_nan_in_state_reports_nonfinite(bec, oracle, gs, prec):
    c = gpu_calc(bec, **config1_model(oracle, ne=2))
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        a, b = s.get_state(bec.SPACE_X)
        s.observables()   # finite state: fine
        a[1, 64, 8, 8] = np.nan
        s.set_state(a, b, bec.SPACE_X)
        s.propagate(2)
        with pytest.raises(bec.BecError) as e:
            s.observables()
        assert e.value.status == bec.api.ERR_NONFINITE
        a[1, 64, 8, 8] = 0.0
        b[0, 3, 2, 1] = np.inf
        s.set_state(a, b, bec.SPACE_X)
        with pytest.raises(bec.BecError) as e:
            s.observables(group=0)
        assert e.value.status == bec.api.ERR_NONFINITE


@pytest.mark.parametrize("method", ["gradient", "split_step"])
def test_ground_state_nan_guess_reports_nonfinite(bec, oracle, gs, method):
    c = gpu_calc(bec, **config1_model(oracle, ne=1))
    guess = gs.copy()
    guess[64, 8, 8] = np.nan
    opt = dict(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL) if method == "gradient" else \
        dict(method=bec.GS_SPLIT_STEP, rule=bec.GS_RULE_REFERENCE)
    with bec.Solver(c, precision=1) as s:
        with pytest.raises(bec.BecError) as e:
            s.calculate_steady_state(guess=guess, tol=1e-6, max_iter=50, precision=1, **opt)
    assert e.value.status == bec.api.ERR_NONFINITE


@pytest.mark.parametrize("method", ["gradient", "split_step"])
def test_ground_state_iteration_cap_reports_not_converged(bec, oracle, method):
    m = config1_model(oracle, ne=1)
    c, oc = gpu_calc(bec, **m), oracle_calc(oracle, **m)
    opt = dict(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL) if method == "gradient" else \
        dict(method=bec.GS_SPLIT_STEP, rule=bec.GS_RULE_RESIDUAL)
    with bec.Solver(c, precision=1) as s:
        psi, rep, st = s.calculate_steady_state(tol=1e-12, max_iter=7, precision=1, **opt)
    assert st == bec.api.ERR_NOT_CONVERGED, (st, rep)
    assert rep["converged"] == 0 and rep["iterations"] == 7
    # the reported residual is the returned state's own, and above the tolerance
    assert np.isfinite(rep["residual"]) and rep["residual"] > 1e-12
    assert rep["residual"] == pytest.approx(oracle.residual(oc, psi)[0], rel=1e-6)
    assert np.all(np.isfinite(psi))


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("noise", [False, True])
def test_hybrid_checkpoint_resumes_bit_for_bit(bec, oracle, gs, prec, noise):
    c = gpu_calc(bec, **config1_model(oracle, ne=8))
    with bec.Solver(c, precision=prec) as ref:
        if noise:
            ref.set_loss_noise(True, seed=0x1055)
        ref.init_evolution(gs, SEED, equilibrate=True)
        ref.propagate(40)
        ha, hb = ref.get_state(bec.SPACE_HYBRID)
        t_mid = ref.t
        step_mid = ref.loss_noise_step() if noise else 0
        ref.propagate(40)
        want = ref.get_state(bec.SPACE_HYBRID)
        want_obs = ref.observables()
    with bec.Solver(c, precision=prec) as res:
        if noise:
            res.set_loss_noise(True, seed=0x1055, first_step=step_mid)
        res.set_state(ha, hb, bec.SPACE_HYBRID)
        res.t = t_mid
        res.propagate(40)
        got = res.get_state(bec.SPACE_HYBRID)
        got_obs = res.observables()
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert got_obs == want_obs
