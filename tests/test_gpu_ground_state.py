"""GPU: calculateSteadyState (wigner.cu:84-175) and the residual-targeted solver."""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc, oracle_calc, rel_l2

pytestmark = pytest.mark.gpu


def test_split_step_reference_rule_matches_oracle(bec, oracle):
    m = config1_model(oracle, ne=1)
    c, oc = gpu_calc(bec, **m), oracle_calc(oracle, **m)
    ref, orep = oracle.ground_state(oc, rule=oracle.RULE_REFERENCE, tol=1e-6)
    with bec.Solver(c, precision=0) as s:
        psi, rep, st = s.calculate_steady_state(method=bec.GS_SPLIT_STEP, rule=bec.GS_RULE_REFERENCE,
                                                tol=1e-6, precision=bec.FP64)
    assert st == 0 and rep["converged"]
    assert rep["iterations"] == orep.iterations
    assert rel_l2(psi, ref) < 1e-10
    assert abs(rep["E_ref"] - orep.E_ref) < 1e-9 * abs(orep.E_ref)
    assert abs(rep["mu_ref"] - orep.mu_ref) < 1e-9 * abs(orep.mu_ref)
    assert abs(rep["E"] - orep.E) < 1e-9 * abs(orep.E)
    assert abs(rep["mu"] - orep.mu) < 1e-9 * abs(orep.mu)
    assert abs(rep["residual"] - orep.residual) < 1e-6 * orep.residual


def test_gradient_flow_reaches_residual(bec, oracle):
    m = config1_model(oracle, ne=1)
    c, oc = gpu_calc(bec, **m), oracle_calc(oracle, **m)
    with bec.Solver(c, precision=0) as s:
        psi, rep, st = s.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                                precision=bec.FP64, max_iter=2000)
    assert st == 0 and rep["converged"], rep
    assert rep["residual"] < 1e-6
    r, mu = oracle.residual(oc, psi)     # independent FP64 check of the returned state
    assert r < 1.05e-6
    assert abs(mu - rep["mu"]) < 1e-9 * abs(mu)
    E, mu2, N = oracle.energy_mu(oc, psi)
    assert abs(N - 30000) < 1e-6 * 30000
    assert abs(E - rep["E"]) < 1e-9 * abs(E)


def test_batched_ground_states_match_single(bec, oracle):
    base = config1_model(oracle, ne=1)
    params = [gpu_calc(bec, **dict(base, N=n, fz=fz)) for n, fz in [(20000, 11.96), (45000, 9.0), (8000, 15.0)]]
    out, reps, st = bec.ground_states_batched(params, method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL,
                                              tol=1e-6, precision=bec.FP64, max_iter=3000)
    assert st == 0
    for i, p in enumerate(params):
        oc = oracle.derive(oracle.model(**dict(base, N=[20000, 45000, 8000][i], fz=[11.96, 9.0, 15.0][i])))
        r, _ = oracle.residual(oc, out[i])
        assert r < 1.05e-6, (i, r)
        assert reps[i]["converged"]
