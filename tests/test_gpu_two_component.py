"""GPU: two-component ground states (bec_ground_states_two_component) against the
FP64 oracle, batched against single solves, and the decoupled limit against the
one-component GPU solver."""
import numpy as np
import pytest

from tests.helpers import config1_model, rel_l2

pytestmark = pytest.mark.gpu


def _model(**kw):
    d = config1_model(None, ne=1)
    d.update(kw)
    return d


def test_two_component_matches_oracle(bec, oracle):
    m = _model()
    c = bec.derive(bec.default_model(**m))
    oc = oracle.derive(oracle.model(**m))
    oa, ob, orep = oracle.ground_state2(oc, tol=1e-8, max_iter=5000)
    out, reps, st = bec.ground_states_two_component([c], tol=1e-8, max_iter=5000, precision=bec.FP64)
    r = reps[0]
    print("two-component GS: gpu", r, "oracle iterations", orep.iterations, "residual", orep.residual)
    assert st == 0 and r["converged"]
    assert r["residual"] < 1e-8
    # same algorithm in FP64: the iteration counts agree up to rounding
    assert abs(r["iterations"] - orep.iterations) <= 3
    assert rel_l2(out[0, 0], oa) < 1e-6 and rel_l2(out[0, 1], ob) < 1e-6
    assert r["mu_a"] == pytest.approx(orep.mu_a, rel=1e-9)
    assert r["mu_b"] == pytest.approx(orep.mu_b, rel=1e-9)
    assert r["E"] == pytest.approx(orep.E, rel=1e-9)
    assert r["N_a"] == pytest.approx(orep.N_a, rel=1e-10)
    # the oracle's residual of the GPU state
    res, _, _ = oracle.residual2(oc, out[0, 0], out[0, 1])
    assert res < 2e-8


def test_two_component_batched_equals_single(bec):
    sets = [bec.derive(bec.default_model(**_model(a12=a12, N=n))) for a12, n in ((95.0, 20000), (100.0, 50000))]
    out, reps, st = bec.ground_states_two_component(sets, frac_a=[0.5, 0.3], tol=1e-7, max_iter=5000,
                                                    precision=bec.FP64)
    assert st == 0
    for i, (cs, f) in enumerate(zip(sets, (0.5, 0.3))):
        o1, r1, st1 = bec.ground_states_two_component([cs], frac_a=f, tol=1e-7, max_iter=5000, precision=bec.FP64)
        assert st1 == 0
        assert reps[i]["iterations"] == r1[0]["iterations"]
        assert np.abs(out[i] - o1[0]).max() < 1e-12 * np.abs(o1[0]).max()
        assert reps[i]["N_a"] == pytest.approx(f * cs.N, rel=1e-10)


def test_decoupled_limit_equals_one_component_solver(bec):
    c = bec.derive(bec.default_model(**_model()))
    c.g12 = 0.0
    out, reps, st = bec.ground_states_two_component([c], tol=1e-9, max_iter=5000, precision=bec.FP64)
    assert st == 0
    c1 = bec.derive(bec.default_model(**_model()))
    c1.N = 15000            # same lattice, TF shape and preconditioner shift
    ref, rep1, st1 = bec.ground_states_batched([c1], method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-9,
                                               max_iter=5000, precision=bec.FP64)
    assert st1 == 0
    assert rel_l2(out[0, 0], ref[0]) < 1e-7
    assert reps[0]["mu_a"] == pytest.approx(rep1[0]["mu"], rel=1e-10)
