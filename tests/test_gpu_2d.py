"""GPU: the 2D program (bec_run_2d, one workgroup per trajectory with the state in
LDS) against the FP64 oracle's restatement of programs/bec_wigner_2d.cpp on the
same parameters and Philox seeds: ground state, per-sample ensemble means and
errors, visibility and mean density."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {
    # the program's constants on 64 x 64 (config 0's lattice), a short ensemble
    "program_64": dict(nvx=64, nvy=64, ne=8, ntW=40, n_meas=3),
    # the residual-stopped ground state, a 32 x 32 lattice, more samples
    "residual_32": dict(nvx=32, nvy=32, xmax=12.0, ymax=12.0, mu=10.0, gs_rule=1, gs_tol=1e-3, gs_max_iter=3000,
                        dtG=0.01, ne=4, ntW=10, n_meas=5),
    # non-square lattice, no noise, detuned
    "rect_16x32": dict(nvx=16, nvy=32, xmax=10.0, ymax=14.0, mu=8.0, E=1.3, Va=0.0, Vb=0.0, ne=2, ntW=5,
                       n_meas=4),
}


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.mark.parametrize("case", list(CASES))
def test_2d_program_matches_oracle(bec, oracle, case):
    kw = CASES[case]
    got = bec.run_2d(bec.params2d(**kw))
    # the same constants on both sides; the oracle's default lattice is the program's 128 x 128
    op = oracle.params2d(**dict(dict(nvx=64, nvy=64), **kw))
    ref = oracle.run2d(op)
    rr = ref["report"]
    rep = got["report"]
    print(case, "gpu", rep, "oracle iterations", rr.gs_iterations, "residual", rr.gs_residual)
    assert rep["gs_iterations"] == rr.gs_iterations
    assert rep["N0"] == pytest.approx(rr.N0, rel=1e-11)
    assert _rel(got["gs"], ref["gs"]) < 1e-11
    assert rep["gs_residual"] == pytest.approx(rr.gs_residual, rel=1e-7, abs=1e-13)
    assert _rel(got["m"], ref["m"]) < 1e-10
    assert np.allclose(got["var"], ref["var"], rtol=1e-6, atol=1e-9 * np.abs(ref["m"]).max())
    assert _rel(got["vis"], ref["vis"]) < 1e-10
    assert _rel(got["avn"], ref["avn"]) < 1e-10


def test_2d_rejects_lattices_beyond_one_workgroup(bec):
    with pytest.raises(bec.BecError):
        bec.run_2d(bec.params2d(nvx=128, nvy=128, ne=1))
