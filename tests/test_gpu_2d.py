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


def test_config0_rb87_matches_oracle(bec, oracle):
    """Config 0: the Rb-87 cloud of 30k atoms (a11 = 100.4, a22 = 95.0, a12 = 97.66 a0)
    in the 2D program's units (bec_fill_2d_params: the radial plane of the
    97.6 / 11.96 Hz trap), 64 x 64, ground state of |1> to residual 1e-8 by the
    renormalised gradient flow, the pi/2 pulse, 2000 steps of 10 us with every
    step sampled, no noise, one trajectory: the GPU against the oracle."""
    model = bec.default_model(N=30000, nvx=64, nvy=64, nvz=2, ne=1, Va=0.0, Vb=0.0, detuning=0.0, dt_evo=1e-5,
                              dt_steady=1e-5)
    p = bec.params2d_from_model(model, n_meas=2000)
    got = bec.run_2d(p)
    op = oracle.Params2D(**{f: getattr(p, f) for f, _ in oracle.Params2D._fields_})
    ref = oracle.run2d(op)
    rr, rep = ref["report"], got["report"]
    print("config 0: gpu", rep, "oracle iterations", rr.gs_iterations, "residual", rr.gs_residual)
    assert rep["gs_residual"] < 1e-8 and rr.gs_residual < 1e-8
    assert abs(rep["gs_iterations"] - rr.gs_iterations) <= 1
    assert rep["N0"] == pytest.approx(30000, rel=1e-10)
    assert _rel(got["gs"], ref["gs"]) < 1e-9
    assert _rel(got["m"], ref["m"]) < 1e-9
    assert _rel(got["vis"], ref["vis"]) < 1e-9
    assert _rel(got["avn"], ref["avn"]) < 1e-9
    # N1 + N2 is conserved by every step and pulse
    assert np.allclose(got["m"].sum(axis=0), 30000, rtol=1e-9)
