"""GPU: the real-time hot path (initEvolution + propagate, wigner.cu:178-227) and the
observables against the FP64 oracle on identical Philox seeds."""
import numpy as np
import pytest

from tests.helpers import config1_model, gpu_calc, oracle_calc, rel_l2  # noqa: F401

pytestmark = pytest.mark.gpu

SEED = 0x5EED


@pytest.fixture(scope="module")
def gs(oracle):
    c = oracle_calc(oracle, **config1_model(oracle, ne=4))
    psi, rep = oracle.ground_state(c, rule=oracle.RULE_REFERENCE, tol=1e-6)
    assert rep.converged
    return psi


def _pair(bec, oracle, prec, **kw):
    m = config1_model(oracle, **kw)
    return gpu_calc(bec, **m), oracle_calc(oracle, **m)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 2e-6)])
def test_init_evolution_matches_oracle(bec, oracle, gs, prec, tol):
    c, oc = _pair(bec, oracle, prec, ne=4)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=True)
        ak, bk = s.get_state(bec.SPACE_K)
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=True)
    assert rel_l2(ak, oa) < tol
    assert rel_l2(bk, ob) < tol


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 2e-5)])
def test_propagate_matches_oracle_per_trajectory(bec, oracle, gs, prec, tol):
    c, oc = _pair(bec, oracle, prec, ne=3, detuning=-41.0)
    nsteps = 25
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.propagate(nsteps)
        ax, bx = s.get_state(bec.SPACE_X)
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=False)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, nsteps)
    oax, obx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
    ea = rel_l2(ax, oax, axis=(1, 2, 3))
    eb = rel_l2(bx, obx, axis=(1, 2, 3))
    assert ea.max() < tol, ea
    assert eb.max() < tol, eb


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 2e-5)])
def test_strong_losses_take_both_loss_factor_forms(bec, oracle, gs, prec, tol):
    """Loss rates 100x config 1's: the per-step loss exponent y = -dt (l111 na^2 + l12 nb)/2 of
    component a then straddles 2^-8, where the FP32 step switches from its cubic expm1 to
    exp(y) - 1, so both forms run in the same waves; each trajectory must still follow the
    FP64 oracle."""
    loss = dict(gamma111=5.4e-28, gamma12=0.78e-11, gamma22=1.5e-12)
    c, oc = _pair(bec, oracle, prec, ne=3, **loss)
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=False)
    na = np.abs(oracle.to_xspace(oc, oa)) ** 2
    nb = np.abs(oracle.to_xspace(oc, ob)) ** 2
    ya = -0.5 * oc.dt_evo * (oc.l111 * na ** 2 + oc.l12 * nb)
    big = np.mean(np.abs(ya) >= 2.0 ** -8)
    assert 0.001 < big < 0.5, big
    nsteps = 25
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.propagate(nsteps)
        ax, bx = s.get_state(bec.SPACE_X)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, nsteps)
    ea = rel_l2(ax, oracle.to_xspace(oc, oa), axis=(1, 2, 3))
    eb = rel_l2(bx, oracle.to_xspace(oc, ob), axis=(1, 2, 3))
    assert ea.max() < tol, ea
    assert eb.max() < tol, eb
    # the losses act: the atoms of component a fall by more than a percent over the run
    n0 = np.sum(na, axis=(1, 2, 3))
    n1 = np.sum(np.abs(ax) ** 2, axis=(1, 2, 3))
    assert np.all(n1 < 0.99 * n0), n1 / n0


@pytest.mark.parametrize("prec", [0, 1])
def test_samples_and_observables(bec, oracle, gs, prec):
    c, oc = _pair(bec, oracle, prec, ne=4)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=True)
        samples = s.propagate(10, sample_every=5)
        ax, bx = s.get_state(bec.SPACE_X)
        obs = s.observables()
        sums = s.trajectory_sums()
        proj = s.projections(alpha=0.0)
    assert samples.shape == (2, 4, 4)
    # the sample pass and get_state run the same in-register transform: sums agree to FP64 rounding
    ref = oracle.traj_sums(oc, ax, bx)
    assert rel_l2(sums, ref) < 1e-12
    assert rel_l2(samples[1], sums) < 1e-12
    # observables from the sums vs the reference's per-cell rotate-and-reduce
    n1, n2 = oracle.rotate_sums(oc, ax, bx, 0.0)
    dV = oc.dx * oc.dy * oc.dz
    assert abs(obs["Na"] - n1 * dV) / abs(n1 * dV) < 1e-9
    assert abs(obs["Nb"] - n2 * dV) / abs(n2 * dV) < 1e-9
    assert abs(obs["N_total"] - (n1 + n2) * dV / oc.ne) < 1e-9 * abs(obs["N_total"])
    assert abs(obs["visibility_grid"] - oracle.visibility_grid(oc, ax, bx)) < 1e-9
    # error bars: standard error of the ensemble mean of the per-trajectory populations
    per = np.array([oracle.rotate_sums(oc, ax[j], bx[j], 0.0) for j in range(oc.ne)]) * dV
    se = per.std(axis=0, ddof=1) / np.sqrt(oc.ne)
    assert obs["Na_stderr"] == pytest.approx(se[0], rel=1e-7)
    assert obs["Nb_stderr"] == pytest.approx(se[1], rel=1e-7)
    # projections vs per-cell ensemble sums
    da, db = oracle.cell_density_sums(oc, ax, bx, 0.0)
    assert rel_l2(proj["xy_a"], da.sum(axis=0)) < 1e-9
    assert rel_l2(proj["zy_b"], db.sum(axis=2).T) < 1e-9
    assert rel_l2(proj["axial_a"], da.sum(axis=(1, 2))) < 1e-9


def test_pulses_match_oracle(bec, oracle, gs):
    c, oc = _pair(bec, oracle, 1, ne=2)
    with bec.Solver(c, precision=1) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        a0, b0 = s.get_state(bec.SPACE_X)
        s.pulse(bec.PULSE_PI)
        a1, b1 = s.get_state(bec.SPACE_X)
        s.pulse(bec.PULSE_HALF_PI)
        a2, b2 = s.get_state(bec.SPACE_X)
    ra, rb = oracle.pi_pulse(a0, b0)
    assert rel_l2(a1, ra) < 1e-13 and rel_l2(b1, rb) < 1e-13
    ra, rb = oracle.half_pi_pulse(ra, rb)
    assert rel_l2(a2, ra) < 1e-13 and rel_l2(b2, rb) < 1e-13


def test_graph_propagation_equals_stream_propagation(bec, oracle, gs):
    c, _ = _pair(bec, oracle, 0, ne=4)
    with bec.Solver(c, precision=0) as s1, bec.Solver(c, precision=0) as s2:
        s1.init_evolution(gs, SEED)
        s2.init_evolution(gs, SEED)
        s1.propagate(20)
        s2.propagate_graph(20, 10)
        a1, b1 = s1.get_state(bec.SPACE_K)
        a2, b2 = s2.get_state(bec.SPACE_K)
    # the graph replays 10-step blocks: K/2 K/2 at each seam instead of K (rounding only)
    assert rel_l2(a2, a1) < 1e-5 and rel_l2(b2, b1) < 1e-5


@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 2e-5)])
def test_midpoint_iterations_match_oracle(bec, oracle, gs, prec, tol):
    """itmax = 2 takes the general midpoint path of the fused nonlinear step."""
    c, oc = _pair(bec, oracle, prec, ne=2, itmax=2, detuning=-41.0)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.propagate(10)
        ax, bx = s.get_state(bec.SPACE_X)
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=False)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, 10)
    assert rel_l2(ax, oracle.to_xspace(oc, oa)) < tol
    assert rel_l2(bx, oracle.to_xspace(oc, ob)) < tol


@pytest.mark.parametrize("prec", [0, 1])
def test_zero_midpoint_iterations_give_zero_state(bec, oracle, gs, prec):
    """itmax = 0 never forms the nonlinear factor: the reference's kernel then
    multiplies by an uninitialised factor (quirk Q10); the oracle and the GPU
    both give zero."""
    c, oc = _pair(bec, oracle, prec, ne=2, itmax=0)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        s.propagate(3)
        ax, bx = s.get_state(bec.SPACE_X)
    oa, ob = oracle.init_evolution(oc, gs, SEED, equilibrate=False)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, 3)
    assert not np.any(oa) and not np.any(ob)
    assert not np.any(ax) and not np.any(bx)


def test_autotune_keeps_state_and_results(bec, oracle, gs):
    """The autotuner restores the state; every schedule computes the same step."""
    c, _ = _pair(bec, oracle, 0, ne=8)
    with bec.Solver(c, precision=0) as s:
        s.init_evolution(gs, SEED)
        a0, b0 = s.get_state(bec.SPACE_K)
        best, res = s.autotune(steps=3)
        a1, b1 = s.get_state(bec.SPACE_K)
        assert np.array_equal(a0, a1) and np.array_equal(b0, b1)
        assert len(res) >= 8 and all(r["ms_per_step"] > 0 for r in res)
        outs = []
        for sch in res[:: max(1, len(res) // 6)]:
            s.set_state(a0, b0, bec.SPACE_K)
            s.set_schedule(**{k: sch[k] for k in ("pw", "z_r1", "z_split", "nt", "xcd", "xy_minw", "z_half",
                                                   "xy_split")})
            s.propagate(5)
            outs.append(s.get_state(bec.SPACE_K)[0])
    for o in outs[1:]:
        assert rel_l2(o, outs[0]) < 1e-6


@pytest.mark.parametrize("nxy,nz", [(16, 64), (16, 128), (16, 256), (32, 64), (32, 128), (32, 256)])
@pytest.mark.parametrize("prec", [0, 1])
def test_autotune_every_supported_lattice(bec, nxy, nz, prec):
    """Every candidate schedule of every supported lattice launches (the candidate
    list respects the z pass's LDS and workgroup limits)."""
    c = gpu_calc(bec, N=30000, nvx=nxy, nvy=nxy, nvz=nz, ne=8)
    with bec.Solver(c, precision=prec) as s:
        psi = np.zeros((nz, nxy, nxy), dtype=complex)
        psi[nz // 2, nxy // 2, nxy // 2] = 1.0
        s.init_evolution(psi, SEED, equilibrate=False)
        best, res = s.autotune(steps=1)
        assert res and all(r["ms_per_step"] > 0 for r in res)
        assert best in res or any(abs(r["ms_per_step"] - best["ms_per_step"]) < 1e-12 for r in res)


def test_uneven_groups_match_single_group_runs(bec, oracle, gs):
    """Two grid groups (128x16x16 and 256x32x32) stepped together -- the graph replays
    one branch per group -- give the same states as each group stepped alone."""
    c1, _ = _pair(bec, oracle, 0, ne=8)
    c2 = gpu_calc(bec, **config1_model(oracle, N=70000, nvx=32, nvy=32, nvz=256, ne=8))
    with bec.Solver(c2, precision=1) as g2:
        gs2, rep2, _ = g2.calculate_steady_state()
    with bec.Solver(None, precision=0, groups=[c1, c2]) as s, bec.Solver(c1, precision=0) as s1, \
            bec.Solver(c2, precision=0) as s2:
        s.init_evolution(gs, SEED, group=0)
        s.init_evolution(gs2, SEED + 1, group=1)
        s1.init_evolution(gs, SEED)
        s2.init_evolution(gs2, SEED + 1)
        s.propagate(70)          # exercises the 32-step graph blocks
        s1.propagate(70)
        s2.propagate(70)
        for grp, single in ((0, s1), (1, s2)):
            a, b = s.get_state(bec.SPACE_K, group=grp)
            a1, b1 = single.get_state(bec.SPACE_K)
            assert rel_l2(a, a1) < 1e-6 and rel_l2(b, b1) < 1e-6


@pytest.mark.parametrize("nxy,nz", [(16, 64), (16, 256), (32, 64), (32, 128), (32, 256)])
@pytest.mark.parametrize("prec,tol", [(1, 1e-11), (0, 5e-5)])
def test_propagate_every_lattice_matches_oracle(bec, oracle, nxy, nz, prec, tol):
    """Every supported lattice other than the headline one, with an odd trajectory
    count (no XCD placement) and with its own ground state (split step, reference
    rule, on the GPU in FP64 and checked to equal the oracle's first)."""
    m = config1_model(oracle, N=30000 if nxy == 16 else 70000, nvx=nxy, nvy=nxy, nvz=nz, ne=3, detuning=-41.0)
    c, oc = gpu_calc(bec, **m), oracle_calc(oracle, **m)
    c1 = gpu_calc(bec, **dict(m, ne=1))
    with bec.Solver(c1, precision=1) as g:
        psi, rep, st = g.calculate_steady_state(method=bec.GS_GRADIENT, rule=bec.GS_RULE_RESIDUAL, tol=1e-6,
                                                precision=1, max_iter=20000)
    assert st == 0
    assert oracle.residual(oc, psi)[0] < 1.05e-6
    nsteps = 10
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(psi, SEED, equilibrate=True)
        s.propagate(nsteps)
        ax, bx = s.get_state(bec.SPACE_X)
        obs = s.observables()
    oa, ob = oracle.init_evolution(oc, psi, SEED, equilibrate=True)
    oa, ob = oracle.propagate(oc, oa, ob, oc.dt_evo, nsteps)
    oax, obx = oracle.to_xspace(oc, oa), oracle.to_xspace(oc, ob)
    ea = rel_l2(ax, oax, axis=(1, 2, 3))
    eb = rel_l2(bx, obx, axis=(1, 2, 3))
    assert max(ea.max(), eb.max()) < tol, (ea, eb)
    assert abs(obs["visibility_grid"] - oracle.visibility_grid(oc, oax, obx)) < max(tol, 1e-10) * 10


@pytest.mark.parametrize("prec", [0, 1])
def test_averages_match_oracle(bec, oracle, gs, prec):
    # calculateAverages (wigner.cu:305-337) on the GPU against the oracle's on the
    # same x-space state (the state get_state returns)
    c, oc = _pair(bec, oracle, prec, ne=5)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=True)
        s.propagate(10)
        got = s.averages()
        ax, bx = s.get_state(bec.SPACE_X)
    ref = oracle.averages(oc, ax, bx)
    for k, v in ref.items():
        assert got[k] == pytest.approx(v, rel=1e-11), k


@pytest.mark.parametrize("nxy", [16, 32])
def test_split_xy_pass_equals_two_component_pass(bec, nxy):
    """The split xy pass (one wavefront per component, densities exchanged in LDS)
    performs the same operations per cell as the two-component pass."""
    c = gpu_calc(bec, **config1_model(bec, nvx=nxy, nvy=nxy, ne=8))
    psi = np.zeros((128, nxy, nxy), dtype=complex)
    z, y, x = np.meshgrid(np.arange(128) - 64, np.arange(nxy) - nxy // 2, np.arange(nxy) - nxy // 2, indexing="ij")
    psi[:] = 8.0 * np.exp(-(x ** 2 + y ** 2) / 8.0 - z ** 2 / 800.0)
    outs = []
    with bec.Solver(c, precision=bec.FP32) as s:
        s.init_evolution(psi, SEED, equilibrate=False)
        a0, b0 = s.get_state(bec.SPACE_K)
        cands = [k for k in s.schedule_candidates() if k["xy_split"] == 1]
        assert cands, "split candidates are offered for FP32"
        base = {k: cands[0][k] for k in ("pw", "z_r1", "z_split", "nt", "xcd", "xy_minw", "z_half")}
        for split in (0, 1):
            s.set_state(a0, b0, bec.SPACE_K)
            s.set_schedule(xy_split=split, **base)
            assert s.schedule()["xy_split"] == split
            s.propagate(12)
            outs.append(s.get_state(bec.SPACE_K))
    (a_two, b_two), (a_split, b_split) = outs
    print("split vs two-component: bitwise", np.array_equal(a_two, a_split) and np.array_equal(b_two, b_split),
          "rel", rel_l2(a_split, a_two), rel_l2(b_split, b_two), flush=True)
    assert rel_l2(a_split, a_two) < 1e-6 and rel_l2(b_split, b_two) < 1e-6


@pytest.mark.parametrize("prec,tol", [(1, 1e-14), (0, 1e-6)])
def test_pulse_rotation_matches_oracle(bec, oracle, gs, prec, tol):
    """bec_pulse_rotation(theta, phi) against orc_rotation on the same state; at
    (pi/2, 0) it is the reference's half-pi pulse."""
    c, _ = _pair(bec, oracle, prec, ne=2)
    with bec.Solver(c, precision=prec) as s:
        s.init_evolution(gs, SEED, equilibrate=False)
        a0, b0 = s.get_state(bec.SPACE_K)
        s.pulse_rotation(1.3, -0.8)
        a1, b1 = s.get_state(bec.SPACE_K)
        s.set_state(a0, b0, bec.SPACE_K)
        s.pulse_rotation(np.pi / 2, 0.0)
        a2, b2 = s.get_state(bec.SPACE_K)
        s.set_state(a0, b0, bec.SPACE_K)
        s.pulse(bec.PULSE_HALF_PI)
        a3, b3 = s.get_state(bec.SPACE_K)
    ra, rb = oracle.rotation(a0, b0, 1.3, -0.8)
    assert rel_l2(a1, ra) < tol and rel_l2(b1, rb) < tol
    assert rel_l2(a2, a3) < tol and rel_l2(b2, b3) < tol
