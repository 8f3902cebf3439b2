"""GPU: what the trained surrogate gains over the method without it.

A reduced network (fewer modes and channels than the production one, same
code path) is trained by the library's own kernels on ground states that the
library's solver generates; on held-out parameter sets the solve started from
its guess must reach the same residual as the Thomas-Fermi-started solve in
fewer iterations, both with the library's default flow (tau 1.0, beta 0.85)
and with the slower original one (tau 0.5, beta 0.6).  The reduced network
measured 143 against 175 iterations with the first and 470 to 489 against 782
with the second; the bars are 0.9 and 0.8 of the Thomas-Fermi count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(nz=128, ny=16, nx=16, mz=8, my=8, mx=8, width=16, layers=2, c_in=16, c_out=2, hidden=32, max_batch=8)


@pytest.fixture(scope="module")
def trained(bec):
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    thetas = S.sample_params(72, seed=11)
    X, Y = S.generate(thetas, chunk=72, log=print)
    net = F.FNO(seed=3, **SMALL)
    hist, secs = S.train(net, X[:64], Y[:64], epochs=40, batch=8, lr=3e-3, log=lambda m: None)
    print("training loss", hist[0], "->", hist[-1], f"({secs:.1f} s)")
    return net, thetas, X, Y, hist


def test_trained_network_cuts_iterations(bec, trained):
    from bec_amd import surrogate as S
    net, thetas, X, Y, hist = trained
    assert hist[-1] < 0.1 * hist[0]
    for flow, bar in ((None, 0.9), (dict(tau=0.5, beta=0.6), 0.8)):
        rows = S.evaluate(net, thetas[64:], X[64:], Y[64:], log=print, tf_opts=flow, nn_opts=flow)
        s = S.summary(rows)
        print("held-out, flow", flow or "default", ":", s)
        assert s["converged_all"]
        assert s["final_residual_max"]["nn"] < 1e-6 and s["final_residual_max"]["tf"] < 1e-6
        assert s["nn_iterations_mean"] < bar * s["tf_iterations_mean"]


def test_features_and_targets_layout(bec):
    """The input features and the guess conversion round-trip the solver's layout."""
    from bec_amd import surrogate as S
    th = S.sample_params(1, seed=5)[0]
    c = S.calc_for(th)
    f = S.features(c, th)
    assert f.shape == (128 * 16 * 16, 16)
    out, reps, st = S.solve([c], tol=1e-5)
    assert st == 0
    t = S.targets(c, out[0])
    g = S.guesses(c, t)
    assert np.abs(g - out[0].reshape(2, -1).real).max() < 1e-5 * np.abs(out[0]).max()
    assert S.guess_error(t, t) == 0.0
    # the Thomas-Fermi channel peaks at the trap centre
    tf = f[:, 0].reshape(128, 16, 16)
    assert np.unravel_index(np.argmax(tf), tf.shape)[0] in (63, 64)


def test_device_features_equal_host_features(bec):
    """bec_fno_gs2_features writes the same inputs as the host construction."""
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    for t in S.sample_params(4, seed=17):
        c = S.calc_for(t)
        host = S.features(c, t)
        d = F.DeviceArray((1,) + host.shape, np.float32)
        F.gs2_features(c, S.scaled(t), d)
        dev = d.download()[0]
        d.free()
        exact = float(np.mean(dev == host))
        print(f"device features equal to the host's: {exact:.6f} of the values")
        # FP64 arithmetic in the same order; at most one FP32 rounding step apart
        assert np.all(np.abs(dev - host) <= np.spacing(np.abs(host).astype(np.float32)))
        assert exact > 0.999


def _vis(sums):
    """2 |sum conj(a) b| / (sum |a|^2 + sum |b|^2) from {sum|a|^2, sum|b|^2, Re, Im sum conj(a) b}."""
    return 2.0 * np.hypot(sums[2], sums[3]) / (sums[0] + sums[1])


def test_network_started_state_evolves_like_thomas_fermi_started(bec, oracle, trained):
    """North-star check of the learned start: a held-out parameter set solved to
    residual 1e-6 from the Thomas-Fermi start (default flow) and from the
    network's guess (the network-start flow), then a pi/2 pulse and 400
    real-time steps of 20 us (FP64 state, no noise, no losses): the visibility
    of the guess-started run is within 1e-4 of the Thomas-Fermi-started one and
    of the FP64 oracle's evolution of the Thomas-Fermi-started state, at every
    sample."""
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    net, thetas, X, Y, _ = trained
    th = thetas[64]
    c = S.calc_for(th)
    out_tf, rep_tf, st_tf = S.solve([c], tol=1e-6)
    dx = F.DeviceArray((1, c.cells, S.C_IN), np.float32)
    dy = F.DeviceArray((1, c.cells, S.C_OUT), np.float32)
    F.gs2_features(c, S.scaled(th), dx)
    net.forward(dx, dy, 1)
    g = S.guesses(c, dy.download()[0])[None]
    dx.free()
    dy.free()
    out_nn, rep_nn, st_nn = S.solve([c], guess=g, tol=1e-6, **S.flow(S.NN_FLOW))
    assert st_tf == 0 and st_nn == 0
    assert rep_tf[0]["residual"] < 1e-6 and rep_nn[0]["residual"] < 1e-6
    print("iterations: Thomas-Fermi", rep_tf[0]["iterations"], "guess", rep_nn[0]["iterations"])

    N, fz, fr, a12 = th
    kw = dict(N=int(round(N)), fx=fr, fy=fr, fz=fz, a12=a12, nvx=16, nvy=16, nvz=128, ne=1, Va=0.0, Vb=0.0,
              gamma111=0.0, gamma12=0.0, gamma22=0.0, detuning=0.0, dt_evo=2e-5, itmax=1)
    cg = bec.derive(bec.default_model(**kw))
    assert abs(cg.xmax - c.xmax) < 1e-12 * c.xmax and abs(cg.zmax - c.zmax) < 1e-12 * c.zmax
    steps, every = 400, 100

    def gpu_run(state):
        with bec.Solver(cg, precision=1) as s:
            s.set_state(state[0][None], state[1][None], space=bec.SPACE_X)
            s.pulse(bec.PULSE_HALF_PI)
            return np.array([_vis(v[0]) for v in s.propagate(steps, sample_every=every)])

    v_tf, v_nn = gpu_run(out_tf[0]), gpu_run(out_nn[0])
    oc = oracle.derive(oracle.model(**kw))
    a, b = oracle.half_pi_pulse(out_tf[0][0][None], out_tf[0][1][None])
    ak, bk = oracle.to_kspace(oc, a), oracle.to_kspace(oc, b)
    v_or = []
    for _ in range(steps // every):
        ak, bk = oracle.propagate(oc, ak, bk, oc.dt_evo, every)
        v_or.append(_vis(oracle.traj_sums(oc, oracle.to_xspace(oc, ak), oracle.to_xspace(oc, bk))[0]))
    v_or = np.array(v_or)
    print("visibility: TF start", v_tf, "guess start", v_nn, "oracle", v_or)
    assert np.max(np.abs(v_tf - v_or)) < 1e-10
    assert np.max(np.abs(v_nn - v_tf)) < 1e-4
    assert np.max(np.abs(v_nn - v_or)) < 1e-4
    assert v_or.min() < 0.999   # the visibility moves over the run: the comparison is not of constants


def test_device_start_equals_host_start(bec, trained):
    """bec_ground_states_two_component_net converts the network's output to the
    solver's start on the GPU: the same bits as the host conversion passed as
    guesses, so the same iterations and the same states, for a batch of sets."""
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    net, thetas, X, Y, _ = trained
    ths = thetas[64:67]
    calcs = [S.calc_for(t) for t in ths]
    cells = calcs[0].cells
    dx = F.DeviceArray((len(ths), cells, S.C_IN), np.float32)
    dy = F.DeviceArray((len(ths), cells, S.C_OUT), np.float32)
    for j, t in enumerate(ths):
        F.gs2_features(calcs[j], S.scaled(t), F.DeviceView(dx, j * cells * S.C_IN * 4))
    net.forward(dx, dy, len(ths))
    pred = dy.download()
    g = np.stack([S.guesses(c, p) for c, p in zip(calcs, pred)])
    host, r_host, st_host = S.solve(calcs, guess=g, **S.flow(S.NN_FLOW))
    dev, r_dev, st_dev = S.solve(calcs, net=dy, **S.flow(S.NN_FLOW))
    dx.free()
    dy.free()
    assert st_host == 0 and st_dev == 0
    assert [r["iterations"] for r in r_host] == [r["iterations"] for r in r_dev]
    assert np.array_equal(host, dev)
