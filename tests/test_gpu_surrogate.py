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


def test_trained_network_cuts_iterations(bec):
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    thetas = S.sample_params(72, seed=11)
    X, Y = S.generate(thetas, chunk=72, log=print)
    net = F.FNO(seed=3, **SMALL)
    hist, secs = S.train(net, X[:64], Y[:64], epochs=40, batch=8, lr=3e-3, log=lambda m: None)
    print("training loss", hist[0], "->", hist[-1], f"({secs:.1f} s)")
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
