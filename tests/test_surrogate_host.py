"""CPU: the learned start's host-side inputs and output conversion
(bec_amd/surrogate.py), checked against what they mean rather than against
their own formulas."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def S():
    from bec_amd import surrogate
    return surrogate


def test_feature_channels(S):
    th = S.sample_params(3, seed=21)
    for t in th:
        c = S.calc_for(t)
        f = S.features(c, t)
        assert f.shape == (c.nvz * c.nvy * c.nvx, S.C_IN) and f.dtype == np.float32
        g = f.reshape(c.nvz, c.nvy, c.nvx, S.C_IN)
        tf, v = g[..., 0].astype(np.float64), g[..., 1].astype(np.float64)
        # Thomas-Fermi amplitude^2 + V/mu = 1 inside the cloud, amplitude 0 outside
        inside = v < 1.0
        assert inside.any() and (~inside).any()
        assert np.allclose(tf[inside] ** 2 + v[inside], 1.0, atol=1e-6)
        assert np.all(tf[~inside] == 0.0)
        # the trap is harmonic: V/mu is a sum of squares of the coordinates
        assert v.min() >= 0.0
        # coordinates: x fastest, z slowest, each spanning [-1, 1] (the lattice includes both ends)
        assert np.allclose(g[0, 0, :, 2], -1 + 2 * np.arange(c.nvx) / (c.nvx - 1), atol=1e-6)
        assert np.allclose(g[0, :, 0, 3], -1 + 2 * np.arange(c.nvy) / (c.nvy - 1), atol=1e-6)
        assert np.allclose(g[:, 0, 0, 4], -1 + 2 * np.arange(c.nvz) / (c.nvz - 1), atol=1e-6)
        assert np.all(g[..., 2] == g[:1, :1, :, 2]) and np.all(g[..., 4] == g[:, :1, :1, 4])
        # parameters scaled into [-1, 1] from their sampling ranges, constant over the lattice
        lo = np.array([r[1] for r in S.RANGES])
        hi = np.array([r[2] for r in S.RANGES])
        want = (2 * (t - lo) / (hi - lo) - 1).astype(np.float32)
        assert np.all(np.abs(want) <= 1.0)
        assert np.all(g[..., 5:9] == want)
        assert np.all(g[..., 9:] == 0.0)


def test_targets_and_guesses_round_trip(S):
    """targets() then guesses() returns the component amplitudes; the difference
    channel carries (a - b)/sqrt 2 scaled by SPIN_WEIGHT."""
    t = S.sample_params(1, seed=4)[0]
    c = S.calc_for(t)
    rng = np.random.default_rng(0)
    u = S.amp_unit(c)
    a = u * rng.random(c.nvz * c.nvy * c.nvx)
    b = u * rng.random(c.nvz * c.nvy * c.nvx)
    state = np.stack([a, b]).reshape(2, c.nvz, c.nvy, c.nvx).astype(np.complex128)
    y = S.targets(c, state)
    assert y.shape == (a.size, 2) and y.dtype == np.float32
    assert np.allclose(y[:, 0], (a + b) / np.sqrt(2) / u, rtol=1e-6)
    assert np.allclose(y[:, 1], S.SPIN_WEIGHT * (a - b) / np.sqrt(2) / u, rtol=1e-6, atol=1e-6)
    g = S.guesses(c, y)
    assert g.shape == (2, a.size) and g.dtype == np.complex128 and np.all(g.imag == 0)
    assert np.abs(g[0].real - a).max() < 1e-6 * u and np.abs(g[1].real - b).max() < 1e-6 * u
    assert S.guess_error(y, y) == 0.0


def test_flow_spec(S):
    assert S.flow("") == {} and S.flow(None) == {}
    assert S.flow("1.0:0.7") == dict(tau=1.0, beta=0.7)
    assert S.flow(S.NN_FLOW)["beta"] < 0.85      # less momentum than the Thomas-Fermi-started flow
