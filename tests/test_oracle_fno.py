"""CPU: pins the FP64 surrogate-network oracle (oracle/fno_oracle.py).

- Its hand-derived gradients agree with central finite differences of its own
  loss, parameter by parameter, for every parameter block (a wrong adjoint,
  a missing conjugate or a transposed weight fails this).
- Its spectral layer with every mode kept and R = identity is the identity map
  (the orthonormal transform pair), and with R = c I it is multiplication by
  Re c; keeping no high z modes removes them exactly.
"""
import numpy as np
import pytest

from oracle import fno_oracle as FO


def _tiny():
    return FO.Config(nz=8, ny=4, nx=4, mz=4, my=4, mx=2, width=4, layers=2, c_in=3, c_out=2, hidden=5)


def test_param_layout_roundtrip():
    cfg = _tiny()
    n = FO.param_count(cfg)
    flat = np.arange(n, dtype=np.float64)
    p = FO.unpack(cfg, flat)
    g = {k: p[k] for k in ("P", "p", "Q1", "q1", "Q2", "q2")}
    g["W"], g["b"], g["R"] = p["W"], p["b"], p["R"]
    assert np.array_equal(FO.pack_grads(cfg, g), flat)


def test_gradients_match_finite_differences():
    cfg = _tiny()
    rng = np.random.default_rng(0)
    n = FO.param_count(cfg)
    flat = rng.normal(size=n) * 0.4
    x = rng.normal(size=(2, cfg.nz, cfg.ny, cfg.nx, cfg.c_in))
    t = rng.normal(size=(2, cfg.nz, cfg.ny, cfg.nx, cfg.c_out))
    loss, g = FO.loss_and_grads(cfg, flat, x, t)
    # every block: the first, a middle and the last entry of each
    W, H = cfg.width, cfg.hidden
    modes = cfg.mz * cfg.my * cfg.mx
    sizes = [cfg.c_in * W, W] + [W * W, W, modes * W * W * 2] * cfg.layers + [W * H, H, H * cfg.c_out, cfg.c_out]
    idx, o = [], 0
    for s in sizes:
        idx += sorted({o, o + s // 2, o + s - 1, o + s // 3})
        o += s
    eps = 1e-6
    for i in idx:
        e = np.zeros(n)
        e[i] = eps
        lp, _ = FO.loss_and_grads(cfg, flat + e, x, t)
        lm, _ = FO.loss_and_grads(cfg, flat - e, x, t)
        fd = (lp - lm) / (2 * eps)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9), (i, fd, g[i])


def test_spectral_identity_and_truncation():
    full = FO.Config(nz=8, ny=4, nx=4, mz=8, my=4, mx=4, width=3, layers=1, c_in=1, c_out=1, hidden=1)
    rng = np.random.default_rng(1)
    h = rng.normal(size=(2, 8, 4, 4, 3))
    eye = np.broadcast_to(np.eye(3), (8, 4, 4, 3, 3)).astype(complex)
    u, _ = FO.spectral(full, h, eye)
    assert np.abs(u - h).max() < 1e-12
    u2, _ = FO.spectral(full, h, (0.5 - 2j) * eye)
    assert np.abs(u2 - 0.5 * h).max() < 1e-12
    # keep |kz| < 2 only: the result has no z frequency beyond 1
    low = FO.Config(nz=8, ny=4, nx=4, mz=4, my=4, mx=4, width=3, layers=1, c_in=1, c_out=1, hidden=1)
    u3, _ = FO.spectral(low, h, np.broadcast_to(np.eye(3), (4, 4, 4, 3, 3)).astype(complex))
    spec = np.fft.fft(u3, axis=1)
    assert np.abs(spec[:, 3:6]).max() < 1e-12          # k = 3, 4, 5 (= -3) removed
    assert np.abs(spec[:, 2]).max() > 1e-3               # k = 2 kept
