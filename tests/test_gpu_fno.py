"""GPU: the surrogate network's forward pass and gradients (bec_fno_*, BF16 MFMA
GEMMs with FP32 accumulation) against the FP64 numpy oracle at a small size,
determinism of the gradients, and an optimiser smoke test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(nz=32, ny=8, nx=8, mz=8, my=8, mx=8, width=8, layers=2, c_in=5, c_out=2, hidden=16, max_batch=3)


def _setup(seed=0):
    from bec_amd import fno as F
    from oracle import fno_oracle as FO
    net = F.FNO(seed=seed, **CFG)
    ocfg = FO.Config.from_struct(net.cfg)
    assert net.n_params == FO.param_count(ocfg)
    rng = np.random.default_rng(seed)
    cells = CFG["nz"] * CFG["ny"] * CFG["nx"]
    x = rng.normal(size=(3, cells, CFG["c_in"])).astype(np.float32)
    t = rng.normal(size=(3, cells, CFG["c_out"])).astype(np.float32)
    return F, FO, net, ocfg, x, t


def _grid(a, cfg):
    return a.reshape(a.shape[0], cfg.nz, cfg.ny, cfg.nx, a.shape[-1]).astype(np.float64)


def _blocks(cfg):
    W, H = cfg.width, cfg.hidden
    modes = cfg.mz * cfg.my * cfg.mx
    names = ["P", "p"]
    sizes = [cfg.c_in * W, W]
    for l in range(cfg.layers):
        names += [f"W{l}", f"b{l}", f"R{l}"]
        sizes += [W * W, W, modes * W * W * 2]
    names += ["Q1", "q1", "Q2", "q2"]
    sizes += [W * H, H, H * cfg.c_out, cfg.c_out]
    o = 0
    for n, s in zip(names, sizes):
        yield n, slice(o, o + s)
        o += s


def test_forward_and_gradients_match_oracle(bec):
    F, FO, net, ocfg, x, t = _setup()
    flat = net.get_params().astype(np.float64)
    y_gpu = net.predict(x)
    y_ref = FO.forward(ocfg, flat, _grid(x, ocfg)).reshape(y_gpu.shape)
    err = np.linalg.norm(y_gpu - y_ref) / np.linalg.norm(y_ref)
    print("forward rel L2 vs FP64 oracle:", err)
    assert err < 2e-2
    dx = F.DeviceArray(x.shape, np.float32, x)
    dt = F.DeviceArray(t.shape, np.float32, t)
    loss = net.loss_backward(dx, dt, 3)
    g_gpu = net.get_grads().astype(np.float64)
    loss_ref, g_ref = FO.loss_and_grads(ocfg, flat, _grid(x, ocfg), _grid(t, ocfg))
    assert loss == pytest.approx(loss_ref, rel=2e-2)
    report = {}
    for name, sl in _blocks(ocfg):
        a, b = g_gpu[sl], g_ref[sl]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        report[name] = (round(rel, 5), round(cos, 6))
        assert rel < 5e-2 and cos > 0.998, (name, rel, cos)
    print("gradient blocks (rel L2, cosine) vs FP64 oracle:", report)


def test_gradients_are_deterministic(bec):
    F, FO, net, ocfg, x, t = _setup(1)
    dx = F.DeviceArray(x.shape, np.float32, x)
    dt = F.DeviceArray(t.shape, np.float32, t)
    l1 = net.loss_backward(dx, dt, 3)
    g1 = net.get_grads()
    l2 = net.loss_backward(dx, dt, 3)
    g2 = net.get_grads()
    assert l1 == l2 and np.array_equal(g1, g2)


def test_adam_reduces_the_loss(bec):
    F, FO, net, ocfg, x, t = _setup(2)
    t = 0.3 * np.tanh(x[..., :2])            # a learnable per-cell target
    dx = F.DeviceArray(x.shape, np.float32, x)
    dt = F.DeviceArray(t.shape, np.float32, t)
    losses = []
    for _ in range(60):
        losses.append(net.loss_backward(dx, dt, 3))
        net.adam_step(3e-3)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]


def test_production_size_forward_matches_oracle(bec):
    """The production network at batch 1 (the per-mode channel mixing on its one-row
    kernel, the plane transforms split in K) and at batch 2 (the mixing on the GEMM):
    the forward pass equals the FP64 oracle to the BF16 bar."""
    from bec_amd import fno as F
    from bec_amd import surrogate as S
    from oracle import fno_oracle as FO
    net = F.FNO(seed=5, max_batch=16, **S.NET)
    ocfg = FO.Config.from_struct(net.cfg)
    flat = net.get_params().astype(np.float64)
    th = S.sample_params(2, seed=9)
    x = np.stack([S.features(S.calc_for(t), t) for t in th])
    for b in (1, 2):
        y = net.predict(x[:b])
        ref = FO.forward(ocfg, flat, _grid(x[:b], ocfg)).reshape(y.shape)
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        print(f"batch {b}: forward rel L2 vs FP64 oracle {err:.2e}")
        assert np.abs(ref).max() > 0 and err < 1e-2
