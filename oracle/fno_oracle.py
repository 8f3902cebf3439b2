"""FP64 numpy restatement of the surrogate network of include/bec_fno.h, forward
and backward, used only by tests/ to check the GPU kernels.  TEST
INFRASTRUCTURE: nothing in the product imports this module.

The spectral path is written with numpy.fft (orthonormal 3D transform, modes
selected by index) rather than the DFT-matrix GEMMs of the GPU path, and the
backward pass with the adjoints derived by hand:
  u_spec = Re A^H (R . A h)    (A: orthonormal DFT restricted to the kept modes)
  dT = A du,  dS = dT conj(R)^T,  dR = sum_b conj(S)^T dT,  dh = Re A^H dS
The hand-derived gradients are themselves pinned by central finite
differences (tests/test_oracle_fno.py).
"""
import math

import numpy as np


class Config:
    def __init__(self, nz, ny, nx, mz, my, mx, width, layers, c_in, c_out, hidden):
        self.nz, self.ny, self.nx = nz, ny, nx
        self.mz, self.my, self.mx = mz, my, mx
        self.width, self.layers = width, layers
        self.c_in, self.c_out, self.hidden = c_in, c_out, hidden

    @classmethod
    def from_struct(cls, s):
        return cls(s.nz, s.ny, s.nx, s.mz, s.my, s.mx, s.width, s.layers, s.c_in, s.c_out, s.hidden)


def _axis_modes(n, m, centered_split):
    """Indices of the kept frequencies along one axis, in the GPU's order."""
    half = m // 2 if centered_split else (m + 1) // 2
    ks = [i if i < half else i - m for i in range(m)]
    return np.array([k % n for k in ks])


def mode_index(cfg):
    return (_axis_modes(cfg.nz, cfg.mz, True), _axis_modes(cfg.ny, cfg.my, False),
            _axis_modes(cfg.nx, cfg.mx, False))


def unpack(cfg, flat):
    W, H = cfg.width, cfg.hidden
    o = 0

    def take(n, shape):
        nonlocal o
        a = flat[o:o + n].reshape(shape)
        o += n
        return a

    p = {"P": take(cfg.c_in * W, (cfg.c_in, W)), "p": take(W, (W,)), "W": [], "b": [], "R": []}
    modes = cfg.mz * cfg.my * cfg.mx
    for _ in range(cfg.layers):
        p["W"].append(take(W * W, (W, W)))
        p["b"].append(take(W, (W,)))
        r = take(modes * W * W * 2, (cfg.mz, cfg.my, cfg.mx, W, W, 2))
        p["R"].append(r[..., 0] + 1j * r[..., 1])
    p["Q1"] = take(W * H, (W, H))
    p["q1"] = take(H, (H,))
    p["Q2"] = take(H * cfg.c_out, (H, cfg.c_out))
    p["q2"] = take(cfg.c_out, (cfg.c_out,))
    assert o == flat.size
    return p


def pack_grads(cfg, g):
    parts = [g["P"].ravel(), g["p"].ravel()]
    for l in range(cfg.layers):
        parts += [g["W"][l].ravel(), g["b"][l].ravel(),
                  np.stack([g["R"][l].real, g["R"][l].imag], axis=-1).ravel()]
    parts += [g["Q1"].ravel(), g["q1"].ravel(), g["Q2"].ravel(), g["q2"].ravel()]
    return np.concatenate(parts)


def param_count(cfg):
    W, H = cfg.width, cfg.hidden
    modes = cfg.mz * cfg.my * cfg.mx
    return cfg.c_in * W + W + cfg.layers * (W * W + W + modes * W * W * 2) + W * H + H + H * cfg.c_out + cfg.c_out


def gelu(x):
    from scipy.special import erf
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    from scipy.special import erf
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def _restrict(cfg, F):
    iz, iy, ix = mode_index(cfg)
    return F[:, iz][:, :, iy][:, :, :, ix]


def _scatter(cfg, S):
    iz, iy, ix = mode_index(cfg)
    B, W = S.shape[0], S.shape[-1]
    F = np.zeros((B, cfg.nz, cfg.ny, cfg.nx, W), dtype=complex)
    F[np.ix_(np.arange(B), iz, iy, ix, np.arange(W))] = S
    return F


def spectral(cfg, h, R):
    """Re A^H (R . A h); returns (u_spec, S) with S = A h on the kept modes."""
    S = _restrict(cfg, np.fft.fftn(h, axes=(1, 2, 3), norm="ortho"))
    T = np.einsum("bzyxi,zyxio->bzyxo", S, R)
    u = np.fft.ifftn(_scatter(cfg, T), axes=(1, 2, 3), norm="ortho").real
    return u, S


def forward(cfg, flat, x, keep=False):
    p = unpack(cfg, np.asarray(flat, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    h = x @ p["P"] + p["p"]
    cache = {"h": [h], "u": [], "S": []}
    for l in range(cfg.layers):
        us, S = spectral(cfg, h, p["R"][l])
        u = us + h @ p["W"][l] + p["b"][l]
        h = gelu(u) if l < cfg.layers - 1 else u
        cache["u"].append(u)
        cache["S"].append(S)
        cache["h"].append(h)
    g1 = h @ p["Q1"] + p["q1"]
    a1 = gelu(g1)
    y = a1 @ p["Q2"] + p["q2"]
    cache.update(g1=g1, a1=a1, p=p)
    return (y, cache) if keep else y


def loss_and_grads(cfg, flat, x, target):
    """Mean squared error over batch, cells and outputs, and its gradient (flat)."""
    x = np.asarray(x, dtype=np.float64)
    y, c = forward(cfg, flat, x, keep=True)
    p = c["p"]
    n = y.size
    d = y - target
    loss = float(np.sum(d * d) / n)
    dy = 2.0 * d / n
    red = tuple(range(y.ndim - 1))

    def tdot(a, b):   # sum over all leading axes of a^T b
        return np.tensordot(a, b, axes=(red, red))

    g = {"W": [None] * cfg.layers, "b": [None] * cfg.layers, "R": [None] * cfg.layers}
    g["Q2"] = tdot(c["a1"], dy)
    g["q2"] = dy.sum(axis=red)
    dg1 = (dy @ p["Q2"].T) * gelu_grad(c["g1"])
    g["Q1"] = tdot(c["h"][-1], dg1)
    g["q1"] = dg1.sum(axis=red)
    dh = dg1 @ p["Q1"].T
    for l in reversed(range(cfg.layers)):
        du = dh * gelu_grad(c["u"][l]) if l < cfg.layers - 1 else dh
        hin = c["h"][l]
        g["W"][l] = tdot(hin, du)
        g["b"][l] = du.sum(axis=red)
        dT = _restrict(cfg, np.fft.fftn(du, axes=(1, 2, 3), norm="ortho"))
        g["R"][l] = np.einsum("bzyxi,bzyxo->zyxio", np.conj(c["S"][l]), dT)
        dS = np.einsum("bzyxo,zyxio->bzyxi", dT, np.conj(p["R"][l]))
        dh = du @ p["W"][l].T + np.fft.ifftn(_scatter(cfg, dS), axes=(1, 2, 3), norm="ortho").real
    g["P"] = tdot(x, dh)
    g["p"] = dh.sum(axis=red)
    return loss, pack_grads(cfg, g)
