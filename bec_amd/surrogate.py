"""Learned ground-state start: parameter sampling, network inputs, training and
the time-to-residual comparison against the Thomas-Fermi start.

    python -m bec_amd.surrogate [--train 2000 --test 200 --epochs 150 --seed 7 --out FILE]

One run generates the two-component ground states with the library's own
solver (bec_ground_states_two_component, FP64, gradient flow), trains the
network of include/bec_fno.h on them with its own BF16 MFMA kernels, and then
solves every held-out parameter set twice to residual 1e-6: from the Thomas-
Fermi start and from the network's guess.  Prints progress lines and one JSON
summary line.

Parameter sets (uniform, seeded): N in [5e3, 1e5], axial trap frequency in
[5, 20] Hz, radial in [50, 150] Hz, a12 in [95, 100] Bohr radii; a11 = 100.4,
a22 = 95.0; half of the atoms in each component; lattice 128 x 16 x 16.

Network inputs per cell (16 channels): the one-component Thomas-Fermi
amplitude sqrt(max(mu - V, 0) / mu), V / mu, the cell coordinates scaled to
[-1, 1] (x, y, z), the four parameters scaled to [-1, 1], and zeros.  Outputs:
the sum and the weighted difference of the two component amplitudes, in units
of sqrt(mu / g11) (to_net / from_net).  The lattice box
scales with mu (xmax = 1.2 sqrt(2 mu)), so these targets are O(1) over the
whole parameter range.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bec_amd as B  # noqa: E402
from bec_amd import fno as F  # noqa: E402

RANGES = (("N", 5e3, 1e5), ("fz", 5.0, 20.0), ("fr", 50.0, 150.0), ("a12", 95.0, 100.0))
C_IN, C_OUT = 16, 2
NET = dict(nz=128, ny=16, nx=16, mz=16, my=16, mx=16, width=32, layers=4, c_in=C_IN, c_out=C_OUT, hidden=64)


def sample_params(n, seed):
    rng = np.random.default_rng(seed)
    lo = np.array([r[1] for r in RANGES])
    hi = np.array([r[2] for r in RANGES])
    return lo + (hi - lo) * rng.random((n, len(RANGES)))


def scaled(theta):
    lo = np.array([r[1] for r in RANGES])
    hi = np.array([r[2] for r in RANGES])
    return 2.0 * (np.asarray(theta) - lo) / (hi - lo) - 1.0


def calc_for(theta, nz=128, nxy=16):
    N, fz, fr, a12 = theta
    m = B.default_model(N=int(round(N)), fx=fr, fy=fr, fz=fz, a12=a12, nvx=nxy, nvy=nxy, nvz=nz, ne=1)
    return B.derive(m)


def features(c, theta):
    """[cells][16] float32 inputs for one parameter set (cell = z ny nx + y nx + x).
    The coordinates stay separable (broadcast, not meshed) until the potential
    is formed; the values are those of the full-grid formulas."""
    x = (-c.xmax + c.dx * np.arange(c.nvx))[None, None, :]
    y = (-c.ymax + c.dy * np.arange(c.nvy))[None, :, None]
    z = (-c.zmax + c.dz * np.arange(c.nvz))[:, None, None]
    V = 0.5 * (x ** 2 + y ** 2 + z ** 2 / c.lambda_ ** 2)
    f = np.empty((c.nvz, c.nvy, c.nvx, C_IN), dtype=np.float32)
    f[..., 0] = np.sqrt(np.maximum(c.mu - V, 0.0) / c.mu)
    f[..., 1] = V / c.mu
    f[..., 2] = x / c.xmax
    f[..., 3] = y / c.ymax
    f[..., 4] = z / c.zmax
    f[..., 5:9] = scaled(theta)
    f[..., 9:] = 0.0
    return f.reshape(-1, C_IN)


# The network predicts the sum s = (a + b)/sqrt 2 and the difference d = (a - b)/sqrt 2 of
# the component amplitudes, with d scaled by SPIN_WEIGHT.  The squared-error loss then
# weighs errors in the relative density SPIN_WEIGHT^2 times more: that mode is soft near
# the miscibility threshold, so its errors are the slow ones for the refining solver,
# while errors of the total density are removed in a few iterations.
SPIN_WEIGHT = 3.0


def amp_unit(c):
    return math.sqrt(c.mu / c.g11)


def to_net(a, b):
    r = 1.0 / math.sqrt(2.0)
    return np.stack([(a + b) * r, (a - b) * r * SPIN_WEIGHT], axis=-1)


def from_net(y):
    r = 1.0 / math.sqrt(2.0)
    s, d = y[..., 0], y[..., 1] / SPIN_WEIGHT
    return (s + d) * r, (s - d) * r


def targets(c, state):
    """state [2][nz][ny][nx] complex (x-space ground state) -> [cells][2] float32 network targets."""
    u = amp_unit(c)
    return to_net(state[0].real.ravel() / u, state[1].real.ravel() / u).astype(np.float32)


def guesses(c, pred):
    """Network output [cells][2] -> solver guess [2][cells] complex."""
    u = amp_unit(c)
    a, b = from_net(np.asarray(pred, dtype=np.float64))
    g = np.zeros((2, a.shape[0]), dtype=np.complex128)
    g.real[0] = a * u
    g.real[1] = b * u
    return g


def guess_error(pred, target):
    """Relative L2 error of the predicted component amplitudes."""
    pa, pb = from_net(np.asarray(pred, dtype=np.float64))
    ta, tb = from_net(np.asarray(target, dtype=np.float64))
    return float(np.sqrt((np.sum((pa - ta) ** 2) + np.sum((pb - tb) ** 2)) / (np.sum(ta ** 2) + np.sum(tb ** 2))))


def solve(calcs, guess=None, tol=1e-6, max_iter=20000, net=None, **opts):
    """Two-component ground states, FP64; opts (tau, beta, alpha) override the coupled flow's defaults.
    guess: host guesses [B][2][cells]; net: the network's output left on the device
    ([B][cells][C_OUT]), converted to the start on the GPU."""
    out, reps, st = B.ground_states_two_component(calcs, guesses=guess, net=net, spin_weight=SPIN_WEIGHT, tol=tol,
                                                  max_iter=max_iter, precision=B.FP64, **opts)
    return out, reps, st


def generate(thetas, tol=1e-7, chunk=200, log=print):
    """Ground states of many parameter sets, solved in batches on the GPU."""
    n = len(thetas)
    X = np.empty((n, 128 * 256, C_IN), dtype=np.float32)
    Y = np.empty((n, 128 * 256, C_OUT), dtype=np.float32)
    for s in range(0, n, chunk):
        part = thetas[s:s + chunk]
        calcs = [calc_for(t) for t in part]
        out, reps, st = solve(calcs, tol=tol, max_iter=40000)
        if st != 0:
            raise RuntimeError(f"ground states did not converge (status {st})")
        for i, (c, t) in enumerate(zip(calcs, part)):
            X[s + i] = features(c, t)
            Y[s + i] = targets(c, out[i])
        log(f"generated {min(s + chunk, n)}/{n} ground states "
            f"(iterations max {max(r['iterations'] for r in reps)})")
    return X, Y


def train(net, X, Y, epochs, batch=16, lr=1e-3, lr_min=1e-5, seed=0, log=print, time_limit=None):
    """Adam with a cosine learning rate; the whole set is resident on the device."""
    n = X.shape[0]
    dX = F.DeviceArray(X.shape, np.float32, X)
    dY = F.DeviceArray(Y.shape, np.float32, Y)
    per_x = X.shape[1] * X.shape[2] * 4
    per_y = Y.shape[1] * Y.shape[2] * 4
    rng = np.random.default_rng(seed)
    nb = n // batch
    total = epochs * nb
    step = 0
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = rng.permutation(nb)
        acc = 0.0
        for j in order:
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, total)))
            vx = F.DeviceView(dX, j * batch * per_x)
            vy = F.DeviceView(dY, j * batch * per_y)
            acc += net.loss_backward(vx, vy, batch)
            net.adam_step(cur)
            step += 1
        hist.append(acc / nb)
        el = time.perf_counter() - t0
        log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.3e} lr {cur:.2e} {el:.0f}s")
        if time_limit and el > time_limit:
            log("time limit reached")
            break
    dX.free()
    dY.free()
    return hist, time.perf_counter() - t0


def predict(net, X, batch=16):
    out = np.empty((X.shape[0], X.shape[1], C_OUT), dtype=np.float32)
    for s in range(0, X.shape[0], batch):
        out[s:s + batch] = net.predict(X[s:s + batch])
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# Step and momentum of the coupled flow when it starts from the network's guess.
# The guess is already close (relative L2 error about 4e-3), so less momentum
# than the Thomas-Fermi-started default converges sooner.  Mean iterations to
# residual 1e-6 over 200 held-out sets: 1.0:0.75 70.8, 1.0:0.7 72.9,
# 1.0:0.8 81.7, 0.9:0.7 79.6, 1.0:0.6 93.8, the default 1.0:0.85 109.3.
# tau 1.1 saves 3% more (1.1:0.7 67.9, 1.1:0.75 68.5) but sits next to the
# edge: 1.2:0.65, 1.2:0.5, 1.5:0.3, 1.0:0.4 and 1.0:0.0 did not all converge.
NN_FLOW = "1.0:0.75"


def flow(spec):
    """'tau:beta' -> solve() options ({} for an empty spec: the library's defaults)."""
    if not spec:
        return {}
    tau, beta = (float(v) for v in spec.split(":"))
    return dict(tau=tau, beta=beta)


def evaluate(net, thetas, X, Y, tol=1e-6, log=print, tf_opts=None, nn_opts=None, device_start=True):
    """Per held-out set: solve from TF and from the network guess (one set per solve).
    inference_ms is the host wall clock of the guess: input features (built on
    the device) and forward pass, then either (device_start) the solver reads the
    network's output on the device and converts it in a kernel before its timed
    iterations, as it fills the Thomas-Fermi start, or the output is downloaded
    and converted to the solver's layout on the host (included in inference_ms).
    tf_opts / nn_opts: the flow's step and momentum for each start."""
    rows = []
    cells = X.shape[1]
    dx = F.DeviceArray((1, cells, C_IN), np.float32)
    dy = F.DeviceArray((1, cells, C_OUT), np.float32)
    for i, t in enumerate(thetas):
        c = calc_for(t)
        _, r_tf, st_tf = solve([c], tol=tol, **(tf_opts or {}))
        t0 = time.perf_counter()
        F.gs2_features(c, scaled(t), dx)
        net.forward(dx, dy, 1)
        if device_start:
            F.device_synchronize()
            t_inf = (time.perf_counter() - t0) * 1e3
            _, r_nn, st_nn = solve([c], net=dy, tol=tol, **(nn_opts or {}))
            pred = dy.download()[0]
        else:
            pred = dy.download()[0]
            g = guesses(c, pred)[None]
            t_inf = (time.perf_counter() - t0) * 1e3
            _, r_nn, st_nn = solve([c], guess=g, tol=tol, **(nn_opts or {}))
        rows.append({"theta": [float(v) for v in t], "tf_iters": r_tf[0]["iterations"],
                     "tf_ms": r_tf[0]["elapsed_ms"], "tf_residual": r_tf[0]["residual"],
                     "nn_iters": r_nn[0]["iterations"], "nn_ms": r_nn[0]["elapsed_ms"],
                     "nn_residual": r_nn[0]["residual"], "inference_ms": t_inf,
                     "guess_rel_l2": guess_error(pred, Y[i]), "converged": bool(st_tf == 0 and st_nn == 0)})
        if (i + 1) % 50 == 0:
            log(f"evaluated {i + 1}/{len(thetas)}")
    return rows


def evaluate_batched(net, thetas, tol=1e-6, tf_opts=None, nn_opts=None, device_start=True):
    """All sets in one batched solve from each start (host wall clock).  The
    network-started time includes the guesses, made in batches of the
    network's max_batch (features on the device, forward pass) and left on the
    device for the solver's start kernel (device_start), or downloaded and
    converted on the host."""
    calcs = [calc_for(t) for t in thetas]
    n, cells = len(calcs), calcs[0].cells
    solve(calcs[:2], tol=tol, **(tf_opts or {}))     # first-call set-up outside the timing
    t0 = time.perf_counter()
    _, r_tf, st_tf = solve(calcs, tol=tol, **(tf_opts or {}))
    t_tf = (time.perf_counter() - t0) * 1e3
    b = net.cfg.max_batch
    dx = F.DeviceArray((b, cells, C_IN), np.float32)
    dy = F.DeviceArray((n if device_start else b, cells, C_OUT), np.float32)
    g = None if device_start else np.empty((n, 2, cells), dtype=np.complex128)
    t0 = time.perf_counter()
    for s0 in range(0, n, b):
        m = min(b, n - s0)
        for j in range(m):
            F.gs2_features(calcs[s0 + j], scaled(thetas[s0 + j]), F.DeviceView(dx, j * cells * C_IN * 4))
        if device_start:
            net.forward(dx, F.DeviceView(dy, s0 * cells * C_OUT * 4), m)
        else:
            net.forward(dx, dy, m)
            pred = dy.download()
            for j in range(m):
                g[s0 + j] = guesses(calcs[s0 + j], pred[j])
    if device_start:
        F.device_synchronize()
    t_inf = (time.perf_counter() - t0) * 1e3
    _, r_nn, st_nn = solve(calcs, guess=g, net=dy if device_start else None, tol=tol, **(nn_opts or {}))
    t_nn = (time.perf_counter() - t0) * 1e3
    dx.free()
    dy.free()
    it_tf = [r["iterations"] for r in r_tf]
    it_nn = [r["iterations"] for r in r_nn]
    return {"sets": n, "converged_all": bool(st_tf == 0 and st_nn == 0),
            "tf_ms": t_tf, "nn_ms_incl_inference": t_nn, "inference_ms": t_inf, "speedup_time": t_tf / t_nn,
            "tf_iterations_mean": float(np.mean(it_tf)), "tf_iterations_max": int(max(it_tf)),
            "nn_iterations_mean": float(np.mean(it_nn)), "nn_iterations_max": int(max(it_nn)),
            "final_residual_max": {"tf": float(max(r["residual"] for r in r_tf)),
                                   "nn": float(max(r["residual"] for r in r_nn))}}


def summary(rows):
    tf_ms = np.array([r["tf_ms"] for r in rows])
    nn_ms = np.array([r["nn_ms"] + r["inference_ms"] for r in rows])
    return {
        "held_out": len(rows),
        "converged_all": all(r["converged"] for r in rows),
        "tf_iterations_mean": float(np.mean([r["tf_iters"] for r in rows])),
        "nn_iterations_mean": float(np.mean([r["nn_iters"] for r in rows])),
        "tf_ms_mean": float(tf_ms.mean()),
        "nn_ms_mean_incl_inference": float(nn_ms.mean()),
        "inference_ms_mean": float(np.mean([r["inference_ms"] for r in rows])),
        "speedup_mean_time": float(tf_ms.mean() / nn_ms.mean()),
        "speedup_median_per_set": float(np.median(tf_ms / nn_ms)),
        "speedup_iterations": float(np.sum([r["tf_iters"] for r in rows]) / np.sum([r["nn_iters"] for r in rows])),
        "final_residual_max": {"tf": float(max(r["tf_residual"] for r in rows)),
                               "nn": float(max(r["nn_residual"] for r in rows))},
        "guess_rel_l2_mean": float(np.mean([r["guess_rel_l2"] for r in rows])),
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--time-limit", type=float, default=None, help="training wall-clock limit, seconds")
    ap.add_argument("--out", default=None, help="write the JSON summary (with per-set rows) here")
    ap.add_argument("--tf", default=None, help="tau:beta of the flow from Thomas-Fermi (default: the library's)")
    ap.add_argument("--nn", default=NN_FLOW, help="tau:beta of the flow from the network guess")
    ap.add_argument("--nn-sweep", default="", help="extra tau:beta settings of the network-started flow to report")
    ap.add_argument("--legacy", default="", help="tau:beta of both starts for a second comparison on the same network")
    ap.add_argument("--batched", action="store_true", help="also solve all held-out sets in one batch per start")
    ap.add_argument("--save-net", default=None, help="write the trained parameters (float32 .npy) here")
    ap.add_argument("--load-net", default=None, help="skip training and load the parameters from this .npy")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    thetas = sample_params(a.train + a.test, a.seed)
    t0 = time.perf_counter()
    X, Y = generate(thetas, log=log)
    t_gen = time.perf_counter() - t0
    net = F.FNO(seed=a.seed, max_batch=a.batch, **NET)
    log(f"network parameters {net.n_params}")
    if a.load_net:
        net.set_params(np.load(a.load_net))
        hist, t_train = [float("nan")], 0.0
    else:
        hist, t_train = train(net, X[:a.train], Y[:a.train], a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                              log=log, time_limit=a.time_limit)
    if a.save_net:
        np.save(a.save_net, np.asarray(net.get_params(), dtype=np.float32))
    tf_opts, nn_opts = flow(a.tf), flow(a.nn)
    rows = evaluate(net, thetas[a.train:], X[a.train:], Y[a.train:], log=log, tf_opts=tf_opts, nn_opts=nn_opts)
    res = {"config": dict(NET, train=a.train, test=a.test, epochs_run=len(hist), batch=a.batch, lr=a.lr,
                          seed=a.seed, tf_flow=tf_opts or "default", nn_flow=nn_opts or "default"),
           "generate_s": t_gen, "train_s": t_train, "loss_first": hist[0], "loss_last": hist[-1],
           "train_loss": hist, "summary": summary(rows)}
    sweep = {}
    for spec in [v for v in a.nn_sweep.split(",") if v]:
        r2 = evaluate(net, thetas[a.train:], X[a.train:], Y[a.train:], log=log, tf_opts=tf_opts, nn_opts=flow(spec))
        sweep[spec] = summary(r2)
        log(f"network-started flow {spec}: {json.dumps(sweep[spec])}")
    if sweep:
        res["nn_sweep"] = sweep
    if a.legacy:
        lo = flow(a.legacy)
        r3 = evaluate(net, thetas[a.train:], X[a.train:], Y[a.train:], log=log, tf_opts=lo, nn_opts=lo)
        res["both_starts_" + a.legacy] = summary(r3)
        log(f"both starts with {a.legacy}: {json.dumps(res['both_starts_' + a.legacy])}")
    if a.batched:
        res["batched"] = evaluate_batched(net, thetas[a.train:], tf_opts=tf_opts, nn_opts=nn_opts)
        log(f"batched: {json.dumps(res['batched'])}")
    if a.out:
        with open(a.out, "w") as f:
            json.dump(dict(res, rows=rows), f)
    print(json.dumps({k: v for k, v in res.items() if k != "train_loss"}))
    return res


if __name__ == "__main__":
    main()
