"""Runs the surrogate network's kernels at the production size for counter and
trace runs:

    rocprofv3 --kernel-trace --stats -- python3 -m bec_amd.profile_fno

A few training steps (forward, loss, backward, Adam) on a batch of 16 and a
few single-sample forward passes, on synthetic inputs.  Prints one JSON line
with the wall-clock times per step.
"""
import argparse
import json
import time

import numpy as np

from bec_amd import fno as F
from bec_amd import surrogate as S


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--batch", type=int, default=16)
    a = ap.parse_args()
    net = F.FNO(seed=1, max_batch=a.batch, **S.NET)
    cells = net.cells
    rng = np.random.default_rng(0)
    x = rng.normal(size=(a.batch, cells, S.C_IN)).astype(np.float32)
    t = 0.1 * rng.normal(size=(a.batch, cells, S.C_OUT)).astype(np.float32)
    dx = F.DeviceArray(x.shape, np.float32, x)
    dt = F.DeviceArray(t.shape, np.float32, t)
    dy = F.DeviceArray(t.shape, np.float32)
    net.loss_backward(dx, dt, a.batch)
    net.adam_step(1e-3)
    net.get_params()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.loss_backward(dx, dt, a.batch)
        net.adam_step(1e-3)
    net.get_params()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        net.forward(dx, dy, 1)
    dy.download()
    t2 = time.perf_counter()
    print(json.dumps({"params": net.n_params, "batch": a.batch, "train_step_ms": (t1 - t0) / a.steps * 1e3,
                      "forward_b1_ms": (t2 - t1) / a.steps * 1e3}), flush=True)


if __name__ == "__main__":
    main()
