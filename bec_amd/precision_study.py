"""Precision study: FP32 state/FFT/nonlinear step with FP64 reductions against an
all-FP64 GPU run, error stated per observable.

    python -m bec_amd.precision_study [--traj 128] [--steps 40000] [--nz 128 --nxy 32] [--dt 2e-5]

Both runs start from the same FP64 ground state and the same Philox seed;
observables (atom numbers after the virtual pulse, visibility on the angle grid
and analytic, mean relative phase) are sampled every --sample-every steps.
The comparison of both against the FP64 CPU oracle is tests/test_gpu_precision.py.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bec_amd as B  # noqa: E402

OBS = ("Na", "Nb", "N_total", "visibility_grid", "visibility", "phase")


def model(nz, nxy, ne, va=1.0, vb=1.0, dt=4e-5, losses=True):
    g = dict(gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14) if losses else \
        dict(gamma111=0.0, gamma12=0.0, gamma22=0.0)
    return dict(N=30000, nvx=nxy, nvy=nxy, nvz=nz, ne=ne, detuning=0.0, dt_evo=dt, dt_steady=1e-5, t_equilib=0.0,
                Va=va, Vb=vb, itmax=1, **g)


def run_gpu(c, gs, prec, seed, steps, every):
    out = []
    with B.Solver(c, precision=prec) as s:
        s.init_evolution(gs, seed, equilibrate=True)
        out.append(s.observables())
        t0 = time.perf_counter()
        for _ in range(steps // every):
            s.propagate(every)
            out.append(s.observables())
        el = time.perf_counter() - t0
        a, b = s.get_state(B.SPACE_X)
    return out, el, a, b


def errors(test, ref, ne):
    """Largest differences per observable over the samples; with vacuum noise and
    several trajectories also the populations' differences in units of the
    combined standard error of the ensemble means (the noisy trajectories of the
    two precisions separate after a while, so only their statistics can agree)."""
    e = {}
    for k in OBS:
        x = np.array([o[k] for o in test])
        y = np.array([o[k] for o in ref])
        if k == "phase":
            d = np.angle(np.exp(1j * (x - y)))
            e[k] = {"max_abs": float(np.abs(d).max())}
        else:
            scale = np.abs(y).max()
            e[k] = {"max_abs": float(np.abs(x - y).max()), "max_rel": float(np.abs(x - y).max() / scale)}
    if ne > 1:
        for k in ("Na", "Nb"):
            x = np.array([o[k] for o in test]) / ne
            y = np.array([o[k] for o in ref]) / ne
            se = np.hypot([o[k + "_stderr"] for o in test], [o[k + "_stderr"] for o in ref])
            z = np.abs(x - y)[1:] / se[1:]
            e[k]["z_max"], e[k]["z_mean"] = float(z.max()), float(z.mean())
    n12 = np.array([o["N1"] + o["N2"] for o in ref]) / max(ne, 1)
    e["N1_plus_N2_per_trajectory"] = {"first": float(n12[0]), "last": float(n12[-1]), "min": float(n12.min())}
    return e


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--traj", type=int, default=128)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--sample-every", type=int, default=250)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--nxy", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--va", type=float, default=1.0)
    ap.add_argument("--vb", type=float, default=1.0)
    ap.add_argument("--dt", type=float, default=4e-5, help="real-time step, s (128x32x32 is stable at 2e-5)")
    ap.add_argument("--no-losses", action="store_true")
    a = ap.parse_args(argv)

    c1 = B.derive(B.default_model(**model(a.nz, a.nxy, 1, dt=a.dt, losses=not a.no_losses)))
    with B.Solver(c1, precision=B.FP64) as g:
        gs, rep, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-8,
                                              precision=B.FP64, max_iter=20000)
    c = B.derive(B.default_model(**model(a.nz, a.nxy, a.traj, a.va, a.vb, a.dt, not a.no_losses)))
    o32, t32, a32, b32 = run_gpu(c, gs, B.FP32, a.seed, a.steps, a.sample_every)
    o64, t64, a64, b64 = run_gpu(c, gs, B.FP64, a.seed, a.steps, a.sample_every)
    res = {"lattice": [a.nz, a.nxy, a.nxy], "trajectories": a.traj, "steps": a.steps, "dt": a.dt,
           "losses": not a.no_losses, "Va": a.va, "Vb": a.vb,
           "sample_every": a.sample_every, "ground_state_residual": rep["residual"],
           "fp32_vs_fp64_gpu": errors(o32, o64, a.traj),
           "state_rel_l2_fp32_vs_fp64": float(np.sqrt((np.linalg.norm(a32 - a64) ** 2 + np.linalg.norm(b32 - b64) ** 2)
                                                      / (np.linalg.norm(a64) ** 2 + np.linalg.norm(b64) ** 2))),
           "seconds": {"fp32": t32, "fp64": t64}}

    print(json.dumps(res))
    return res


if __name__ == "__main__":
    main()
