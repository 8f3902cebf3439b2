"""Precision study: FP32 state/FFT/nonlinear step with FP64 reductions against an
all-FP64 GPU run and the FP64 CPU oracle, error stated per observable.

    python -m bec_amd.precision_study [--traj 128] [--steps 20000] [--nz 128 --nxy 32]

Both GPU runs start from the same FP64 ground state and the same Philox seed;
observables (atom numbers after the virtual pulse, visibility on the angle grid
and analytic, mean relative phase) are sampled every --sample-every steps.
The oracle comparison runs on --oracle-traj trajectories for --oracle-steps
steps (CPU time).  Prints one JSON line.
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


def model(nz, nxy, ne):
    return dict(N=30000, nvx=nxy, nvy=nxy, nvz=nz, ne=ne, gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14,
                detuning=0.0, dt_evo=4e-5, dt_steady=1e-5, t_equilib=0.0, Va=1.0, Vb=1.0, itmax=1)


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


def errors(test, ref):
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
    return e


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--traj", type=int, default=128)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--sample-every", type=int, default=250)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--nxy", type=int, default=32)
    ap.add_argument("--seed", type=int, default=0x5EED)
    ap.add_argument("--oracle-traj", type=int, default=2)
    ap.add_argument("--oracle-steps", type=int, default=200)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)

    c1 = B.derive(B.default_model(**model(a.nz, a.nxy, 1)))
    with B.Solver(c1, precision=B.FP64) as g:
        gs, rep, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-8,
                                              precision=B.FP64, max_iter=20000)
    c = B.derive(B.default_model(**model(a.nz, a.nxy, a.traj)))
    o32, t32, a32, b32 = run_gpu(c, gs, B.FP32, a.seed, a.steps, a.sample_every)
    o64, t64, a64, b64 = run_gpu(c, gs, B.FP64, a.seed, a.steps, a.sample_every)
    res = {"lattice": [a.nz, a.nxy, a.nxy], "trajectories": a.traj, "steps": a.steps,
           "sample_every": a.sample_every, "ground_state_residual": rep["residual"],
           "fp32_vs_fp64_gpu": errors(o32, o64),
           "state_rel_l2_fp32_vs_fp64": float(np.sqrt((np.linalg.norm(a32 - a64) ** 2 + np.linalg.norm(b32 - b64) ** 2)
                                                      / (np.linalg.norm(a64) ** 2 + np.linalg.norm(b64) ** 2))),
           "seconds": {"fp32": t32, "fp64": t64}}

    if not a.no_oracle:
        from oracle import oracle as O
        oc = O.derive(O.model(**model(a.nz, a.nxy, a.oracle_traj)))
        cs = B.derive(B.default_model(**model(a.nz, a.nxy, a.oracle_traj)))
        oa, ob = O.init_evolution(oc, gs, a.seed, equilibrate=True)
        oa, ob = O.propagate(oc, oa, ob, oc.dt_evo, a.oracle_steps)
        ox, oy = O.to_xspace(oc, oa), O.to_xspace(oc, ob)
        per = {}
        for prec, name in ((B.FP32, "fp32"), (B.FP64, "fp64")):
            with B.Solver(cs, precision=prec) as s:
                s.init_evolution(gs, a.seed, equilibrate=True)
                s.propagate(a.oracle_steps)
                ga, gb = s.get_state(B.SPACE_X)
                obs = s.observables()
            dV = oc.dx * oc.dy * oc.dz
            n1, n2 = O.rotate_sums(oc, ox, oy, 0.0)
            sums = O.traj_sums(oc, ox, oy).sum(axis=0)
            vac = a.oracle_traj * oc.cells * oc.V / (2 * dV)
            vis = 2 * np.hypot(sums[2], sums[3]) / (sums[0] + sums[1] - 2 * vac)
            per[name] = {
                "state_rel_l2_per_traj_max": float(max(
                    np.linalg.norm(ga[i] - ox[i]) / np.linalg.norm(ox[i]) for i in range(a.oracle_traj))),
                "Na_rel": abs(obs["Na"] - n1 * dV) / abs(n1 * dV),
                "Nb_rel": abs(obs["Nb"] - n2 * dV) / abs(n2 * dV),
                "visibility_grid_abs": abs(obs["visibility_grid"] - O.visibility_grid(oc, ox, oy)),
                "visibility_abs": abs(obs["visibility"] - vis),
                "phase_abs": abs(float(np.angle(np.exp(1j * (obs["phase"] - np.arctan2(sums[3], sums[2])))))),
            }
        res["vs_fp64_oracle"] = {"trajectories": a.oracle_traj, "steps": a.oracle_steps, **per}
    print(json.dumps(res))
    return res


if __name__ == "__main__":
    main()
