"""Launches each hot-pass kernel a few times on a large state, plus a plain
copy kernel of known traffic, for hardware-counter runs:

    rocprofv3 --pmc FETCH_SIZE -- python3 -m bec_amd.profile_passes
    rocprofv3 --pmc WRITE_SIZE -- python3 -m bec_amd.profile_passes
    rocprofv3 --kernel-trace --stats -- python3 -m bec_amd.profile_passes

Prints one JSON line with the event-timed durations.  The copy kernel moves
exactly `copy_bytes` read and `copy_bytes` written, which calibrates the
counters: traffic = counter(pass) * copy_bytes / counter(copy).
"""
import argparse
import json
import sys

import bec_amd as B


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--traj", type=int, default=2048)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--nxy", type=int, default=16, help="lattice side in x and y")
    ap.add_argument("--nz", type=int, default=128, help="lattice size along z")
    ap.add_argument("--atoms", type=int, default=30000)
    ap.add_argument("--autotune", action="store_true", help="time the autotuner's choice of schedule")
    ap.add_argument("--schedule", default="",
                    help="pw,z_r1,z_split,nt,xcd,xy_minw,z_half[,xy_split] (default: built-in)")
    a = ap.parse_args()
    kw = dict(N=a.atoms, nvx=a.nxy, nvy=a.nxy, nvz=a.nz, ne=1, gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14,
              detuning=0.0, dt_evo=4e-5, dt_steady=1e-5, t_equilib=0.0, Va=1.0, Vb=1.0, itmax=1)
    c1 = B.derive(B.default_model(**kw))
    with B.Solver(c1, precision=B.FP64) as g:
        psi, _, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-6,
                                             precision=B.FP64)
    kw["ne"] = a.traj
    c = B.derive(B.default_model(**kw))
    s = B.Solver(c, precision=B.FP32)
    s.init_evolution(psi, 0x5EED, equilibrate=False)
    if a.schedule:
        keys = ("pw", "z_r1", "z_split", "nt", "xcd", "xy_minw", "z_half", "xy_split")
        s.set_schedule(**dict(zip(keys, (int(v) for v in a.schedule.split(",")))))
    if a.autotune:
        s.autotune(steps=8)
    ms = s.time_passes(iters=a.iters)
    copy_bytes = a.copy_mib << 20
    ms_copy = B.time_copy(copy_bytes, a.iters)
    out = {"traj": a.traj, "lattice": [a.nz, a.nxy, a.nxy], "iters": a.iters, "schedule": s.schedule(), "ms": ms, "copy_bytes": copy_bytes,
           "copy_ms": ms_copy, "copy_GBs": 2 * copy_bytes / (ms_copy * 1e-3) / 1e9,
           "pass_bytes": a.traj * 2 * a.nz * a.nxy * a.nxy * 8 * 2}
    s.close()
    print(json.dumps(out))
    sys.stdout.flush()


if __name__ == "__main__":
    main()
