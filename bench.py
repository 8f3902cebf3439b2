#!/usr/bin/env python3
"""Benchmark of the split-step Wigner GPE propagator on one MI355X.

Workload (the headline configuration): 64 trajectories of a two-component
Rb-87 condensate on a 128x16x16 lattice, complex64 state, N = 30000, losses
K111 = 5.4e-30 cm^6/s, K12 = 0.78e-13 cm^3/s, K22 = 1.5e-14 cm^3/s, Philox
vacuum noise (seed 0x5EED), pi/2 pulse, dt = 40 us; ensemble sums sampled
every 250 steps.  One "step" = one propagate() over all trajectories.

Prints one JSON line: trajectory-steps/s (whole job), the HBM roofline of the
dominant pass kernel measured with HIP events on a >= 1 GiB state (2048
trajectories, beyond the 256 MB Infinity Cache), ground-state time to
residual 1e-6, and the FP64 CPU split-step timed on this host as a baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ALGO_BYTES_PER_TRAJ_STEP = 2 * 32768 * 8 * 2   # read + write both components once (128x16x16, c64)
HBM_PEAK = 8.0e12
BF16_PEAK = 2.5e15   # MI355X dense BF16 matrix peak
SEED = 0x5EED


def log(msg):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def config1(**kw):
    d = dict(N=30000, nvx=16, nvy=16, nvz=128, ne=64, gamma111=5.4e-30, gamma12=0.78e-13, gamma22=1.5e-14,
             detuning=0.0, dt_evo=4e-5, dt_steady=1e-5, t_equilib=0.0, Va=1.0, Vb=1.0, itmax=1)
    d.update(kw)
    return d


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic():
    """HBM bytes per launch of each pass, from the committed rocprofv3 counter summary."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic_passes_2048traj_128x16x16.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return {k: v["bytes_per_launch"] for k, v in d.get("kernels", {}).items()}


def uneven_groups(B, psi_a, steps, autotune=True):
    """256 trajectories at 128x16x16 (N = 30000) and 128 at 256x32x32 (N = 70000), both at 20 us
    steps (the 256x32x32 lattice is unstable at 40 us), no sampling, complex64.  Each launch form
    is timed over the same steps:
      plain_per_grid      one context per grid, one after the other, default schedules, no graphs
      tuned_per_grid      one context per grid, one after the other, autotuned schedules, graphs
      branches            one context, a concurrent graph branch per grid, autotuned schedules
      one_launch          one context, one launch per pass over both grids' tiles (multi-group
                          table, fixed schedules), graphs
      one_launch_no_graph the same launch without graphs
    The schedules are tuned once per grid and given to every form that uses them."""
    ka = config1(ne=256, dt_evo=2e-5)
    kb = config1(ne=128, N=70000, nvx=32, nvy=32, nvz=256, dt_evo=2e-5)
    ca = B.derive(B.default_model(**ka))
    cb1 = B.derive(B.default_model(**dict(kb, ne=1)))
    with B.Solver(cb1, precision=B.FP64) as g:
        psi_b, rep_b, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-6,
                                                   precision=B.FP64, max_iter=20000)
    cb = B.derive(B.default_model(**kb))
    grids = (("128x16x16", ca, psi_a), ("256x32x32", cb, psi_b))
    ntot = ca.ne + cb.ne

    def timed(solver, n):
        solver.propagate(64)
        solver.synchronize()
        t0 = time.perf_counter()
        solver.propagate(n)
        solver.synchronize()
        return time.perf_counter() - t0

    def tps(sec):
        return ntot * steps / sec

    def one_grid(c, psi, sched, graphs):
        p = B.Solver(c, precision=B.FP32)
        p.set_graphs(graphs)
        p.init_evolution(psi, SEED, equilibrate=False)
        if sched is not None:
            p.set_schedule(**sched)
        sec = timed(p, steps)
        p.close()
        return sec

    scheds = {}
    for name, c, psi in grids:
        p = B.Solver(c, precision=B.FP32)
        p.init_evolution(psi, SEED, equilibrate=False)
        best = p.autotune(steps=8)[0] if autotune else p.schedule()
        scheds[name] = {k: v for k, v in best.items() if k != "ms_per_step"}
        p.close()
    forms = {}
    sec = {n: one_grid(c, psi, None, False) for n, c, psi in grids}
    forms["plain_per_grid"] = {"traj_steps_per_s": tps(sum(sec.values())), "seconds_per_grid": sec}
    sec = {n: one_grid(c, psi, scheds[n], True) for n, c, psi in grids}
    forms["tuned_per_grid"] = {"traj_steps_per_s": tps(sum(sec.values())), "seconds_per_grid": sec}
    for form, multi, graphs in (("branches", False, True), ("one_launch", True, True),
                                ("one_launch_no_graph", True, False)):
        s = B.Solver(ca, precision=B.FP32, groups=[ca, cb])
        s.init_evolution(psi_a, SEED, group=0, equilibrate=False)
        s.init_evolution(psi_b, SEED, group=1, equilibrate=False)
        if not multi:
            for g, (n, _, _) in enumerate(grids):
                s.set_schedule(group=g, **scheds[n])
        s.set_graphs(graphs)
        s.set_multi_launch(multi)
        forms[form] = {"traj_steps_per_s": tps(timed(s, steps))}
        s.close()
    best_sep = max(forms[k]["traj_steps_per_s"] for k in ("plain_per_grid", "tuned_per_grid", "branches"))
    return {"workload": "256 traj x 128x16x16 (N=30000) + 128 traj x 256x32x32 (N=70000), complex64, dt 20 us",
            "steps": steps, "forms": forms, "schedules": scheds,
            "one_launch_over_best_other": forms["one_launch"]["traj_steps_per_s"] / best_sep,
            "ground_state_256x32x32": {"iterations": rep_b["iterations"], "residual": rep_b["residual"]}}


SURROGATE_SUMMARIES = [os.path.join(ROOT, "profiles", n) for n in
                       ("surrogate_groundstate_200heldout.json", "surrogate_flow_sweep_200heldout.json")]


def program_2d(B, with_cpu):
    """Config 0: the 2D program (programs/bec_wigner_2d.cpp) for the Rb-87 cloud of 30k atoms
    (a11 = 100.4, a22 = 95.0, a12 = 97.66 a0) in the trap's radial plane (bec_fill_2d_params),
    64 x 64, complex128: ground state of |1> to residual 1e-8 (renormalised gradient flow), pi/2
    pulse, 2000 steps of 10 us, every step sampled, one trajectory without noise -- on the GPU
    (one workgroup, state in LDS) and the FP64 oracle on the host cores; plus the same evolution
    for 256 trajectories with vacuum noise on the GPU (one workgroup each)."""
    model = B.default_model(N=30000, nvx=64, nvy=64, nvz=2, ne=1, Va=0.0, Vb=0.0, detuning=0.0, dt_evo=1e-5,
                            dt_steady=1e-5)
    p = B.params2d_from_model(model, n_meas=2000)
    B.run_2d(B.params2d_from_model(model, n_meas=10))        # first-call setup
    t0 = time.perf_counter()
    one = B.run_2d(p)
    t_gpu = time.perf_counter() - t0
    pe = B.params2d_from_model(model, n_meas=2000, ne=256, Va=1.0, Vb=1.0)
    t0 = time.perf_counter()
    ens = B.run_2d(pe, want_avn=False)
    t_ens = time.perf_counter() - t0
    r = one["report"]
    out = {"workload": "config 0: Rb-87 30k atoms in the 2D program's units, 64x64 complex128, ground state to "
                       "residual 1e-8, pi/2 pulse, 2000 sampled steps of 10 us; 1 trajectory (and 256 with "
                       "vacuum noise)",
           "gpu_ms_1traj": t_gpu * 1e3, "gpu_device_ms_1traj": r["elapsed_ms"],
           "ground_state": {"iterations": r["gs_iterations"], "residual": r["gs_residual"], "N0": r["N0"],
                            "mu": r["gs_mu"]},
           "visibility_min": float(one["vis"].min()), "visibility_end": float(one["vis"][-1]),
           "gpu_ms_256traj": t_ens * 1e3, "traj_steps_per_s_256": 256 * 2000 / t_ens,
           "N1_N2_end_1traj": [float(one["m"][0, -1]), float(one["m"][1, -1])],
           "N1_N2_end_256traj": [float(ens["m"][0, -1]), float(ens["m"][1, -1])]}
    if with_cpu:
        from oracle import oracle as O
        threads = host_threads()
        O.set_threads(threads)
        op = O.Params2D(**{f: getattr(p, f) for f, _ in O.Params2D._fields_})
        t0 = time.perf_counter()
        ref = O.run2d(op)
        t_cpu = time.perf_counter() - t0
        out.update(cpu_ms_1traj=t_cpu * 1e3, cpu_threads=threads, cpu_kind="port",
                   max_rel_diff_N=float(np.max(np.abs(one["m"] - ref["m"]) / np.abs(ref["m"]).max())),
                   max_abs_diff_visibility=float(np.max(np.abs(one["vis"] - ref["vis"]))))
    return out


def program_1d(B, with_cpu):
    """The 1D program (programs/bec_wigner_2d.cpp/wigner_1d.cpp) as it stands: 41 cells, 100
    ground-state steps, 3000 equilibration steps and 3001 sampled steps for 20 trajectories, on the
    GPU (one wavefront per trajectory) and, for comparison, the FP64 oracle on one host core; plus
    the same run for 4096 trajectories on the GPU."""
    B.run_1d(B.params1d(ntW=10, ne=2))                       # first-call setup
    t0 = time.perf_counter()
    one = B.run_1d(B.params1d())
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    big = B.run_1d(B.params1d(ne=4096))
    t_big = time.perf_counter() - t0
    steps = 2 * 3001 - 1
    out = {"workload": "41 cells complex128, 100 ground-state steps, 3000 + 3001 steps; 20 trajectories "
                       "(and 4096)",
           "gpu_ms_20traj": t_gpu * 1e3, "gpu_device_ms_20traj": one["report"]["elapsed_ms"],
           "gpu_ms_4096traj": t_big * 1e3, "traj_steps_per_s_4096": 4096 * steps / t_big,
           "N1_N2_last_sample": [float(one["m"][0, -1]), float(one["m"][1, -1])]}
    if with_cpu:
        from oracle import oracle as O
        t0 = time.perf_counter()
        ref = O.run1d(O.params1d())
        t_cpu = time.perf_counter() - t0
        out.update(cpu_ms_20traj=t_cpu * 1e3, cpu_threads=1, cpu_kind="port",
                   max_rel_diff_N1_plus_N2=float(np.max(np.abs(one["m"].sum(0) - ref["m"].sum(0)))
                                                 / ref["m"].sum(0)[0]))
    return out


def loss_noise_cost(B, psi, traj, steps):
    """The headline workload with the loss-induced Wigner noise switched on against the same run without
    it (same context, schedule and graphs), plus the xy-pass time at 2048 trajectories with the noise."""
    c = B.derive(B.default_model(**config1(ne=traj)))
    out = {"workload": f"{traj} trajectories x 128x16x16, {steps} steps, graphs on"}
    with B.Solver(c, precision=B.FP32) as s:
        s.init_evolution(psi, SEED, equilibrate=False)
        for noise in (False, True):
            s.set_loss_noise(noise, seed=0x1055)
            s.propagate(200)
            s.synchronize()
            t0 = time.perf_counter()
            s.propagate(steps)
            s.synchronize()
            out["ms_per_step_noise" if noise else "ms_per_step_plain"] = (time.perf_counter() - t0) / steps * 1e3
    cb = B.derive(B.default_model(**config1(ne=2048)))
    with B.Solver(cb, precision=B.FP32) as s:
        s.init_evolution(psi, SEED, equilibrate=False)
        ms = s.time_passes(iters=20)
        s.set_loss_noise(True, seed=0x1055)
        msn = s.time_passes(iters=20)
    out["xy_pass_ms_2048traj"] = {"plain": ms["xy_pass"], "noise": msn["xy_pass"]}
    out["noise_over_plain"] = out["ms_per_step_noise"] / out["ms_per_step_plain"]
    return out


def two_component_ground_state(B):
    """Two-component ground state (N/2 per component) at the headline parameters, gradient flow
    from the Thomas-Fermi start to residual 1e-6, FP64 (second solve timed)."""
    c = B.derive(B.default_model(**config1(ne=1)))
    for _ in range(2):
        out, reps, st = B.ground_states_two_component([c], tol=1e-6, max_iter=20000, precision=B.FP64)
    r = reps[0]
    return {"time_to_residual_1e-6_ms": r["elapsed_ms"], "iterations": r["iterations"], "residual": r["residual"],
            "converged": bool(st == 0), "start": "Thomas-Fermi",
            "method": "preconditioned imaginary-time gradient flow, FP64, coupled components"}


def gemm_rate():
    """The surrogate network's GEMM (bec_gemm: BF16 operands, FP32 accumulation,
    v_mfma_f32_16x16x32_bf16) on a square problem and on the network's per-cell layer shape,
    event-timed; TFLOP/s against MI355X's dense BF16 peak (2.5 PFLOP/s)."""
    from bec_amd import fno as F
    rng = np.random.default_rng(7)
    out = {"peak_TFLOPs": BF16_PEAK / 1e12, "kernel": "gemm_kernel: 64x64 tiles, 4 waves, K tile 32, LDS-staged"}
    cases = (("square_4096_bf16", 4096, 4096, 4096, True),
             ("per_cell_layer_b16", 16 * 32768, 64, 64, False))
    for name, M, N, K, bf in cases:
        A = rng.normal(size=(M, K)).astype(np.float32)
        Bt = rng.normal(size=(N, K)).astype(np.float32)       # B[k][n] at n * K + k
        Ah, Bh = (F.to_bf16_bits(A), F.to_bf16_bits(Bt)) if bf else (A, F.to_bf16_bits(Bt))
        dA = F.DeviceArray(Ah.shape, Ah.dtype, Ah)
        dB = F.DeviceArray(Bh.shape, Bh.dtype, Bh)
        dC = F.DeviceArray((M, N), np.float32)
        ms = F.gemm(dA, dB, dC, M, N, K, F.m1(K), F.m1(1), F.m1(1), F.m1(K), F.m1(N), F.m1(1), time_iters=20)
        tf = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        out[name] = {"M": M, "N": N, "K": K, "A": "bf16" if bf else "fp32 (rounded to bf16 on load)",
                     "B": "bf16", "ms": ms, "TFLOPs": tf, "frac_of_peak": tf * 1e12 / BF16_PEAK}
        for arr in (dA, dB, dC):
            arr.free()
    return out


def surrogate_summary():
    """The committed results of `python -m bec_amd.surrogate` (not measured by this run): the
    Thomas-Fermi-started solve against the guess-started one, for each flow setting reported."""
    keep = ("held_out", "converged_all", "tf_iterations_mean", "nn_iterations_mean", "tf_ms_mean",
            "nn_ms_mean_incl_inference", "speedup_mean_time", "speedup_iterations")
    cut = lambda s: {k: s[k] for k in keep if k in s}
    out = {"measured_by_this_run": False,
           "note": "copied from the committed summaries below; training the network takes minutes, so the "
                   "benchmark does not re-run it",
           "sources": [], "commands": []}
    for path in SURROGATE_SUMMARIES:
        if not os.path.exists(path):
            continue
        with open(path) as f:
            d = json.load(f)
        out["sources"].append(os.path.relpath(path, ROOT))
        out["commands"].append(d.get("command"))
        nn = d["config"].get("nn_flow")
        key = ("same_flow_both_starts" if nn == "default"
               else f"network_start_flow_{nn['tau']}:{nn['beta']}")
        out.setdefault(key, cut(d["summary"]))
        for spec, s in (d.get("nn_sweep") or {}).items():
            out.setdefault(f"network_start_flow_{spec}", cut(s))
        for k, s in d.items():
            if k.startswith("both_starts_"):
                out.setdefault(k, cut(s))
        if "batched" in d:
            out.setdefault("batched_all_sets", d["batched"])
    return out if out["sources"] else None


def host_threads():
    """The host cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS
    where the environment sets the process's share of a shared host."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def cpu_baseline(psi, seconds):
    """FP64 oracle split-step (same algorithm and seeds) timed on this host's cores: OpenMP over
    trajectories, at least one trajectory per thread."""
    from oracle import oracle as O
    threads = host_threads()
    O.set_threads(threads)
    ne = max(16, threads)
    c = O.derive(O.model(**config1(ne=ne)))
    a, b = O.init_evolution(c, psi, SEED, equilibrate=True)
    t0 = time.perf_counter()
    a, b = O.propagate(c, a, b, c.dt_evo, 2)
    est = (time.perf_counter() - t0) / 2
    steps = max(2, int(seconds / max(est, 1e-6)))
    t0 = time.perf_counter()
    a, b = O.propagate(c, a, b, c.dt_evo, steps)
    el = time.perf_counter() - t0
    return {"value": ne * steps / el, "unit": "trajectory-steps/s", "cores": threads, "kind": "port",
            "sample": f"{ne} trajectories x {steps} steps at 128x16x16 (FP64 oracle, OpenMP {threads} threads, "
                      f"{cpu_model()}), {el:.1f} s",
            "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else None, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": cpu_model(),
            "precision": "FP64 (the oracle has no FP32 build)"}


def graph_use(a):
    """What the timed call replays: a segment of up to 64 steps between samples is one
    captured graph, a longer one runs 32-step block graphs plus single launches."""
    if a.no_graphs:
        return "none: one launch per pass"
    seg = min(a.sample_every, a.steps) if a.sample_every > 0 else a.steps
    if seg <= 64:
        return f"one graph per {seg}-step segment"
    return f"32-step block graphs inside {seg}-step segments"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--traj", type=int, default=64)
    ap.add_argument("--sample-every", type=int, default=250)
    ap.add_argument("--roofline-traj", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-groups comparison")
    ap.add_argument("--uneven-steps", type=int, default=400)
    ap.add_argument("--uneven-only", action="store_true",
                    help="run only the uneven-groups comparison and print its JSON (for profiling)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}): trajectories are independent and "
              "would shard across GPUs with one reduction of the FP64 sample sums, which this build does not run",
              file=sys.stderr)
        sys.exit(2)

    import bec_amd as B
    if B.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    if a.uneven_only:
        c1 = B.derive(B.default_model(**config1(ne=1)))
        with B.Solver(c1, precision=B.FP64) as g:
            psi, _, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-6,
                                                 precision=B.FP64, max_iter=5000)
        print(json.dumps(uneven_groups(B, psi, a.uneven_steps, autotune=not a.no_autotune)))
        return

    # ground state on the GPU: plain reference solve and residual-targeted solve
    c1 = B.derive(B.default_model(**config1(ne=1)))
    with B.Solver(c1, precision=B.FP64) as g:
        psi, rep_grad, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-6,
                                                    precision=B.FP64, max_iter=5000)
        psi_ref, rep_ref, _ = g.calculate_steady_state(method=B.GS_SPLIT_STEP, rule=B.GS_RULE_REFERENCE,
                                                       tol=1e-6, precision=B.FP64)
        # second run of each: timings without first-call setup
        psi, rep_grad, _ = g.calculate_steady_state(method=B.GS_GRADIENT, rule=B.GS_RULE_RESIDUAL, tol=1e-6,
                                                    precision=B.FP64, max_iter=5000)

    log("ground states done")
    # headline workload
    c = B.derive(B.default_model(**config1(ne=a.traj)))
    s = B.Solver(c, precision=B.FP32)
    s.set_graphs(not a.no_graphs)
    s.init_evolution(psi, SEED, equilibrate=True)
    # 200-step trials (5 ms each): the leading schedules here are about 1% apart
    sched = s.autotune(steps=200)[0] if not a.no_autotune else s.schedule()
    s.prepare(a.steps, sample_every=a.sample_every)   # capture the timed call's step graphs, untimed
    # the warmup runs last, so that the timed call follows work on the GPU rather than host-side capture
    s.propagate(a.warmup, sample_every=a.sample_every)
    s.synchronize()
    t0 = time.perf_counter()
    samples = s.propagate(a.steps, sample_every=a.sample_every)
    s.synchronize()
    el = time.perf_counter() - t0
    obs = s.observables()
    value = a.traj * a.steps / el
    ms_cfg = s.time_passes(iters=200)
    s.close()

    log(f"headline run done: {value:.4g} trajectory-steps/s")
    # roofline on a state beyond the Infinity Cache
    cb = B.derive(B.default_model(**config1(ne=a.roofline_traj)))
    sb = B.Solver(cb, precision=B.FP32)
    sb.init_evolution(psi, SEED, equilibrate=False)
    # 40-step trials (28 ms each here): 8-step ones could not separate schedules 0.7% apart
    sched_big = sb.autotune(steps=40)[0] if not a.no_autotune else sb.schedule()
    sb.propagate(64)
    ms_big = sb.time_passes(iters=50)
    sb.synchronize()
    nbig = 1000
    t0 = time.perf_counter()
    sb.propagate(nbig)
    sb.synchronize()
    big_rate = a.roofline_traj * nbig / (time.perf_counter() - t0)
    sb.close()
    copy_ms = B.time_copy(1 << 30, 10)
    per_launch = a.roofline_traj * ALGO_BYTES_PER_TRAJ_STEP
    kern = {"z_pass": ms_big["z_pass"], "xy_pass": ms_big["xy_pass"]}
    dom = max(kern, key=kern.get)
    achieved = per_launch / (kern[dom] * 1e-3)
    traffic = load_traffic()
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": (traffic or {}).get(dom), "kernel": dom,
            "workload": f"{a.roofline_traj} trajectories x 128x16x16 x 2 components, complex64 "
                        f"({per_launch / 2**30:.2f} GiB read + written per launch)",
            "ms_per_launch": kern,
            "other_kernel_GBs": {k: per_launch / (v * 1e-3) / 1e9 for k, v in kern.items()},
            "step_traj_steps_per_s": big_rate,
            "step_algorithmic_frac": big_rate * ALGO_BYTES_PER_TRAJ_STEP / HBM_PEAK,
            "passes_per_step": 2, "schedule": sched_big,
            "copy_kernel_GBs": 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9}

    log("roofline leg done")
    uneven = None if a.no_uneven else uneven_groups(B, psi, a.uneven_steps, autotune=not a.no_autotune)
    log("uneven-groups leg done")
    gs2 = two_component_ground_state(B)
    log("two-component ground state done")
    lnoise = loss_noise_cost(B, psi, a.traj, 2000)
    log("loss-noise leg done")
    gemm = gemm_rate()
    log("GEMM leg done")
    p2d = program_2d(B, not a.no_cpu)
    log("2D program done")
    p1d = program_1d(B, not a.no_cpu)
    log("1D program done")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(psi, a.cpu_seconds)

    out = {
        "metric": "trajectory-steps/s at 128x16x16x2 + HBM GB/s; ground-state time to 1e-6 residual",
        "value": value, "unit": "trajectory-steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (Philox vacuum noise on the GPU-computed ground state)",
        "config": {"workload": f"{a.traj} trajectories x 128x16x16 two-component Rb-87, complex64, N=30000, "
                               f"losses on, Philox seed 0x5EED, dt=40us, FP64 sums sampled every "
                               f"{a.sample_every} steps ({int(samples.shape[0])} samples in the timed "
                               f"{a.steps} steps)",
                   "trajectories": a.traj, "lattice": [128, 16, 16], "sample_every": a.sample_every,
                   "samples_taken": int(samples.shape[0]),
                   "graphs_used": graph_use(a),
                   "schedule": sched},
        "roofline": roof,
        "cpu_baseline": cpu,
        "hbm_GBs_config_size": {k: a.traj * ALGO_BYTES_PER_TRAJ_STEP / (v * 1e-3) / 1e9
                                for k, v in ms_cfg.items() if k != "sample_pass"},
        "ground_state": {
            "time_to_residual_1e-6_ms": rep_grad["elapsed_ms"], "iterations": rep_grad["iterations"],
            "residual": rep_grad["residual"], "method": "preconditioned imaginary-time gradient flow, FP64",
            "reference_rule_ms": rep_ref["elapsed_ms"], "reference_rule_iterations": rep_ref["iterations"],
            "reference_rule_residual": rep_ref["residual"]},
        "ground_state_two_component": gs2,
        "loss_noise": lnoise,
        "surrogate_start": surrogate_summary(),
        "surrogate_gemm": gemm,
        "uneven_groups": uneven,
        "program_2d": p2d,
        "program_1d": p1d,
        "final_observables": {k: obs[k] for k in ("t", "N_total", "visibility", "visibility_grid")},
        "n_samples": int(samples.shape[0]),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
