#!/usr/bin/env python3
"""Benchmark: drop-impact time steps per second at 1024x512, every projection
(MAC, FEM, implicit viscosity u and v) converged to relative residual 1e-10.

One step = fluid_2d::step_forward(dt_reg) on GPU-resident state (gas-layer
update, Godunov passes, MAC projection, boundary conditions, two implicit
viscous solves, FEM projection).  Workload: 32 cSt drop, R=1 mm, V=0.5 m/s,
h_min=5 um, x_sym, implicit viscosity, synthetic (no input data needed).

`value` is the plain FP64 path.  Beside it the line carries the mixed-
precision form, the mixed-precision + learned-guess form (network trained in
the run on pairs the solver records), the batched form (64 simulations of
uneven sizes), the operator-level solves and the CPU baseline (the FP64
oracle, a C restatement of the reference step, on a bounded sample).
Prints one JSON line; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "steps/s at 1024x512 to 1e-10 residual: plain FP64 vs mixed+learned; HBM %"
PEAK_HBM = 8000.0  # GB/s, MI355X


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel(a, b):
    import numpy as np
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def field_errors(vdi, a, b):
    fields = {"u": vdi.U, "v": vdi.V, "p": vdi.P, "gas_pg": vdi.PG, "gas_h": vdi.H, "gas_p": vdi.GP,
              "gas_tau": vdi.GTAU}
    return {name: rel(a.get(aid), b.get(aid)) for name, aid in fields.items()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m, n, seconds, warmup, schedules):
    """Times the oracle on the host (test infrastructure used as the checker
    and as the reported CPU baseline), with the GPU run's V(nu1, nu2) per
    solver: a bounded sample with all the threads this process may use
    (OMP_NUM_THREADS, threads bound close), then one with a single thread."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    from vdropimpact_amd.params import drop_impact_params

    p = drop_impact_params(m, n)

    def sample(nthreads, secs):
        used = oracle_lib.lib().vdo_set_threads(nthreads)
        s = oracle_lib.OracleSim(p, rel_tol=1e-10)
        for k, nu in enumerate(schedules):
            s.set_schedule(k, *nu)
        s.init_fields()
        for _ in range(warmup):
            s.step_forward(p.dt_reg)
        k = 0
        t0 = time.perf_counter()
        while True:
            s.step_forward(p.dt_reg)
            k += 1
            el = time.perf_counter() - t0
            if el >= secs and k >= 2:
                break
        s.close()
        return k / el, k, el, used

    rate, k, el, used = sample(threads, seconds)
    rate1, k1, el1, _ = sample(1, seconds / 2)
    return {"value": rate, "unit": "steps/s", "cores": used, "kind": "port",
            "steps_per_s_1_thread": rate1,
            "host": {"nproc": os.cpu_count(), "cpu_model": cpu_model(),
                     "OMP_NUM_THREADS": used, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")},
            "sample": f"steps {warmup + 1}..{warmup + k} of the {m}x{n} workload ({k} steps, {el:.1f} s) with "
                      f"{used} threads bound close, and steps {warmup + 1}..{warmup + k1} ({el1:.1f} s) with 1 "
                      f"thread; every solve to 1e-10 with V(nu1,nu2) per solver {schedules}"}


def operator_level(vdi, reps=3):
    """Operator-level solves (projection and Helmholtz systems alone): the
    64-mode Fourier inputs, every solve from zero to relative residual 1e-10,
    FP64 at 1024x1024, 2048x1024 and 4096x2048, mixed FP32/FP64 at 2048x1024
    and 4096x2048.  GB/s = V-cycles x algorithmic bytes per V-cycle / time;
    level0_down / level0_up: the level-0 launches alone, algorithmic bytes
    per launch over their device-clock duration (first workgroup start to
    last workgroup end).
    Then the advection passes alone on the same field at 1024x1024 and
    2048x1024 (GB/s = algorithmic bytes per launch / launch time)."""
    import numpy as np
    from vdropimpact_amd import workloads
    from vdropimpact_amd.params import VdiParams
    out = []
    cases = [(form, m, n) for form, grids in ((vdi.FORM_FP64, [(1024, 1024), (2048, 1024), (4096, 2048)]),
                                              (vdi.FORM_MIXED, [(2048, 1024), (4096, 2048)])) for m, n in grids]
    names = {0: "mac", 1: "fem", 2: "helmholtz"}
    for form, m, n in cases:
        for kind in (0, 1, 2):
            log(f"operator level: {names[kind]} {m}x{n} {'fp64' if form == 0 else 'mixed'}")
            M, N, cx, cy, b = workloads.operator_case(kind, m, n)
            g = vdi.MGSystem(kind, M, N, cx, cy, 1)
            g.set_form(form)
            g.set_b(b)
            zero = np.zeros(M * N)
            g.set_z(zero)
            g.autotune(3)
            times, cyc = [], 0
            for _ in range(reps):
                g.set_z(zero)
                g.synchronize()
                t0 = time.perf_counter()
                cyc = g.solve(1e-10, 100)
                times.append(time.perf_counter() - t0)
            t = min(times)
            gbs = cyc * g.cycle_bytes() / t / 1e9
            rec = {"system": names[kind], "grid": [m, n], "form": "fp64" if form == 0 else "mixed",
                   "vcycles": cyc, "ms_per_solve": 1e3 * t, "GB_per_s": gbs, "frac_of_8TBps": gbs / PEAK_HBM}
            # the level-0 launches alone: device-clock stamps over 5 V-cycles
            for cls, key in ((vdi.KCLASS_MG_DOWN0, "level0_down"), (vdi.KCLASS_MG_UP0, "level0_up")):
                g.time_kernels(cls, 64)
                g.set_z(zero)
                g.vcycles(5)
                nl, ms, by = g.kernel_times()
                g.time_kernels(-1, 0)
                if nl:
                    lg = by / (ms * 1e-3) / 1e9
                    rec[key] = {"us_per_launch": 1e3 * ms / nl, "alg_MB_per_launch": by / nl / 1e6,
                                "GB_per_s": lg, "frac_of_8TBps": lg / PEAK_HBM}
            out.append(rec)
            g.close()
    for m, n in ((1024, 1024), (2048, 1024)):
        p = VdiParams(m=m, n=n, x_sym=1, implicit_visc=1, ax=0.0, bx=1.0, ay=0.0, by=n / m, rho=1.0, mu=1e-3,
                      dt_reg=0.2 / m, alpha=1 / 1.4, mug=1.8e-5, Pamb=1e5, h0=5e-6, R=1e-3, V=0.5, sigma=0.02)
        sim = vdi.Fluid2D(p)
        u, v, pn, z = workloads.advection_state(m, n)
        for aid, arr in ((vdi.U, u), (vdi.V, v), (vdi.P, pn)):
            sim.set(aid, arr)
        sim.set(vdi.Z_MAC, z.ravel())
        # reads u, v, p; writes ul, vl, ud, vd, c0, c1, src: 10 words a cell
        # reads z, ul, vl, ud, vd, u, v, c0, c1; writes c0, c1: 11 words a cell
        for which, name, words in ((vdi.PASS_GODUNOV, "godunov K1-K6", 10),
                                   (vdi.PASS_CORRECT_ADVECT, "correct+advect K7-K8", 11)):
            sim.run_pass(which, p.dt_reg, 3)
            ms = sim.run_pass(which, p.dt_reg, 50)
            gbs = words * 8.0 * m * n / (ms * 1e-3) / 1e9
            out.append({"system": name, "grid": [m, n], "form": "fp64", "ms_per_launch": ms, "GB_per_s": gbs,
                        "frac_of_8TBps": gbs / PEAK_HBM})
        sim.close()
    return out


def measured_traffic(kernel_class):
    """HBM bytes per launch of the dominant kernel function (or class) from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE per dispatch, each
    scaled by the factor the same counter reads for a plain copy kernel of
    known size), or None when no summary for this class is present."""
    path = os.path.join(REPO, "profiles", "pmc_traffic_1024x512.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    k = d.get("kernels", {}).get(kernel_class) or d.get("classes", {}).get(kernel_class)
    return None if k is None else k.get("hbm_bytes_per_launch")


def copy_probe(vdi, n, reps):
    ms = vdi.copy_probe(n, reps)
    return {"doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n, "ms_per_copy": ms,
            "GB_per_s": 16 * n / (ms * 1e-3) / 1e9}


def time_steps(sim, dt, steps):
    sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step_forward(dt)
    sim.synchronize()
    return time.perf_counter() - t0


def run_plain(vdi, args, p):
    sim = vdi.Fluid2D(p, rel_tol=1e-10)
    if args.no_graphs:
        sim.set_graphs(False)
    if args.serial_visco:
        sim.set_concurrency(False)
    sim.init_fields()
    schedules = [list(sim.schedule(k)) for k in (vdi.MAC, vdi.FEM, vdi.VX, vdi.VY)]
    nstep = 0

    def steps(k):
        nonlocal nstep
        for _ in range(k):
            sim.step_forward(p.dt_reg)
        nstep += k

    steps(args.warmup)
    if args.autotune > 0:
        sim.autotune(args.autotune)
        steps(3)
    log(f"plain: {args.warmup} warm-up steps, tiles tuned, schedules {schedules}")
    # the FP64 level launches the tuner chose, per solver and level (tile or
    # streaming launch)
    tiles = {}
    for k, name in ((vdi.MAC, "mac"), (vdi.FEM, "fem"), (vdi.VX, "vx"), (vdi.VY, "vy")):
        idx = sim.tile_indices(k)
        lc = (len(idx) - 1) // 4
        tiles[name] = {"down": [vdi.tile_label(i) for i in idx[:lc]], "up": [vdi.tile_label(i) for i in idx[lc:2 * lc]]}
    # Which kernel function dominates the step: every launch of one function
    # (all levels) is timed by the kernels' own wall-clock stamps over a few
    # steps; the level classes are surveyed too.
    class_ms, class_n = {}, {}
    for kc in list(vdi.KFUNC_NAMES) + [vdi.KCLASS_MG_DOWN0, vdi.KCLASS_MG_UP0, vdi.KCLASS_MG_MID]:
        sim.time_kernels(kc, 8192)
        steps(3)
        n_, ms, _b = sim.kernel_times()
        class_ms[kc] = ms / 3.0
        class_n[kc] = n_ / 3.0
    funcs = {k: v for k, v in class_ms.items() if k in vdi.KFUNC_NAMES}
    dominant = max(funcs, key=funcs.get)
    names = {**vdi.KCLASS_NAMES, **vdi.KFUNC_NAMES}
    log("kernel functions / classes, ms/step:", {names[k]: round(v, 4) for k, v in class_ms.items()})
    log("kernel functions / classes, launches/step:", {names[k]: round(v, 2) for k, v in class_n.items()})
    # Timed region with the dominant function stamped (a region of its own:
    # the stamps' atomics slow the step).
    sim.time_kernels(dominant, 16384)
    steps(3)   # step graphs captured with the stamps
    sim.reset_stats()
    el = time_steps(sim, p.dt_reg, args.steps)
    nstep += args.steps
    launches, kms, kbytes = sim.kernel_times()
    sim.time_kernels(-1)
    steps(3)
    # The clean timed region: `value`.
    sim.reset_stats()
    g0, r0 = sim.counters()
    b0 = sim.budgets()
    first = nstep + 1
    el_clean = time_steps(sim, p.dt_reg, args.steps)
    nstep += args.steps
    g1, r1 = sim.counters()
    st = sim.stats()
    cyc = {name: st.total_cycles[k] / max(1, st.solves[k]) for k, name in enumerate(["mac", "fem", "vx", "vy"])}
    height_passes = st.last_height_passes
    budgets = sim.budgets()
    log(f"plain: {args.steps} steps in {el:.3f} s (stamped), {el_clean:.3f} s (no stamps), "
        f"simulation steps {first}..{nstep}, V-cycles per solve {cyc}")
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    sim.close()
    traffic = measured_traffic(names[dominant])
    return {
        # the stamps (two atomics per workgroup of the stamped class) slow the
        # step; the value is the same timed region without them
        "value": args.steps / el_clean, "ms_per_step": 1e3 * el_clean / args.steps,
        "steps_per_s_stamped": args.steps / el,
        "timed_simulation_steps": [first, nstep],
        "avg_vcycles": cyc,
        "captured_steps": g1 - g0, "redone_steps": r1 - r0,
        "height_passes_last_step": height_passes,
        "budgets_mac_vx_vy_fem": {"before": b0, "after": budgets},
        "schedules": schedules,
        "tiles_fp64": tiles,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "kernel": names[dominant], "launches": launches,
                     "avg_launch_us": 1e3 * kms / max(1, launches),
                     "alg_bytes_per_launch": kbytes / max(1, launches),
                     "duration": "first workgroup start to last workgroup end (device wall clock), "
                                 "launches that returned at once on a converged flag left out",
                     "ms_per_step": {names[k]: v for k, v in class_ms.items()}},
    }


def run_forms(vdi, args, p):
    """Mixed precision, and mixed precision + learned guess, each against a
    plain FP64 simulation run in lockstep from the same initial state (field
    errors after the same number of steps)."""
    ref = vdi.Fluid2D(p, rel_tol=1e-10)
    mix = vdi.Fluid2D(p, rel_tol=1e-10)
    lrn = vdi.Fluid2D(p, rel_tol=1e-10)
    for s in (ref, mix, lrn):
        s.init_fields()
    mix.set_form(vdi.FORM_MIXED)
    lrn.set_form(vdi.FORM_MIXED)
    lrn.guess_attach(capacity=args.pairs, seed=5)
    lrn.guess_mode(1)
    for _ in range(args.pairs + 2):
        for s in (ref, mix, lrn):
            s.step_forward(p.dt_reg)
    t0 = time.perf_counter()
    loss = lrn.guess_train(epochs=args.epochs, batch=4, lr=1e-3)
    t_train = time.perf_counter() - t0
    gst = lrn.guess_stats()
    log(f"learned guess: {gst['pairs']} pairs, {args.epochs} epochs in {t_train:.1f} s, loss {loss:.4g}")
    lrn.guess_mode(2)
    if args.autotune > 0:
        for s in (ref, mix, lrn):
            s.autotune(args.autotune)
    for s in (ref, mix, lrn):
        for _ in range(3):
            s.step_forward(p.dt_reg)
        s.reset_stats()
    g0 = lrn.guess_stats()
    res = {}
    for name, s in (("fp64", ref), ("mixed", mix), ("mixed_learned", lrn)):
        el = time_steps(s, p.dt_reg, args.steps)
        st = s.stats()
        res[name] = {"steps_per_s": args.steps / el,
                     "avg_vcycles": {n_: st.total_cycles[k] / max(1, st.solves[k])
                                     for k, n_ in enumerate(["mac", "fem", "vx", "vy"])}}
        log(f"{name}: {args.steps / el:.1f} steps/s")
    g1 = lrn.guess_stats()
    out = {
        "mixed": dict(res["mixed"], field_error_vs_fp64=field_errors(vdi, mix, ref)),
        "mixed_learned": dict(res["mixed_learned"], field_error_vs_fp64=field_errors(vdi, lrn, ref),
                              training={"pairs": gst["pairs"], "epochs": args.epochs, "seconds": t_train,
                                        "final_loss": loss, "grid": [p.m, p.n]},
                              guess_accepted=g1["accepted"] - g0["accepted"],
                              guess_rejected=g1["rejected"] - g0["rejected"]),
        "fp64_lockstep": res["fp64"],
    }
    base = res["fp64"]["steps_per_s"]
    out["speedup_mixed_vs_fp64"] = res["mixed"]["steps_per_s"] / base
    out["speedup_mixed_learned_vs_fp64"] = res["mixed_learned"]["steps_per_s"] / base
    out["mac_vcycles_saved_per_step"] = (res["mixed"]["avg_vcycles"]["mac"] -
                                         res["mixed_learned"]["avg_vcycles"]["mac"])
    for s in (ref, mix, lrn):
        s.close()
    return out


def run_batch(vdi, args):
    import numpy as np
    from vdropimpact_amd import workloads
    cfgs, prms = workloads.batch_members(args.batch_sims)
    b = vdi.Batch(prms, vdi.FORM_MIXED)
    b.init_fields()
    for _ in range(4):
        b.step()
    b.autotune(2)
    for _ in range(3):
        b.step()
    # Each of the 64 slots takes every batched step.  A member whose gas film
    # closes ends its run with the gas layer's Newton error, as the
    # reference's run of that simulation would; its slot then starts the same
    # simulation over (a sweep slot), and the step it failed is not counted.
    spans, graph, ends = [], [], []
    steps_done = 0
    t0 = time.perf_counter()
    for k in range(args.batch_steps):
        ended = []
        try:
            b.step()
        except vdi.VdiError as e:
            if e.status != vdi.ERR_NEWTON_NOCONV:
                raise
            ended = [int(s) for s in np.nonzero(b.member_status() != vdi.OK)[0]]
        g, _busy = b.step_times()
        graph.append(g)
        sp = b.member_spans()
        spans.extend(float(sp[s]) for s in range(len(prms)) if s not in ended)
        steps_done += len(prms) - len(ended)
        for s in ended:
            ends.append({"member": s, "grid": f"{prms[s].m}x{prms[s].n}", "timed_step": k})
            b.restart(s)
    el = time.perf_counter() - t0
    spans = np.array(spans)
    _nb, redo = b.counters()
    sizes = [f"{p.m}x{p.n}" for p in prms]
    out = {"simulations": len(prms), "steps_each": args.batch_steps, "form": "mixed",
           "aggregate_steps_per_s": steps_done / el,
           "member_steps_completed": steps_done,
           "runs_ended_and_restarted": ends,
           "batched_step_ms_p50": float(np.percentile(graph, 50)),
           "batched_step_ms_p99": float(np.percentile(graph, 99)),
           # per simulation and step: from the start of its first work item to
           # the end of its last one in the batched step (device clock)
           "sim_step_ms_p50": float(np.percentile(spans, 50)),
           "sim_step_ms_p99": float(np.percentile(spans, 99)),
           "grids": {s: sizes.count(s) for s in sorted(set(sizes))},
           "members_redone": redo}
    log(f"batch: {out['aggregate_steps_per_s']:.0f} member-steps/s")
    b.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--m", type=int, default=1024)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-operator", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the mixed and mixed+learned forms")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serial-visco", action="store_true",
                    help="captured steps run the two viscous solves one after the other")
    ap.add_argument("--no-graphs", action="store_true",
                    help="run every step on the direct path (for profilers that cannot follow graph launches)")
    ap.add_argument("--autotune", type=int, default=3, help="autotuner repetitions (0: default tile shapes)")
    ap.add_argument("--pairs", type=int, default=64, help="training pairs recorded for the learned guess")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch-sims", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many doubles (counter calibration) and exit")
    ap.add_argument("--operator-only", action="store_true",
                    help="only run the operator-level solves and print their records (profiling) and exit")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this solver runs one simulation on one GPU; "
              "multi-GPU runs are not part of this build (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import vdropimpact_amd as vdi
    from vdropimpact_amd.params import drop_impact_params

    if args.copy_probe > 0:
        print(json.dumps(copy_probe(vdi, args.copy_probe, 10)), flush=True)
        return
    if args.operator_only:
        print(json.dumps({"operator_level": operator_level(vdi)}), flush=True)
        return
    p = drop_impact_params(args.m, args.n)
    plain = run_plain(vdi, args, p)
    out = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"single drop {args.m}x{args.n}, 32 cSt, V=0.5 m/s, R=1 mm, h_min=5 um, "
                               "x_sym, implicit viscosity, all solves to rel. residual 1e-10",
                   "grid": [args.m, args.n], "form": "fp64", "rel_tol": 1e-10,
                   "timed_simulation_steps": plain["timed_simulation_steps"],
                   "avg_vcycles": plain["avg_vcycles"],
                   "captured_steps": plain["captured_steps"], "redone_steps": plain["redone_steps"],
                   "gas_height_passes_last_step": plain["height_passes_last_step"],
                   "vcycle_schedule_mac_fem_vx_vy": plain["schedules"],
                   "level_launches_fp64": plain["tiles_fp64"],
                   "vcycle_budgets_mac_vx_vy_fem": plain["budgets_mac_vx_vy_fem"],
                   "concurrent_branches": not args.serial_visco,
                   "steps_per_s_with_kernel_stamps": plain["steps_per_s_stamped"]},
        "roofline": plain["roofline"],
        "cpu_baseline": None,
    }
    if not args.no_forms:
        log("mixed and mixed+learned forms ...")
        out["forms"] = run_forms(vdi, args, p)
    if not args.no_batch:
        log("batched form ...")
        out["batched"] = run_batch(vdi, args)
    if not args.no_operator:
        log("operator-level solves ...")
        out["operator_level"] = operator_level(vdi)
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.m, args.n, args.cpu_seconds, 3, plain["schedules"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
