#!/usr/bin/env python3
"""Throughput of SHOCK's RHS / RK-stage path on one MI355X.

Workload (the headline configuration): 3D single block 256^3 on a periodic curvilinear grid
x = xi + 0.05 sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta), WENO5 + viscous fluxes, FP64,
classic RK4, Re = 2000, Mach 0.5, decaying turbulence from a seeded random-phase velocity
field (E(k) ~ k^4 exp(-2 (k/4)^2), turbulent Mach 0.5, uniform rho and p), time step from
CFL 0.5 on each step's starting state.  A "step" is one full RK4 time step (4 stages) followed
by checkNAN; checkNAN and the CFL maximum of the new state are fused into the last stage's
update.  The metric is interior cell-updates per second per RK stage with the state resident in
HBM.

Prints one JSON line: the FP64 form, a `mixed` object with the mixed FP32/FP64 form on the same
workload, the CPU baseline (the C oracle with OpenMP on a bounded sample of the same workload,
on the host's CPU share) and `mms2d_64`, the 2D 64^2 periodic manufactured-field case on the GPU and
on one host core.  Per-kernel durations come from HIP events recorded on the library's stream
during the timed region.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from shock_amd import cases, shockx  # noqa: E402

HBM_PEAK_GBS = 8000.0
CLOCK_HZ = 2.4e9          # MI355X peak engine clock
FP64 = 8


def algo_bytes(name, n, ng, time_order=4, diag=True):
    """Algorithmic HBM bytes of one launch of kernel class `name` for an n^3 block with padded
    extent ng^3 (row padding excluded): each input array element read once, each output element
    written once.  diag: the metric is diagonal (the reference's 3D metric), so a sweep reads one
    metric entry of its direction plus J.  The last viscous sweep of a stage also applies the
    stage update (U_n, Q, Q_sum, U_s+1); the fused zeta kernel (WENO and viscous zeta sweeps plus
    the update) reads the viscous sweep's inputs plus rho and e of the stage input."""
    nint = n ** 3
    npad = ng ** 3
    ghost = npad - nint
    met = 2 if diag else 4                     # metric values read per node by a WENO sweep
    qsum = 2 * 5 * (time_order - 1) / time_order if time_order == 4 else 0.0
    if name == "weno_xi":            # rho,u,v,w,e + metric in, Q (5) out
        return nint * (5 + met + 5) * FP64
    if name in ("weno_eta", "weno_zeta"):   # same + Q read-modify-write
        return nint * (5 + met + 5 + 5) * FP64
    if name == "ghost_values":       # rho,u,v,w,e in (interior), T, mu out (all padded nodes), ghosts filled
        return (nint * 5 + npad * 2 + ghost * 5) * FP64
    if name == "visc_grad":          # u,v,w (+T for the general metric) in (padded), 9 or 12 derivatives out
        return npad * (3 if diag else 4) * FP64 + (9 if diag else 12) * n * ng * ng * FP64
    vin = 5 + (4 + 4 if diag else 8 + 10)       # u,v,w,T,mu + cross derivatives + metric
    if name in ("visc_xi", "visc_eta"):          # + Q (momentum, energy) read-modify-write
        return nint * (vin + 8) * FP64
    if name == "visc_zeta":                      # + Q (5) read, U_n (5) read, Q_sum, U_s+1 (5) written
        return nint * (vin + 5 + 5 + qsum + 5) * FP64
    if name == "weno_visc_zeta":                 # fused zeta sweeps: the viscous inputs + rho, e
        return nint * (vin + 2 + 5 + 5 + qsum + 5) * FP64
    if name == "weno_visc_eta":                  # fused eta sweeps: the viscous inputs + rho, e, Q (5) r/w
        return nint * (vin + 2 + 5 + 5) * FP64
    if name == "visc_grad_xi":                   # the march with viscous xi: u,v,w (padded); T, mu,
        # the diagonal metric and J, Q (momentum, energy) r/w per interior node; 7 derivatives out
        return (npad * 3 + nint * (2 + 4 + 8)) * FP64 + 7 * n * ng * ng * FP64
    if name == "rk_update":          # U_n, Q, J in (Qsum r/w on RK4 stages), U_s+1 out
        return nint * (5 + 5 + 1 + 5 + qsum) * FP64
    if name == "ghost_periodic":     # 5 fields read + written on ghost nodes
        return ghost * 5 * 2 * FP64
    return None


def cpu_baseline(n_sample, steps, seed, threads):
    """C oracle (OpenMP, `threads` threads) on an n_sample^3 block of the same workload."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # noqa: E402  (test infrastructure: CPU baseline leg only)
    orc.lib().orc_set_threads(threads)
    nn = (n_sample,) * 3
    o = orc.Oracle(dims=3, S=5, n=nn, time_order=4, mach=0.5, reynolds=2000.0)
    xe, ye, ze = cases.periodic_ext_coords(3, 5, nn, cases.sine_bump_3d(0.05))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    prim = cases.turbulence_3d_state(3, 5, nn, seed, mach=0.5)
    tau = o.cfl_dt(0.5, met, prim)
    un = [a.copy() for a in prim]
    urk = [a.copy() for a in prim]
    scratch = o.zeros(30)
    o.step(tau, met, un, urk, None, scratch)     # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(tau, met, un, urk, None, scratch)
        for a, b in zip(un, urk):
            a[...] = b
    dt = time.perf_counter() - t0
    return n_sample ** 3 * 4 * steps / dt, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="interior points per direction")
    ap.add_argument("--seed", type=int, default=20251202)
    ap.add_argument("--cpu-n", type=int, default=128, help="CPU baseline sample size per direction")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host's CPU share (affinity set, at most 16)")
    ap.add_argument("--cpu1-n", type=int, default=64, help="one-core CPU sample size per direction (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP events")
    ap.add_argument("--cfl", type=float, default=0.5, help="CFL number of the time step")
    ap.add_argument("--fixed-tau", action="store_true",
                    help="time the steps at the first step's tau (the reference's fixed tau) instead of a CFL "
                         "step per step")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-form run of the default workload")
    ap.add_argument("--no-mms2d", action="store_true", help="skip the 2D 64^2 periodic manufactured-field case")
    ap.add_argument("--no-serve", action="store_true", help="skip the short serving run of the default line")
    ap.add_argument("--no-power", action="store_true",
                    help="skip the board power / clock sample (amd-smi beside ~3 s of further steps)")
    ap.add_argument("--workload", choices=("turbulence", "mms-ladder", "batch", "serve", "shock", "compat"),
                    default="turbulence",
                    help="turbulence: the 256^3 throughput workload (default); mms-ladder: the "
                         "manufactured-solution convergence ladder in both arithmetic forms; batch: 96 "
                         "uneven cases per graph-replayed step; serve: Poisson solve requests; shock: "
                         "the 384^3 WENO9 Mach-3 shock / vorticity block with inflow and outflow; compat: "
                         "the default workload through the reference-named host loop (state in host "
                         "memory between steps)")
    ap.add_argument("--cases", type=int, default=96, help="batch: number of cases; serve: requests")
    ap.add_argument("--rate", default="0.5", help="serve: arrival rates (requests per second, comma-separated)")
    ap.add_argument("--shock-frac", type=float, default=0.5, help="serve: fraction of shock requests")
    ap.add_argument("--n-lo", type=int, default=64, help="serve: smallest request (nodes per direction)")
    ap.add_argument("--n-hi", type=int, default=256, help="serve: largest request")
    ap.add_argument("--ladder", default="32,64,128", help="mms-ladder: rungs (nodes per direction)")
    ap.add_argument("--scheme", type=int, choices=(5, 9), default=9, help="mms-ladder: WENO order")
    ap.add_argument("--ladder-forms", choices=("both", "fp64", "mixed"), default="both",
                    help="mms-ladder: the arithmetic forms to march (both: the error ratio and speedup too)")
    ap.add_argument("--t-end", type=float, default=2.0, help="mms-ladder: marching time per rung")
    ap.add_argument("--stretch", type=float, default=0.0, help="mms-ladder: grid stretching amplitude (0: uniform)")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="arithmetic form of the path (mixed: FP32 WENO nonlinear part)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="launch a plain copy kernel of one padded field this many times after setup "
                         "(reference traffic for hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU layout is "
              "a design (one block per GPU, halo exchange per stage) and is not built.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "mms-ladder":
        return mms_ladder(args)
    if args.workload == "batch":
        return batch_workload(args)
    if args.workload == "serve":
        return serve_workload(args)
    if args.workload == "shock":
        return shock_workload(args)
    if args.workload == "compat":
        return compat_workload(args)

    fp64 = turbulence_run(args, "fp64" if args.precision == "fp64" else "mixed")
    out = fp64["line"]
    if args.precision == "fp64" and not args.no_mixed:
        # the mixed FP32/FP64 form on the same workload, with its own roofline object
        mixed = turbulence_run(args, "mixed")
        m = mixed["line"]
        out["mixed"] = {"value": m["value"], "ms_per_step": m["ms_per_step"], "dtype": m["dtype"],
                        "speedup_vs_fp64": round(m["value"] / out["value"], 4),
                        "roofline": m["roofline"], "kernels": m["kernels"],
                        "time_advanced": m["config"]["time_advanced"]}
    if not args.no_cpu_baseline:
        out["host"] = host_info()
        threads = args.cpu_threads or out["host"]["cpu_share"]
        rate, secs = cpu_baseline(args.cpu_n, args.cpu_steps, args.seed, threads)
        out["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "cell-updates/s/RK-stage", "cores": threads, "kind": "port",
            "sample": f"C oracle (OpenMP, {threads} threads: the host's CPU share), {args.cpu_n}^3 block of the "
                      f"same workload (WENO5+viscous, FP64, RK4, fixed tau), {args.cpu_steps} steps = "
                      f"{4 * args.cpu_steps} stages, {secs:.1f}s"}
        if args.cpu1_n:
            rate1, secs1 = cpu_baseline(args.cpu1_n, 1, args.seed, 1)
            out["cpu_baseline_1core"] = {
                "value": round(rate1, 1), "unit": "cell-updates/s/RK-stage", "cores": 1, "kind": "port",
                "sample": f"C oracle, one thread, {args.cpu1_n}^3 block, 1 step = 4 stages, {secs1:.1f}s"}
        if not args.no_mms2d:
            out["mms2d_64"] = mms2d_64_run(args)
    if not args.no_serve:
        out["serve"] = serve_short(args)
    print(json.dumps(out), flush=True)


def serve_short(args):
    """A short serving run beside the throughput line: 10 Poisson requests at 2 per second, 3D
    periodic turbulence or shock cases of 48^3-96^3, FP64; time-to-solution p50/p99 and the
    aggregate rate."""
    from shock_amd import serve
    reqs = serve.make_requests(10, 2.0, args.seed, n_lo=48, n_hi=96, t_lo=0.05, t_hi=0.2, shock_frac=0.5)
    stats, done = serve.Server().run(reqs)
    return {"workload": "10 Poisson requests at 2/s, 3D periodic turbulence or Mach 1.5-3.5 shock cases, "
                        "48^3-96^3 (log-uniform), t_end U(0.05, 0.2), W5/W9, FP64",
            "requests": stats["requests"], "failed": stats["failed"],
            "p50_s": round(stats["p50_s"], 3), "p99_s": round(stats["p99_s"], 3),
            "queue_wait_p50_s": round(stats["queue_wait_p50_s"], 3),
            "cell_updates_per_s": round(stats["cell_updates_per_s"], 1), "wall_s": round(stats["wall_s"], 2)}


def host_info():
    """Host cores as the bench sees them: nproc, the threads this process may run on, the CPU
    model, and the share the CPU baseline uses (the affinity set, at most 16 threads: the GPU
    host grants each GPU a 16-thread share whatever nproc reports)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model, "cpu_share": min(16, aff)}


def mms2d_64_run(args):
    """The CPU-runnable case: 2D 64x64 periodic unit square with the manufactured field,
    WENO5 + viscous, RK4, CFL 0.5, 100 steps; the GPU path and the C oracle on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # noqa: E402  (test infrastructure: CPU baseline leg only)
    n, S, steps = 64, 5, 100
    nn = (n, n, 1)
    xe, ye, ze = cases.periodic_ext_coords(2, S, nn, cases.cartesian_unit(2))
    x, y, _ = cases.interior_coords(2, S, nn, xe, ye, ze)
    prim = cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y), mach=0.5)
    with shockx.Case(dims=2, space_order=S, time_order=4, n=nn, mach=0.5, reynolds=1000.0) as g:
        g.build_metric(xe, ye, ze)
        g.upload_state(prim)
        g.run_cfl(0.5, 5)
        g.synchronize()
        t0 = time.perf_counter()
        t_gpu = g.run_cfl(0.5, steps)
        g.synchronize()
        gpu_s = time.perf_counter() - t0
    # the same case as a batch of one: one captured HIP graph per step at the initial CFL step,
    # checkNAN read once at the end (the small-case path; shx_run reads it and takes a new CFL
    # step after every step)
    with shockx.Case(dims=2, space_order=S, time_order=4, n=nn, mach=0.5, reynolds=1000.0) as g:
        g.build_metric(xe, ye, ze)
        g.upload_state(prim)
        tau0 = g.cfl_dt(0.5)
        with shockx.Batch([g], [tau0]) as b:
            b.run(5)
            b.synchronize()
            t0 = time.perf_counter()
            b.run(steps)
            b.synchronize()
            graph_s = time.perf_counter() - t0
    orc.lib().orc_set_threads(1)
    o = orc.Oracle(dims=2, S=S, n=nn, time_order=4, mach=0.5, reynolds=1000.0)
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    state = [a.copy() for a in prim]
    o.ghost_periodic(state)
    t0 = time.perf_counter()
    t_cpu = 0.0
    for _ in range(steps):
        dt = o.cfl_dt(0.5, met, state)
        state = o.run(dt, met, state, 1)
        t_cpu += dt
    cpu_s = time.perf_counter() - t0
    cells = n * n * 4 * steps
    return {"workload": "2D 64x64 periodic manufactured field, WENO5 + viscous, RK4, CFL 0.5, 100 steps",
            "gpu": {"value": round(cells / gpu_s, 1), "ms_per_step": round(1e3 * gpu_s / steps, 4),
                    "time_advanced": t_gpu},
            "gpu_graph": {"value": round(cells / graph_s, 1), "ms_per_step": round(1e3 * graph_s / steps, 4),
                          "time_step": tau0,
                          "note": "batch of one: a captured HIP graph per step, fixed step, checkNAN read at the end"},
            "cpu_baseline": {"value": round(cells / cpu_s, 1), "unit": "cell-updates/s/RK-stage", "cores": 1,
                             "kind": "port", "sample": f"C oracle, one thread, the whole run ({cpu_s:.2f}s; "
                                                       f"CFL step per step, time {t_cpu:.6g})"}}


def power_sample(run, seconds=3.0):
    """Board power and graphics clock sampled with amd-smi (a child process every ~0.3 s) while
    run(k) keeps the GPU busy with more steps of the same workload, after the timed region;
    None where amd-smi is absent or prints nothing usable."""
    import shutil
    import subprocess
    import threading
    exe = shutil.which("amd-smi") or ("/opt/rocm/bin/amd-smi" if os.path.exists("/opt/rocm/bin/amd-smi") else None)
    if not exe:
        return None
    samples = []
    stop = threading.Event()

    def poll():
        while not stop.is_set():
            try:
                out = subprocess.run([exe, "metric", "-g", "0", "-p", "-c"], capture_output=True, text=True,
                                     timeout=10).stdout
            except (OSError, subprocess.SubprocessError):
                return
            pw = re.search(r"SOCKET_POWER:\s*([\d.]+)", out)
            clks = [float(x) for x in re.findall(r"GFX_\d+:\s*\n\s*CLK:\s*([\d.]+)", out)]
            if pw and clks:
                samples.append((float(pw.group(1)), sum(clks) / len(clks)))
            stop.wait(0.3)

    th = threading.Thread(target=poll, daemon=True)
    th.start()
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < seconds:
        run(10)
        steps += 10
    stop.set()
    th.join(timeout=15)
    if not samples:
        return None
    pw = [a for a, _ in samples]
    ck = [b for _, b in samples]
    return {"tool": "amd-smi metric -p -c", "samples": len(samples), "steps": steps,
            "board_w_mean": round(sum(pw) / len(pw), 1), "board_w_max": max(pw),
            "gfx_mhz_mean": round(sum(ck) / len(ck), 1), "gfx_mhz_min": round(min(ck), 1),
            "board_w_rated": 1400, "gfx_mhz_peak": 2400}


def turbulence_run(args, precision):
    """The 256^3 turbulence workload in one arithmetic form: setup, warm-up, timed CFL steps with
    the fused end-of-step checks; returns the JSON line's fields."""
    n = args.n
    S, time_order = 5, 4
    nn = (n, n, n)
    t_setup = time.perf_counter()
    case = shockx.Case(dims=3, space_order=S, time_order=time_order, n=nn, gamma=1.4, mach=0.5,
                       reynolds=2000.0, prandtl=0.72, T0=300.0, quirks=shockx.QUIRKS_REFERENCE,
                       precision=shockx.MIXED if precision == "mixed" else shockx.FP64)
    xe, ye, ze = cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(0.05))
    case.build_metric(xe, ye, ze)
    del xe, ye, ze
    prim = cases.turbulence_3d_state(3, S, nn, args.seed, mach=0.5)
    case.upload_state(prim)
    del prim
    tau = case.cfl_dt(args.cfl)
    tiles = case.tune(3)
    t_setup = time.perf_counter() - t_setup
    print(f"[bench] {precision} setup {t_setup:.1f}s dt0={tau:.6g} n={n} weno tiles {tiles}", file=sys.stderr,
          flush=True)

    if args.copy_probe:
        case.copy_probe(args.copy_probe)

    def advance(k):
        if args.fixed_tau:
            case.run(tau, k)
            return tau * k
        return case.run_cfl(args.cfl, k)

    advance(args.warmup)
    case.synchronize()
    if not args.no_timing:
        case.set_timing(True)
    case.synchronize()
    t0 = time.perf_counter()
    t_adv = advance(args.steps)
    case.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    kt = case.kernel_times() if not args.no_timing else {}
    print(f"[bench] {precision}: timed {args.steps} steps in {elapsed:.3f}s (checkNAN every step: no "
          f"non-finite density; time advanced {t_adv:.6g})", file=sys.stderr, flush=True)
    power = None
    if precision == "fp64" and not args.no_power:
        if not args.no_timing:
            case.set_timing(False)
        power = power_sample(lambda k: (advance(k), case.synchronize()))
        print(f"[bench] power beside {power['steps'] if power else 0} more steps: {power}", file=sys.stderr, flush=True)

    stages = args.steps * time_order
    value = n ** 3 * stages / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    roofline = None
    valu = None
    kernels = {}
    if kt:
        ng = n + S + 1
        for name, (ms, cnt) in kt.items():
            if cnt == 0:
                continue
            avg = ms / cnt
            b = algo_bytes(name, n, ng, time_order)
            kernels[name] = {"avg_ms": round(avg, 4), "launches": cnt,
                             "share": round(ms / (1e3 * elapsed), 4),
                             "algo_GBs": round(b / (avg * 1e-3) / 1e9, 1) if b else None}
        dom = max(kernels, key=lambda k: kt[k][0])
        avg = kt[dom][0] / kt[dom][1]
        b = algo_bytes(dom, n, ng, time_order)
        ach = b / (avg * 1e-3) / 1e9
        traffic = None
        tfile = os.path.join(ROOT, "profiles", f"traffic_turbulence{n}_{precision}.json")
        if os.path.exists(tfile):
            with open(tfile) as f:
                traffic = json.load(f).get(dom)
        roofline = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "algo_bytes": int(b), "avg_ms": round(avg, 4)}
        # FP64 vector pipe use beside the HBM fraction: FP64 VALU instructions per launch (hardware
        # counters of the same kernels, profiles/sq_counters_*.json) x 4 cycles per wave64 FP64
        # instruction over 1024 SIMDs x 2.4 GHz x the live launch time.
        sfile = os.path.join(ROOT, "profiles", f"sq_counters_turbulence{n}_{precision}.json")
        if os.path.exists(sfile):
            with open(sfile) as f:
                sq = json.load(f)
            for name, kd in kernels.items():
                if name in sq:
                    kd["fp64_pipe_busy"] = round(sq[name]["fp64_valu_instr_per_launch"] * 4.0
                                                 / (1024 * kd["avg_ms"] * 1e-3 * CLOCK_HZ), 3)
            if dom in sq:
                valu = {"kernel": dom, "fp64_pipe_busy": kernels[dom]["fp64_pipe_busy"],
                        "fp64_valu_instr_per_launch": sq[dom]["fp64_valu_instr_per_launch"],
                        "peak_instr_per_s": 1024 * CLOCK_HZ / 4.0}
    case.close()

    # stage-level fraction of the HBM roofline at the compulsory 260 B/node/stage (fused RK4 stage)
    stage_frac = value * 260.0 / (HBM_PEAK_GBS * 1e9)
    line = {
        "metric": "cell-updates/s/RK-stage", "value": round(value, 1), "unit": "cell-updates/s/RK-stage",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if precision == "fp64" else "f64+f32",
        "data": "synthetic (seeded random-phase turbulence)",
        "config": {"workload": f"{n}^3 periodic curvilinear block, WENO5 + viscous, "
                               f"{'FP64' if precision == 'fp64' else 'mixed FP32/FP64'}, RK4, Re=2000",
                   "grid": [n, n, n], "scheme": "WENO5", "viscous": True, "time_order": time_order,
                   "precision": precision, "quirks": "reference",
                   "time_step": (f"fixed tau {tau:.6g}" if args.fixed_tau
                                 else f"CFL {args.cfl} per step (maximum fused into the last stage)"),
                   "checks_per_step": "checkNAN and CFL maximum fused into the last stage's update",
                   "dt_first": tau, "time_advanced": t_adv, "seed": args.seed},
        "roofline": roofline,
        "valu": valu,
        "stage_hbm_frac_260B": round(stage_frac, 4),
        "weno_tile_variants": tiles,
        "kernels": kernels,
    }
    if power is not None:
        line["power"] = power
    return {"line": line}


def batch_cases(count, seed, precision, lo=32 ** 2, hi=160 ** 3):
    """Batch workload: sizes log-uniform in node count from 32^2 (2D) to 160^3 (3D) -- counts up to
    256^2 become 2D squares, larger ones 3D cubes; Mach U(0.2, 3), Reynolds logU(1e2, 1e4), WENO5 or
    WENO9; 2D cases start from the manufactured field of the 2D parity case, 3D cases from the
    random-phase turbulence on the curvilinear box.  lo, hi: the node-count range."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        S = int(rng.choice([5, 9]))
        mach = float(rng.uniform(0.2, 3.0))
        re = float(math.exp(rng.uniform(math.log(1e2), math.log(1e4))))
        if N <= 256 ** 2:
            n = max(S + 1, int(round(math.sqrt(N))))
            nn = (n, n, 1)
            c = shockx.Case(dims=2, space_order=S, time_order=4, n=nn, mach=mach, reynolds=re, precision=precision)
            xe, ye, ze = cases.periodic_ext_coords(2, S, nn, cases.cartesian_unit(2))
            c.build_metric(xe, ye, ze)
            x, y, z = cases.interior_coords(2, S, nn, xe, ye, ze)
            c.upload_state(cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y), mach=mach))
            cells = n * n
        else:
            n = int(round(N ** (1.0 / 3.0)))
            nn = (n, n, n)
            c = shockx.Case(dims=3, space_order=S, time_order=4, n=nn, mach=mach, reynolds=re, precision=precision)
            c.build_metric(*cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(0.05)))
            c.upload_state(cases.turbulence_3d_state(3, S, nn, int(rng.integers(1, 2 ** 31 - 1)), mach=mach))
            cells = n ** 3
        # the convective CFL step (the reference's), capped by the explicit viscous limit: a
        # fine, low-Reynolds, high-Mach draw is otherwise unstable and fails the per-step checkNAN
        h = 1.0 / n
        dims = 2 if N <= 256 ** 2 else 3
        tau = min(c.cfl_dt(0.5), 0.2 * re * h * h / dims)
        out.append((c, cells, tau, (dims, n, S, mach, re)))
    return out


def batch_workload(args):
    """All cases advanced together from one graph replay per step, against the same cases stepped
    one after another through the single-case path: the configured mix (32^2 .. 160^3), and a mix
    of small cases only (16^2 .. 40^3), where launches rather than the largest case set the time."""
    out = batch_mix(args, 32 ** 2, 160 ** 3)
    small = batch_mix(args, 16 ** 2, 40 ** 3)
    out["small_mix"] = {k: small[k] for k in ("value", "ms_per_step", "sequential_value", "batch_speedup")}
    out["small_mix"]["config"] = small["config"]
    print(json.dumps(out), flush=True)


def batch_mix(args, lo, hi):
    prec = shockx.MIXED if args.precision == "mixed" else shockx.FP64
    t0 = time.perf_counter()
    cs = batch_cases(args.cases, args.seed, prec, lo, hi)
    cells = sum(x[1] for x in cs)
    print(f"[batch] {len(cs)} cases, {cells:.3e} nodes, setup {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    taus = [x[2] for x in cs]
    # the same cases through the single-case path, one after another, each step queued on the
    # case's stream without a check or a synchronization (the batch's steps below do more: they
    # count non-finite densities, read once per run)
    for c, _, tau, _ in cs:
        c.calc_runge_kutta(tau)
    for c, _, _, _ in cs:
        c.synchronize()
    t0 = time.perf_counter()
    for c, _, tau, _ in cs:
        for _ in range(args.steps):
            c.calc_runge_kutta(tau)
    for c, _, _, _ in cs:
        c.synchronize()
    seq = time.perf_counter() - t0
    b = shockx.Batch([x[0] for x in cs], taus)
    t0 = time.perf_counter()
    factors = b.tune(1)
    tune_s = time.perf_counter() - t0
    b.run(args.warmup)
    b.synchronize()
    t0 = time.perf_counter()
    b.run(args.steps)
    try:
        b.synchronize()                  # checkNAN of every case over the timed steps
    except shockx.ShxError as e:
        bad = [k for k, cnt in enumerate(b.nan_counts()) if cnt > 0]
        raise SystemExit(f"batch cases {[(k, cs[k][3]) for k in bad]} (dims, n, S, Mach, Re): {e}")
    wall = time.perf_counter() - t0
    b.close()
    value = cells * 4 * args.steps / wall
    mix = {"2d_w5": 0, "2d_w9": 0, "3d_w5": 0, "3d_w9": 0}
    for _, _, _, (d, n, S, _, _) in cs:
        mix[f"{d}d_w{S}"] += 1
    out = {"metric": "cell-updates/s/RK-stage", "value": round(value, 1), "unit": "cell-updates/s/RK-stage",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if prec == shockx.FP64 else "f64+f32", "data": "synthetic (seeded cases)",
           "config": {"workload": f"batch of {len(cs)} independent cases ({lo} to {hi} nodes, 2D up to 256^2, W5/W9, "
                                  f"Mach 0.2-3, Re 1e2-1e4), one graph replay per RK step",
                      "nodes": cells, "mix": mix, "seed": args.seed},
           "sequential_value": round(cells * 4 * args.steps / seq, 1),
           "batch_speedup": round(seq / wall, 3), "tuned_resident_blocks_per_cu": factors,
           "tune_s": round(tune_s, 2)}
    for c, _, _, _ in cs:
        c.close()
    print(f"[batch] {lo}..{hi} nodes: {json.dumps(out)}", file=sys.stderr, flush=True)
    return out


def serve_workload(args):
    """Poisson solve requests through the continuous-batching server, one run per arrival rate."""
    from shock_amd import serve
    prec = shockx.MIXED if args.precision == "mixed" else shockx.FP64
    rates = [float(x) for x in str(args.rate).split(",")]
    rows = []
    for rate in rates:
        reqs = serve.make_requests(args.cases, rate, args.seed, n_lo=args.n_lo, n_hi=args.n_hi, precision=prec,
                                   shock_frac=args.shock_frac)
        srv = serve.Server(progress_s=30.0)
        stats, done = srv.run(reqs)
        for r in srv.failed:
            print(f"[serve] request {r.rid} failed ({r.failed}): {r.kind}, n {r.n}, S {r.S}, Mach {r.mach:.2f}, "
                  f"Re {r.reynolds:.0f}, tau {r.tau:.3g}, {r.steps} steps", file=sys.stderr, flush=True)
        row = {"rate": rate, "cell_updates_per_s": round(stats["cell_updates_per_s"], 1),
               "p50_s": round(stats["p50_s"], 3), "p99_s": round(stats["p99_s"], 3), "mean_s": round(stats["mean_s"], 3),
               "queue_wait_p50_s": round(stats["queue_wait_p50_s"], 3), "wall_s": round(stats["wall_s"], 2),
               "requests": stats["requests"], "failed": stats["failed"], "rejected": stats["rejected"],
               "shock_requests": sum(r.kind == "shock" for r in done)}
        print(f"[serve] {json.dumps(row)}", file=sys.stderr, flush=True)
        rows.append(row)
    top = rows[-1]
    out = {"metric": "time-to-solution p50/p99 and aggregate cell-updates/s/RK-stage",
           "value": top["cell_updates_per_s"], "unit": "cell-updates/s/RK-stage", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f64" if prec == shockx.FP64 else "f64+f32",
           "data": "synthetic (seeded requests)",
           "config": {"workload": f"{args.cases} Poisson requests per rate ({args.rate}/s), 3D periodic turbulence "
                                  f"or Mach 1.5-3.5 shock cases (fraction {args.shock_frac}), "
                                  f"{args.n_lo}^3-{args.n_hi}^3 (log-uniform), t_end U(0.5, 2), W5/W9",
                      "seed": args.seed},
           "p50_s": top["p50_s"], "p99_s": top["p99_s"], "sweep": rows}
    print(json.dumps(out), flush=True)


def shock_workload(args):
    """Mach-3 normal shock meeting a vorticity field, 384^3 WENO9 + viscous (Re 5000), supersonic
    inflow and Riemann outflow in x, periodic in y and z, perturbed curvilinear grid."""
    prec = shockx.MIXED if args.precision == "mixed" else shockx.FP64
    n = 384 if args.n == 256 else args.n
    nn = (n, n, n)
    S = 9
    t0 = time.perf_counter()
    coords, prim, inflow = cases.shock_3d(S, nn, args.seed)
    bc = (shockx.BC_INFLOW_SUPERSONIC, shockx.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0)
    case = shockx.Case(dims=3, space_order=S, time_order=4, n=nn, gamma=1.4, mach=3.0, reynolds=5000.0,
                       quirks=shockx.QUIRKS_REFERENCE, bc=bc, inflow=inflow, precision=prec)
    case.build_metric(*coords)
    case.upload_state(prim)
    del coords, prim
    tau = case.cfl_dt(0.5)
    tiles = case.tune(2)
    print(f"[bench] shock tiles {tiles}", file=sys.stderr, flush=True)
    print(f"[bench] shock setup {time.perf_counter() - t0:.1f}s tau={tau:.4g} n={n}", file=sys.stderr, flush=True)
    case.run(tau, args.warmup)
    case.synchronize()
    if not args.no_timing:
        case.set_timing(True)
    t0 = time.perf_counter()
    case.run(tau, args.steps)
    case.synchronize()
    wall = time.perf_counter() - t0
    kt = case.kernel_times() if not args.no_timing else {}
    nan = case.check_nan()
    case.close()
    value = n ** 3 * 4 * args.steps / wall
    kernels = {k: {"avg_ms": round(v[0] / v[1], 4), "launches": v[1]} for k, v in kt.items() if v[1]}
    out = {"metric": "cell-updates/s/RK-stage", "value": round(value, 1), "unit": "cell-updates/s/RK-stage",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if prec == shockx.FP64 else "f64+f32", "data": "synthetic (seeded shock/vorticity)",
           "config": {"workload": f"{n}^3 Mach-3 normal shock + vorticity (rms 0.1 u1), WENO9 + viscous, Re=5000, "
                                  f"inflow/Riemann outflow in x, periodic y/z, perturbed grid, RK4",
                      "precision": args.precision, "seed": args.seed, "tau": tau},
           "non_finite_nodes": nan, "kernels": kernels}
    print(json.dumps(out), flush=True)


def compat_workload(args):
    """The default workload (n^3, WENO5 + viscous, FP64, RK4, the reference's quirks, periodic) at a
    fixed tau through the reference-named entry points and the reference's step loop
    (shock_amd/compat/compat_run: CalcRungeKutta, checkNAN, WriteValuesFromU1ToU2 on host arrays):
    every step hands the state over in host memory, so its rate includes the PCIe transfers and the
    loop's own host copies.  Printed beside the resident rate of the same steps (shx_run)."""
    import subprocess
    import tempfile
    n, S, time_order = args.n, 5, 4
    nn = (n, n, n)
    steps = max(2, min(args.steps, 10))
    case = shockx.Case(dims=3, space_order=S, time_order=time_order, n=nn, gamma=1.4, mach=0.5,
                       reynolds=2000.0, prandtl=0.72, T0=300.0, quirks=shockx.QUIRKS_REFERENCE)
    xe, ye, ze = cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(0.05))
    case.build_metric(xe, ye, ze)
    del xe, ye, ze
    prim = cases.turbulence_3d_state(3, S, nn, args.seed, mach=0.5)
    case.upload_state(prim)
    tau = case.cfl_dt(args.cfl)
    case.run(tau, 1)
    case.synchronize()
    t0 = time.perf_counter()
    case.run(tau, steps)
    case.synchronize()
    resident_ms = 1e3 * (time.perf_counter() - t0) / steps
    met = case.download_metric()
    case.close()
    hdr = np.array([3, S, n, n, n, time_order, 0, steps + 1] + [0] * 6, dtype=np.int32)
    par = np.array([1.4, 0.5, 2000.0, 0.72, 300.0, tau, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0], dtype=np.float64)
    tmp = tempfile.mkdtemp(prefix="shockx_compat_")
    path, out = os.path.join(tmp, "case.bin"), os.path.join(tmp, "out.bin")
    try:
        with open(path, "wb") as f:
            f.write(hdr.tobytes())
            f.write(par.tobytes())
            for a in list(met) + list(prim):
                f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        del met, prim
        exe = os.path.join(ROOT, "shock_amd", "compat", "compat_run")
        print(f"[bench] compat loop: {steps + 1} steps of {n}^3 at tau {tau:.6g}", file=sys.stderr, flush=True)
        r = subprocess.run([exe, path, out], capture_output=True, text=True, timeout=900)
        print(r.stdout.strip(), file=sys.stderr, flush=True)
        if r.returncode:
            print(r.stderr, file=sys.stderr)
            sys.exit(r.returncode)
    finally:
        for fn in (path, out):
            if os.path.exists(fn):
                os.remove(fn)
        os.rmdir(tmp)
    m = re.search(r"first step ([0-9.]+) ms, then ([0-9.]+) ms per step", r.stdout)
    first_ms, ms = float(m.group(1)), float(m.group(2))
    line = {"metric": "cell-updates/s/RK-stage", "value": round(n ** 3 * time_order / (ms * 1e-3), 1),
            "unit": "cell-updates/s/RK-stage", "n_gpus": 1, "steps": steps, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded random-phase turbulence)",
            "config": {"workload": f"{n}^3 periodic curvilinear block, WENO5 + viscous, FP64, RK4, Re=2000, "
                                   "fixed tau, through the reference-named host loop (compat_run)",
                       "state": "host memory between steps (PCIe-inclusive)"},
            "first_step_ms": round(first_ms, 3),
            "resident_ms_per_step": round(resident_ms, 3),
            "resident_value": round(n ** 3 * time_order / (resident_ms * 1e-3), 1)}
    print(json.dumps(line), flush=True)


def mms_ladder(args):
    """Manufactured-solution ladder in FP64 and mixed forms: one JSON line with the errors, the
    observed orders and the throughput of each rung."""
    from shock_amd import ladder
    rungs = [int(x) for x in args.ladder.split(",")]
    res = {}
    forms = [f for f in (("fp64", shockx.FP64), ("mixed", shockx.MIXED)) if args.ladder_forms in ("both", f[0])]
    for prec, pv in forms:
        rows = []
        for n in rungs:
            r = ladder.rung(n, S=args.scheme, precision=pv, t_end=args.t_end, amp=args.stretch)
            print(f"[ladder] {prec} n={n} steps={r['steps']} L2_rho={r['full']['L2_rho']:.3e} "
                  f"Linf_rho={r['full']['Linf_rho']:.3e} L2_p={r['full']['L2_p']:.3e} "
                  f"drho_last={r['last_chunk_drho']:.2e} {r['cell_updates_per_s']:.3e} cu/s",
                  file=sys.stderr, flush=True)
            rows.append(r)
        res[prec] = {"rungs": rows, "order_L2_rho": ladder.orders(rows, "L2_rho"),
                     "order_Linf_rho": ladder.orders(rows, "Linf_rho"),
                     "order_L2_p": ladder.orders(rows, "L2_p"),
                     "order_L2_rho_last_plane": ladder.orders(rows, "L2_rho", "last_plane")}
    both = "fp64" in res and "mixed" in res
    ratio = [m["full"]["L2_rho"] / f["full"]["L2_rho"] for m, f in zip(res["mixed"]["rungs"], res["fp64"]["rungs"])] \
        if both else None
    speed = [m["cell_updates_per_s"] / f["cell_updates_per_s"] for m, f in zip(res["mixed"]["rungs"], res["fp64"]["rungs"])] \
        if both else None
    out = {"metric": "manufactured-solution error ladder", "unit": "L2/Linf of rho and p", "n_gpus": 1,
           "dtype": "f64 and f64+f32", "data": "manufactured solution (reference subsonic NS set)",
           "config": {"workload": f"3D MMS ladder {args.ladder}, WENO{args.scheme} + viscous, RK4, "
                                  f"t_end={args.t_end}, {'uniform' if args.stretch == 0 else 'stretched'} grid, "
                                  f"Dirichlet faces",
                      "scheme": f"WENO{args.scheme}", "quirks": "diag3d metric only"},
           "fp64": res.get("fp64"), "mixed": res.get("mixed"), "mixed_over_fp64_L2_rho": ratio,
           "mixed_speedup": speed}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
