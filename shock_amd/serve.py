"""Continuous-batching solve server over the C ABI (no reference counterpart: SHOCK runs one
case per job).

Requests arrive as a Poisson stream.  Each asks for a 3D solve of size n^3 up to an end time;
it is admitted when its HBM footprint fits the budget next to the cases already resident (and
a slot is free), its state stays on the device, and every resident periodic case is advanced
by the shared launches of one batch (shx_batch).  A request retires when it has taken its
number of steps; the batch is rebuilt whenever the resident set changes.  New cases are created
and uploaded while the resident batch's chunk runs on the device.  A request is a
periodic block of decaying turbulence or a normal shock meeting a vorticity field with
inflow/outflow faces in x.

Each request's steps and time step are fixed at admission (CFL on its initial state), so its
result does not depend on what else is resident: served alone or in a mix, it runs the same
kernels on the same data.
"""
import math
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import cases, shockx

FIELDS_PER_CASE = 49          # 3 x 5 state + Q + Q_sum + 10 metric + 12 derivatives + T + mu


def footprint_bytes(n, S):
    """Device bytes of one 3D case with n^3 interior nodes (the arrays shx_create allocates)."""
    ng = n + S + 1
    pitch = (ng + 15) // 16 * 16
    return FIELDS_PER_CASE * 8 * ng * ng * pitch


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the run
    n: int
    S: int
    t_end: float
    mach: float
    reynolds: float
    seed: int
    precision: int = shockx.FP64
    kind: str = "turbulence"       # "turbulence" (periodic) or "shock" (inflow/outflow in x)
    # filled in by the server
    admitted: float = float("nan")
    done: float = float("nan")
    steps: int = 0
    tau: float = 0.0
    result: list = field(default_factory=list)
    failed: str = ""               # why the request ended without a result (checkNAN)


def make_requests(count, rate, seed, n_lo=64, n_hi=256, t_lo=0.5, t_hi=2.0, precision=shockx.FP64,
                  shock_frac=0.0):
    """Poisson arrivals (rate per second); sizes log-uniform in [n_lo, n_hi]; end times uniform in
    [t_lo, t_hi]; WENO5 or WENO9 at random; Reynolds log-uniform 1e2-1e4.  A fraction shock_frac
    of the requests are shock cases (upstream Mach uniform 1.5-3.5), the others random-phase
    turbulence on the periodic curvilinear box (turbulent Mach uniform 0.2-3)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        S = int(rng.choice([5, 9]))
        t_end = float(rng.uniform(t_lo, t_hi))
        re = float(math.exp(rng.uniform(math.log(1e2), math.log(1e4))))
        shock = bool(rng.uniform() < shock_frac)
        mach = float(rng.uniform(1.5, 3.5)) if shock else float(rng.uniform(0.2, 3.0))
        out.append(Request(rid=i, arrival=t, n=n, S=S, t_end=t_end, mach=mach, reynolds=re,
                           seed=int(rng.integers(1, 2 ** 31 - 1)), precision=precision,
                           kind="shock" if shock else "turbulence"))
    return out


def prepare(r):
    """Host inputs of one request (grid, initial state and inflow state): built off the critical
    path."""
    nn = (r.n, r.n, r.n)
    if r.kind == "shock":
        return cases.shock_3d(r.S, nn, r.seed, mach=r.mach)
    coords = cases.periodic_ext_coords(3, r.S, nn, cases.sine_bump_3d(0.05))
    return coords, cases.turbulence_3d_state(3, r.S, nn, r.seed, mach=r.mach), None


def open_case(r, prep=None, cfl=0.5):
    """Create, upload and size one request's case: the periodic curvilinear box with turbulence,
    or the shock case with supersonic inflow and Riemann outflow in x."""
    nn = (r.n, r.n, r.n)
    coords, state, inflow = prep if prep is not None else prepare(r)
    if r.kind == "shock":
        c = shockx.Case(dims=3, space_order=r.S, time_order=4, n=nn, gamma=1.4, mach=r.mach, reynolds=r.reynolds,
                        quirks=shockx.QUIRKS_REFERENCE, bc=(shockx.BC_INFLOW_SUPERSONIC, shockx.BC_OUTFLOW_RIEMANN,
                                                            0, 0, 0, 0), inflow=inflow, precision=r.precision)
    else:
        c = shockx.Case(dims=3, space_order=r.S, time_order=4, n=nn, mach=r.mach, reynolds=r.reynolds,
                        precision=r.precision)
    c.build_metric(*coords)
    c.upload_state(state)
    # the convective CFL step (the reference's), capped by the explicit viscous limit on the unit
    # box (spacing 1/n): low-Reynolds requests on fine grids are otherwise unstable
    h = 1.0 / r.n
    tau = min(c.cfl_dt(cfl), 0.2 * r.reynolds * h * h / 3.0)
    r.steps = max(1, int(math.ceil(r.t_end / tau)))
    r.tau = r.t_end / r.steps          # land exactly on t_end
    return c


def run_alone(r):
    """Reference result of one request: its case advanced by itself."""
    c = open_case(r)
    c.run(r.tau, r.steps)
    st = c.download_state()
    c.close()
    return st


class Server:
    """budget_bytes: device memory the resident cases may use; max_slots: resident cases at most;
    chunk: steps between scheduling decisions.  open_fn / batch_cls / footprint_fn are the case
    factory, the batch type and the footprint model (replaceable for host-only tests)."""

    def __init__(self, budget_bytes=270 * 2 ** 30, max_slots=64, chunk=20, keep_results=False,
                 open_fn=open_case, batch_cls=None, footprint_fn=footprint_bytes, progress_s=None):
        self.progress_s = progress_s        # seconds between progress lines on stderr (None: quiet)
        self.budget = budget_bytes
        self.max_slots = max_slots
        self.chunk = chunk
        self.keep = keep_results
        self.open_fn = open_fn
        self.batch_cls = batch_cls or shockx.Batch
        self.footprint = footprint_fn

    def run(self, requests):
        with ThreadPoolExecutor(max_workers=2) as pool:
            return self._run(requests, pool)

    def _run(self, requests, pool):
        """One loop turn: take arrivals; start the next chunk of the resident batch (its graph
        replays are queued on the device and run while the host goes on); admit waiting requests
        that fit, creating and uploading their cases while that chunk runs; wait for the chunk;
        retire the requests it finished; rebuild the batch when the resident set changed."""
        t0 = time.perf_counter()
        preps = {}
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting = []                        # arrived, not yet admitted (FIFO)
        active = {}                         # rid -> [request, case, steps left], in the batch
        admitted = []                       # opened this turn, join the batch after the chunk
        used = 0
        batch = None
        batch_ids = []
        finished = []
        failed = []
        rejected = []
        cell_updates = 0.0
        last_note = 0.0
        while pending or waiting or active:
            now = time.perf_counter() - t0
            if self.progress_s and now - last_note >= self.progress_s:
                last_note = now
                print(f"[serve] t={now:.0f}s: {len(active)} active, {len(waiting)} waiting, {len(pending)} not yet "
                      f"arrived, {len(finished)} finished, {len(failed)} failed", file=sys.stderr, flush=True)
            while pending and pending[0].arrival <= now:
                r = pending.pop(0)
                if self.footprint(r.n, r.S) > self.budget:
                    # can never fit, even on an empty device: refused on arrival (queued, it would
                    # hold the line for ever)
                    rejected.append(r)
                    continue
                waiting.append(r)
                if self.open_fn is open_case:
                    preps[r.rid] = pool.submit(prepare, r)
            k = 0
            if active:
                if batch is None:
                    batch_ids = sorted(active)
                    batch = self.batch_cls([active[i][1] for i in batch_ids], [active[i][0].tau for i in batch_ids])
                k = min(self.chunk, min(active[i][2] for i in batch_ids))
                batch.run(k)                # queued: the device works while admissions proceed
            # admit in arrival order while memory and slots allow
            while waiting and len(active) + len(admitted) < self.max_slots:
                r = waiting[0]
                need = self.footprint(r.n, r.S)
                if used + need > self.budget:
                    break
                waiting.pop(0)
                if r.rid in preps:
                    c = self.open_fn(r, preps.pop(r.rid).result())
                else:
                    c = self.open_fn(r)
                r.admitted = time.perf_counter() - t0
                admitted.append([r, c, r.steps])
                used += need
            if k:
                blown = set()
                try:
                    batch.synchronize()
                except shockx.ShxError as e:
                    # the fused checkNAN found non-finite densities: those requests end as failed
                    # (their states hold 1e6 where checkNAN wrote it); the others go on
                    if e.code != shockx.ERR_NAN:
                        raise
                    blown = {i for i, cnt in zip(batch_ids, batch.nan_counts()) if cnt > 0}
                for i in batch_ids:
                    active[i][2] -= k
                    cell_updates += active[i][0].n ** 3 * 4 * k
                retired = [i for i in batch_ids if active[i][2] == 0 or i in blown]
                if retired or admitted:
                    batch.close()
                    batch = None
                for i in retired:
                    r, c, _ = active.pop(i)
                    if self.keep:
                        r.result = c.download_state()
                    else:
                        c.synchronize()
                    r.done = time.perf_counter() - t0
                    c.close()
                    used -= self.footprint(r.n, r.S)
                    if i in blown:
                        r.failed = "checkNAN: non-finite density"
                        failed.append(r)
                    else:
                        finished.append(r)
            elif batch is not None and admitted:
                batch.close()
                batch = None
            for e in admitted:
                active[e[0].rid] = e
            admitted = []
            if not active and pending and not waiting:
                time.sleep(max(0.0, min(0.05, pending[0].arrival - (time.perf_counter() - t0))))
        if batch is not None:
            batch.close()
        wall = time.perf_counter() - t0
        tts = np.array([r.done - r.arrival for r in finished] or [float("nan")])
        self.failed = failed
        return {
            "requests": len(finished), "rejected": len(rejected), "failed": len(failed), "wall_s": wall,
            "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
            "mean_s": float(np.mean(tts)), "cell_updates_per_s": cell_updates / wall,
            "queue_wait_p50_s": float(np.percentile([r.admitted - r.arrival for r in finished] or [float("nan")], 50)),
        }, finished
