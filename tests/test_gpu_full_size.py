"""Parity at the throughput workload's real shape: the 256^3 periodic curvilinear block (WENO5 +
viscous, FP64, RK4, Re 2000, Mach 0.5, seeded random-phase turbulence, the reference's quirks)
against the oracle after one step, and a 128^3 block of the same workload after ten steps.  The
oracle runs on up to 16 host threads (about 15 s per 256^3 step); its results do not depend on
the thread count.  Bars: 1e-12 relative L-infinity per conserved variable (one step) and 1e-10
(many steps)."""
import os

import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf
from shock_amd import cases, shockx

pytestmark = pytest.mark.gpu

SEED = 20251202


def _workload(n):
    nn = (n, n, n)
    o = orc.Oracle(dims=3, S=5, n=nn, time_order=4, mach=0.5, reynolds=2000.0)
    orc.lib().orc_set_threads(min(16, len(os.sched_getaffinity(0))))
    coords = cases.periodic_ext_coords(3, 5, nn, cases.sine_bump_3d(0.05))
    met = o.metric_from_coords(*coords)
    o.ghost_periodic(met)
    prim = cases.turbulence_3d_state(3, 5, nn, SEED, mach=0.5)
    return o, nn, coords, met, prim


@pytest.mark.parametrize("n,steps,bar", [(256, 1, 1e-12), (128, 10, 1e-10)])
def test_turbulence_workload_against_oracle(n, steps, bar):
    o, nn, coords, met, prim = _workload(n)
    tau = o.cfl_dt(0.5, met, prim)
    with shockx.Case(dims=3, space_order=5, time_order=4, n=nn, gamma=1.4, mach=0.5, reynolds=2000.0,
                     prandtl=0.72, T0=300.0, quirks=shockx.QUIRKS_REFERENCE) as g:
        g.build_metric(*coords)
        g.upload_state(prim)
        assert g.cfl_dt(0.5) == pytest.approx(tau, rel=1e-15, abs=0)
        g.run(tau, steps)
        out = g.download_state()
    print(f"[full size] GPU done for {n}^3, {steps} step(s); oracle running", flush=True)
    ref = o.run(tau, met, prim, steps)
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    print(f"[full size] {n}^3 {steps} step(s) vs oracle: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert np.isfinite(out[0][sl]).all()
    assert max(errs) <= bar
