"""End-of-step checks fused into the last stage's update (shx_run, shx_run_cfl).

* checkNAN (src/Functions.c:5419-5467) after every step of shx_run: a NaN planted in the start
  state spreads over its stencil in one step; the run stops after that step with SHX_ERR_NAN, and
  the state, with the non-finite nodes set to 1e6, equals the oracle's step followed by its
  checkNAN; the count is the oracle's.
* CFL stepping (shx_run_cfl): each step's time step is CFL / max node CFL sum of the state it
  starts from.  After the first step that maximum comes from the fused reduction of the previous
  step's last stage; it must equal the separate reduction (shx_cfl_dt) bit for bit, for each kind
  of last-stage kernel (fused zeta kernel for WENO5, WENO9 and the mixed form, last viscous sweep
  with a general metric, 2D viscous in FP64 and mixed, inviscid update).  Several steps match the oracle stepping with its own CFL step each step.
"""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, setup_2d_mms, setup_3d_turb
from shock_amd import shockx

pytestmark = pytest.mark.gpu


def _gpu_case(o, coords, prim, nn, dims, quirks, precision=shockx.FP64):
    g = shockx.Case(dims=dims, space_order=o.c.S, time_order=4, n=nn, mach=o.c.mach, reynolds=o.c.reynolds,
                    inviscid=bool(o.c.inviscid), quirks=quirks, precision=precision)
    g.build_metric(*coords)
    g.upload_state(prim)
    return g


CASES = [
    ("3d_fused_zeta", 3, False, orc.QUIRKS_REFERENCE, 5, shockx.FP64),
    ("3d_fused_zeta_w9", 3, False, orc.QUIRKS_REFERENCE, 9, shockx.FP64),
    ("3d_fused_zeta_mixed", 3, False, orc.QUIRKS_REFERENCE, 5, shockx.MIXED),
    ("3d_general_metric", 3, False, 0, 5, shockx.FP64),
    ("3d_inviscid", 3, True, orc.QUIRKS_REFERENCE, 5, shockx.FP64),
    ("2d_viscous", 2, False, 0, 5, shockx.FP64),
    ("2d_viscous_mixed", 2, False, 0, 5, shockx.MIXED),
]


def _setup(dims, inviscid, quirks, S=5):
    if dims == 3:
        return setup_3d_turb(n=16, S=S, inviscid=inviscid, quirks=quirks)
    return setup_2d_mms(n=24, S=S, inviscid=inviscid)


@pytest.mark.parametrize("name,dims,inviscid,quirks,S,prec", CASES, ids=[c[0] for c in CASES])
def test_fused_cfl_max_equals_separate_reduction(name, dims, inviscid, quirks, S, prec):
    o, coords, met, prim, nn = _setup(dims, inviscid, quirks, S)
    with _gpu_case(o, coords, prim, nn, dims, quirks, prec) as a, \
            _gpu_case(o, coords, prim, nn, dims, quirks, prec) as b:
        t2 = a.run_cfl(0.5, 2)                 # second step from the fused maximum
        t1 = b.run_cfl(0.5, 1)
        dt1 = b.cfl_dt(0.5)                    # separate reduction on the same state
        print(f"[endstep] {name}: dt0 {t1:.17g} dt1 fused {t2 - t1:.17g} separate {dt1:.17g}", flush=True)
        assert t2 == t1 + dt1
        sa, sb = a.download_state(), b.download_state()
        b.run(dt1, 1)
        assert all(np.array_equal(x, y) for x, y in zip(a.download_state(), b.download_state()))
        assert not all(np.array_equal(x, y) for x, y in zip(sa, sb))


def test_cfl_stepping_against_oracle():
    o, coords, met, prim, nn = setup_3d_turb(n=16, S=5)
    steps = 6
    state = [a.copy() for a in prim]
    t_ref = 0.0
    for _ in range(steps):
        dt = o.cfl_dt(0.5, met, state)
        state = o.run(dt, met, state, 1)
        t_ref += dt
    with _gpu_case(o, coords, prim, nn, 3, orc.QUIRKS_REFERENCE) as g:
        t = g.run_cfl(0.5, steps)
        out = g.download_state()
    errs = rel_linf(conserved(out), conserved(state), interior(o, 3, nn))
    print(f"[endstep] CFL stepping {steps} steps: time {t:.17g} oracle {t_ref:.17g}; state "
          + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert abs(t - t_ref) <= 1e-13 * t_ref
    assert max(errs) <= 1e-12


@pytest.mark.parametrize("cfl_mode", [False, True])
def test_run_stops_at_nan_with_checknan_state(cfl_mode):
    """A NaN planted in the start state: the run stops after the first step with SHX_ERR_NAN, and
    the state is that step's (shx_calc_runge_kutta, no check, on a twin case) with exactly the
    nodes of non-finite density set to rho = u = v = w = 1e6 and their count reported.  Which nodes
    a NaN reaches in one step depends on how each operation treats NaN (a clamp through fmax drops
    it), so the node set is the GPU step's own; the oracle's step is checked to find NaN too."""
    o, coords, met, prim, nn = setup_3d_turb(n=16, S=5)
    tau = o.cfl_dt(0.5, met, prim)
    bad = [a.copy() for a in prim]
    bad[0][o.g + 5, o.g + 6, o.g + 7] = np.nan
    o.ghost_periodic(bad)
    ref = [a.copy() for a in o.run(tau, met, bad, 1)]
    assert o.check_nan(ref) > 1
    sl = interior(o, 3, nn)
    with _gpu_case(o, coords, prim, nn, 3, orc.QUIRKS_REFERENCE) as g, \
            _gpu_case(o, coords, prim, nn, 3, orc.QUIRKS_REFERENCE) as raw:
        g.run(tau, 2)                          # finite state: no error
        g.upload_state(bad)
        raw.upload_state(bad)
        with pytest.raises(shockx.ShxError) as ei:
            if cfl_mode:
                g.run_cfl(0.5, 3)
            else:
                g.run(tau, 3)
        assert ei.value.code == shockx.ERR_NAN
        msg = str(ei.value)
        print(f"[endstep] {msg}", flush=True)
        if cfl_mode:
            # the CFL maximum of the start state skips the NaN node (fmax); the step then finds it
            dt = raw.cfl_dt(0.5)
        else:
            dt = tau
        raw.calc_runge_kutta(dt)
        want = raw.download_state()
        out = g.download_state()
    nonfinite = ~np.isfinite(want[0][sl])
    n_bad = int(np.count_nonzero(nonfinite))
    assert n_bad > 1
    assert f"checkNAN: {n_bad} interior nodes" in msg and "after step 1 of 3" in msg
    for m in range(4):
        w = want[m][sl].copy()
        w[nonfinite] = 1e6
        assert np.array_equal(out[m][sl], w, equal_nan=True), m
    assert np.array_equal(out[4][sl], want[4][sl], equal_nan=True)
