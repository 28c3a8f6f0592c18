"""The reference-named entry points (shock_amd/compat/shock_compat.h: CalcRungeKutta, checkNAN,
WriteValuesFromU1ToU2 over restated strct_* structs) driven by the reference's step loop
(shock_amd/compat/compat_run.c), against the oracle on the same cases."""
import os
import subprocess

import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, setup_2d_mms, setup_3d_turb
from shock_amd import build, cases

pytestmark = pytest.mark.gpu

# face codes of compat_run: 0 neighbour, 1 Inflow (supersonic), 2 Outflow (Riemann), 3 WallInviscid,
# 4 WallViscous, 5 Farfield
CODE = {orc.BC_PERIODIC: 0, orc.BC_INFLOW_SUPERSONIC: 1, orc.BC_OUTFLOW_RIEMANN: 2, orc.BC_WALL_SLIP: 3,
        orc.BC_WALL_NOSLIP: 4, orc.BC_FARFIELD: 5}


def _run(tmp_path, o, met, prim, tau, steps, bc, inflow=(1.0, 1.0, 0.0, 0.0, 1.0)):
    c = o.c
    hdr = np.array([c.dims, c.S, c.n[0], c.n[1], c.n[2], c.time_order, c.inviscid, steps]
                   + [CODE[b] for b in bc], dtype=np.int32)
    par = np.array([c.gamma, c.mach, c.reynolds, c.prandtl, c.T0, tau, *inflow, 1.0, 0.0, 1.0], dtype=np.float64)
    case = os.path.join(tmp_path, "case.bin")
    out = os.path.join(tmp_path, "out.bin")
    with open(case, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        for a in list(met) + list(prim):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    assert os.path.exists(build.COMPAT_RUN), "build with python -m shock_amd.build"
    r = subprocess.run([build.COMPAT_RUN, case, out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    data = np.fromfile(out, dtype=np.float64).reshape(5, *o.shape)
    return [data[m] for m in range(5)]


@pytest.mark.parametrize("S,time_order", [(5, 4), (9, 4), (5, 3)])
def test_compat_step_loop_periodic_2d(tmp_path, S, time_order):
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=S, time_order=time_order)
    tau = o.cfl_dt(0.5, met, prim)
    got = _run(tmp_path, o, met, prim, tau, 10, (orc.BC_PERIODIC,) * 6)
    want = o.run(tau, met, prim, 10)
    errs = rel_linf(conserved(got), conserved(want), interior(o, 2, nn))
    print(f"[compat] periodic 2D W{S} RK{time_order}, 10 steps: " + " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs[:3] + errs[4:]) <= 1e-11


def test_compat_step_loop_faces_3d(tmp_path):
    # inflow / outflow in x, slip walls in y, periodic z, on a curvilinear grid; the metric the
    # compat layer receives is the reference's (mirrored at walls)
    n = (14, 10, 12)
    inflow = (1.0, 2.5, 0.0, 0.0, 1.0)
    bc = (orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RIEMANN, orc.BC_WALL_SLIP, orc.BC_WALL_SLIP, 0, 0)
    o = orc.Oracle(dims=3, S=5, n=n, bc=bc, inflow=inflow, mach=0.5, reynolds=1000.0)
    xe, ye, ze = cases.ext_coords(3, 5, n, cases.sine_bump_3d(0.05), periodic=(False, False, True))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    o.mirror_metric(met)
    from test_gpu_boundaries import _stream
    prim = _stream(o, 3, (xe, ye, ze), inflow)
    tau = 0.2 / (n[0] * 5.0)
    got = _run(tmp_path, o, met, prim, tau, 5, bc, inflow)
    want = o.run(tau, met, prim, 5)
    errs = rel_linf(conserved(got), conserved(want), interior(o, 3, n))
    print("[compat] inflow/outflow + slip walls 3D, 5 steps: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-11


def test_compat_step_loop_periodic_3d_large_fields(tmp_path):
    # fields of 5 MB (86^3 padded nodes): the host passes of the loop (shadow compare and copy,
    # WriteValuesFromU1ToU2) and the staged transfers run on several host threads
    o, coords, met, prim, nn = setup_3d_turb(n=80, S=5)
    orc.lib().orc_set_threads(min(16, len(os.sched_getaffinity(0))))
    tau = o.cfl_dt(0.5, met, prim)
    got = _run(tmp_path, o, met, prim, tau, 3, (orc.BC_PERIODIC,) * 6)
    want = o.run(tau, met, prim, 3)
    errs = rel_linf(conserved(got), conserved(want), interior(o, 3, nn))
    print("[compat] periodic 3D 80^3, 3 steps: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-12
