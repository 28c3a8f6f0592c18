"""Batches with boundary faces: supersonic and subsonic inflow, Riemann, normal and Rudy outflow,
slip, no-slip and isothermal walls, farfield and manufactured-solution faces (with the source) mixed with a periodic case in one batch.  Each case
must reach exactly the state it reaches when stepped alone (the single-case path's parity against
the oracle is covered by the boundary tests); three of the cases are also held against the oracle
directly."""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, setup_3d_turb
from shock_amd import cases, mms, shockx
from test_gpu_boundaries import IN_NORMAL, IN_RUDY, INOUT, NOSLIP, SLIP, _mapping, _setup, _stream

pytestmark = pytest.mark.gpu

INFLOW = (1.0, 2.5, 0.0, 0.0, 1.0)


def _faces_case(dims, S, n, bc, inviscid, quirks, mp):
    o, g, met, coords = _setup(dims, S, n, bc, INFLOW, inviscid, quirks, mapping=_mapping(dims, mp))
    prim = _stream(o, dims, coords, INFLOW)
    g.upload_state(prim)
    return {"o": o, "c": g, "met": met, "prim": prim, "n": n, "src": None, "bs": None, "tag": f"faces {bc[:2 * dims]} S={S}"}


def _mms_case(S, n, quirks):
    bc = (orc.BC_MMS,) * 6
    o, g, met, coords = _setup(3, S, n, bc, INFLOW, False, quirks, mms_on=True, reynolds=100.0)
    x, y, z = cases.interior_coords(3, S, n, *coords)
    P = mms.reference_params(3, False, False, mach=0.5)
    ex = mms.exact_state(P, x, y, z, 1.4, 0.5)
    src = mms.source(P, x, y, z, 1.4, 0.5, 100.0, 0.72, False)
    g.upload_boundary_state(ex)
    g.upload_source(src)
    g.upload_state(ex[:5])
    return {"o": o, "c": g, "met": met, "prim": ex[:5], "n": n, "src": src, "bs": ex, "tag": f"mms S={S}"}


def _periodic_case(n, S, seed):
    o, (xe, ye, ze), met, prim, nn = setup_3d_turb(n=n, S=S, inviscid=False, seed=seed)
    c = shockx.Case(dims=3, space_order=S, time_order=4, n=nn, gamma=1.4, mach=o.c.mach, reynolds=o.c.reynolds)
    c.build_metric(xe, ye, ze)
    c.upload_state(prim)
    return {"o": o, "c": c, "met": met, "prim": prim, "n": nn, "src": None, "bs": None, "tag": "periodic"}


def test_batch_with_faces_equals_cases_alone():
    made = [
        _faces_case(2, 5, (24, 20, 1), INOUT + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, None),
        _faces_case(2, 9, (24, 20, 1), INOUT + INOUT + (0, 0), False, orc.QUIRKS_REFERENCE, None),
        _faces_case(3, 5, (14, 12, 12), INOUT + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
        _faces_case(3, 9, (14, 12, 12), INOUT + (0, 0, 0, 0), True, orc.QUIRKS_REFERENCE, "bump"),
        _faces_case(3, 5, (14, 12, 12), IN_RUDY + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
        _faces_case(2, 5, (24, 20, 1), IN_RUDY + IN_NORMAL + (0, 0), True, orc.QUIRKS_REFERENCE, None),
        _faces_case(3, 5, (14, 12, 12), INOUT + SLIP + (0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
        _faces_case(2, 9, (24, 20, 1), INOUT + NOSLIP + (0, 0), False, orc.QUIRKS_REFERENCE, None),
        _mms_case(5, (16, 16, 16), orc.QUIRK_METRIC_DIAG3D),
        _periodic_case(12, 5, 11),
        _faces_case(2, 5, (24, 20, 1), (orc.BC_INFLOW_SUB_RIEMANN, orc.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0), False,
                    orc.QUIRKS_REFERENCE, None),
        _faces_case(3, 5, (14, 12, 12), INOUT + (orc.BC_FARFIELD, orc.BC_WALL_ISOTHERMAL, 0, 0), False,
                    orc.QUIRKS_REFERENCE, "bump"),
        # the full 3D metric (no diagonal quirk): derivative march with T derivatives, W5 and W9
        _faces_case(3, 5, (14, 12, 12), INOUT + (0, 0, 0, 0), False, 0, "bump"),
        _faces_case(3, 9, (14, 12, 12), INOUT + (0, 0, 0, 0), False, 0, "bump"),
    ]
    steps = 3
    taus = [0.2 / (m["n"][0] * 5.0) for m in made]
    alone = []
    for m, tau in zip(made, taus):
        m["c"].run(tau, steps)
        alone.append(m["c"].download_state())
        m["c"].upload_state(m["prim"])
    with shockx.Batch([m["c"] for m in made], taus) as b:
        b.run(steps)
        b.synchronize()
        for k, (m, tau, ref) in enumerate(zip(made, taus, alone)):
            o, dims = m["o"], m["o"].c.dims
            got = m["c"].download_state()
            sl = interior(o, dims, m["n"])
            diff = max(float(np.max(np.abs(a[sl] - r[sl]))) for a, r in zip(got, ref))
            print(f"[batch-faces] case {k} {m['tag']}: max |batch - alone| = {diff:.2e}", flush=True)
            assert diff == 0.0
            if k in (0, 4, 6, 8, 10, 11):
                want = o.run(tau, m["met"], m["prim"], steps, src=m["src"], bstate=m["bs"])
                errs = rel_linf(conserved(got), conserved(want), sl)
                if dims == 2:
                    errs = errs[:3] + errs[4:]
                print(f"[batch-faces] case {k} vs oracle: " + " ".join(f"{e:.1e}" for e in errs), flush=True)
                assert max(errs) <= 1e-11
        # the batch keeps stepping from its own state, ghosts included
        b.run(2)
        b.synchronize()
        after = [m["c"].download_state() for m in made]
    for m, tau, ref, st in zip(made, taus, alone, after):
        m["c"].upload_state(m["prim"])
        m["c"].run(tau, steps + 2)
        sl = interior(m["o"], m["o"].c.dims, m["n"])
        again = m["c"].download_state()
        diff = max(float(np.max(np.abs(a[sl] - r[sl]))) for a, r in zip(st, again))
        print(f"[batch-faces] {m['tag']}: 5 steps batch vs alone {diff:.2e}", flush=True)
        assert diff == 0.0
        m["c"].close()
