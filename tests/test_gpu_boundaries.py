"""GPU parity of the boundary writers (supersonic inflow; normal, Riemann and isentropic subsonic
inflow; Riemann, normal and Rudy subsonic outflow; slip, no-slip and isothermal walls; farfield;
manufactured-solution faces) and the manufactured-solution path: libshockx.so against
the oracle on identical inputs (same bar as the periodic parity tests: relative L-infinity per
conserved variable <= 1e-12 after one step, <= 1e-10 after many).

The stage residual itself is compared at 1e-11 here rather than 1e-12: the writers store p and c
for ghost nodes while the flux kernels recompute them from rho, u, v, w, e (equal to the stored
values within an ulp), and the residual of a perturbed uniform stream is a small difference of
large fluxes.  Recomputing p and c in the oracle itself moves its residual by up to 4e-13 for
these states.

Every test also measures the oracle's own conditioning: the oracle is rerun on inputs perturbed
by one ulp of random noise and that spread is printed beside the GPU's difference.  The stage
residual may differ by up to three times that spread where it exceeds 1e-11; stepped states are
held to the fixed bars with no allowance (cases whose spread reaches the bar would be listed in
SPREAD_EXCEPTIONS with their own bar; WENO9 with the reference's epsilon, 10 sqrt(DBL_MIN), has a
discontinuous weight map and is where that can happen)."""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf
from shock_amd import cases, mms, shockx

pytestmark = pytest.mark.gpu

GAMMA, MACH = 1.4, 0.5
INOUT = (orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RIEMANN)
IN_NORMAL = (orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_NORMAL)
IN_RUDY = (orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RUDY)
SLIP = (orc.BC_WALL_SLIP, orc.BC_WALL_SLIP)
NOSLIP = (orc.BC_WALL_NOSLIP, orc.BC_WALL_NOSLIP)
OUTFLOW = (0.95, 3.0)          # p_out, AlphaNonRef
PRANDTL, T_WALL = 0.9, 1.3


def _noise(o, dims, n, met, prim, steps, tau=0.0, src=None, bstate=None, trials=3):
    """largest per-field relative change of the oracle's output (steps = 0: the stage residual)
    under one ulp of random input noise"""
    sl = interior(o, dims, n)

    def out(p):
        if steps == 0:
            return o.rhs(met, [a.copy() for a in p], src, bstate)[0]
        return conserved(o.run(tau, met, p, steps, src=src, bstate=bstate))

    base = out(prim)
    rng = np.random.default_rng(12345)
    worst = np.zeros(5)
    for _ in range(trials):
        pert = [a * (1.0 + 2.2e-16 * rng.standard_normal(a.shape)) for a in prim]
        worst = np.maximum(worst, rel_linf(out(pert), base, sl))
    return worst


# Stepped states are held to the fixed bar (1e-12 after one step, 1e-11 after ten) with no
# allowance, except for the cases listed here with their own bar: those where the oracle's own
# one-ulp spread (printed by every test) comes within reach of the fixed bar.  Key: the test tag.
SPREAD_EXCEPTIONS = {}


def _bar(errs, fixed, noise, dims, steps=0, tag=""):
    """steps == 0 (the stage residual, a small difference of large fluxes): the fixed bar or three
    times the oracle's one-ulp spread; steps >= 1: the fixed bar (or the listed exception)."""
    keep = [0, 1, 2, 4] if dims == 2 else [0, 1, 2, 3, 4]
    if steps == 0:
        return all(errs[i] <= max(fixed, 3.0 * noise[i]) for i in keep)
    bar = SPREAD_EXCEPTIONS.get((tag, steps), fixed)
    return all(errs[i] <= bar for i in keep)


def _report(tag, errs):
    print(f"[parity] {tag}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)


def _setup(dims, S, n, bc, inflow, inviscid=False, quirks=orc.QUIRKS_REFERENCE, mms_on=False,
           mapping=None, reynolds=1000.0, prandtl=0.72):
    o = orc.Oracle(dims=dims, S=S, n=n, bc=bc, inflow=inflow, mach=MACH, reynolds=reynolds,
                   inviscid=int(inviscid), quirks=quirks, mms=int(mms_on), outflow=OUTFLOW,
                   prandtl=prandtl, T_wall=T_WALL)
    per = tuple(bc[2 * d] == orc.BC_PERIODIC for d in range(3))
    xe, ye, ze = cases.ext_coords(dims, S, n, mapping or cases.cartesian_unit(dims), periodic=per)
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    o.mirror_metric(met)
    g = shockx.Case(dims=dims, space_order=S, time_order=4, n=n, gamma=GAMMA, mach=MACH,
                    reynolds=reynolds, prandtl=prandtl, inviscid=inviscid, manufactured=mms_on, quirks=quirks,
                    bc=bc, inflow=inflow, outflow=OUTFLOW, T_wall=T_WALL)
    g.build_metric(xe, ye, ze)
    return o, g, met, (xe, ye, ze)


def _stream(o, dims, coords, inflow, amp=0.3):
    """inflow state plus smooth perturbations on every node"""
    x, y, z = cases.interior_coords(dims, o.c.S, (o.c.n[0], o.c.n[1], o.c.n[2]), *coords)
    tp = 2 * np.pi
    rho = inflow[0] * (1 + amp * np.sin(tp * x) * np.cos(tp * y))
    u = inflow[1] * (1 + amp * np.cos(tp * y + 0.3) * np.sin(tp * z + 0.1))
    v = amp * np.sin(tp * x) * np.sin(tp * z + 0.2)
    w = amp * np.cos(tp * x) * np.sin(tp * y) if dims == 3 else np.zeros_like(x)
    p = inflow[4] * (1 + amp * np.cos(tp * (x + y)))
    e = cases.energy(u, v, w, p, rho, GAMMA, MACH)
    return [np.ascontiguousarray(a) for a in (rho, u, v, w, e)]


CASES = [
    # dims, S, n, bc, inviscid, quirks, mapping
    (2, 5, (24, 20, 1), INOUT + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, None),
    (2, 9, (24, 20, 1), INOUT + INOUT + (0, 0), False, orc.QUIRKS_REFERENCE, None),
    (3, 5, (14, 12, 12), INOUT + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    (3, 9, (14, 12, 12), INOUT + (0, 0, 0, 0), True, orc.QUIRKS_REFERENCE, "bump"),
    (3, 5, (12, 12, 12), INOUT + INOUT + INOUT, False, 0, "bump"),
    # subsonic outflow writers (stored p and c on their ghosts)
    (2, 5, (24, 20, 1), IN_NORMAL + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, None),
    (2, 9, (24, 20, 1), IN_RUDY + IN_NORMAL + (0, 0), True, orc.QUIRKS_REFERENCE, None),
    (3, 5, (14, 12, 12), IN_RUDY + (0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    (3, 5, (12, 12, 12), IN_NORMAL + IN_RUDY + INOUT, False, 0, "bump"),
    # walls (mirrored ghost metric, corrector on the ghosts' convective momentum terms)
    (2, 5, (24, 20, 1), INOUT + SLIP + (0, 0), False, orc.QUIRKS_REFERENCE, None),
    (2, 9, (24, 20, 1), INOUT + NOSLIP + (0, 0), False, orc.QUIRKS_REFERENCE, None),
    (2, 5, (20, 18, 1), SLIP + SLIP + (0, 0), True, orc.QUIRKS_REFERENCE, None),
    (3, 5, (14, 12, 12), INOUT + SLIP + (0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    (3, 9, (14, 12, 12), INOUT + NOSLIP + (0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    (3, 5, (12, 12, 12), INOUT + (orc.BC_WALL_NOSLIP, orc.BC_WALL_SLIP, orc.BC_WALL_SLIP, orc.BC_WALL_NOSLIP),
     False, 0, "bump"),
    # subsonic inflow writers
    (2, 5, (24, 20, 1), (orc.BC_INFLOW_SUB_NORMAL, orc.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, None),
    (2, 9, (24, 20, 1), (orc.BC_INFLOW_SUB_RIEMANN, orc.BC_OUTFLOW_NORMAL, 0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, None),
    (3, 5, (14, 12, 12), (orc.BC_INFLOW_SUB_ISENTROPIC, orc.BC_OUTFLOW_RUDY, 0, 0, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    # farfield faces
    (2, 5, (24, 20, 1), INOUT + (orc.BC_FARFIELD, orc.BC_FARFIELD, 0, 0), False, orc.QUIRKS_REFERENCE, None),
    (3, 9, (14, 12, 12), INOUT + (orc.BC_FARFIELD, orc.BC_FARFIELD, 0, 0), False, orc.QUIRKS_REFERENCE, "bump"),
    # isothermal walls (T_wall on the ghosts; the heat-conduction factor of Prandtl 0.72 against
    # the case's 0.9 below the upper wall)
    (2, 5, (24, 20, 1), INOUT + (orc.BC_WALL_ISOTHERMAL, orc.BC_WALL_ISOTHERMAL, 0, 0), False,
     orc.QUIRKS_REFERENCE, None),
    (3, 5, (14, 12, 12), INOUT + (orc.BC_WALL_ISOTHERMAL, orc.BC_WALL_ISOTHERMAL, 0, 0), False,
     orc.QUIRKS_REFERENCE, "bump"),
    (3, 5, (12, 12, 12), INOUT + (orc.BC_WALL_SLIP, orc.BC_WALL_ISOTHERMAL, orc.BC_WALL_ISOTHERMAL,
                                  orc.BC_WALL_NOSLIP), False, 0, "bump"),
    # (an isentropic inflow on a face along the stream is not well posed: its density formula
    # leaves the real numbers within ten steps, in the oracle as on the GPU)
    (3, 5, (12, 12, 12), (orc.BC_INFLOW_SUB_RIEMANN, orc.BC_OUTFLOW_RIEMANN, orc.BC_WALL_SLIP,
                          orc.BC_WALL_SLIP, orc.BC_INFLOW_SUB_NORMAL, orc.BC_WALL_NOSLIP), False, 0, "bump"),
]


def _mapping(dims, tag):
    return cases.sine_bump_3d(0.05) if tag == "bump" else cases.cartesian_unit(dims)


@pytest.mark.parametrize("dims,S,n,bc,inviscid,quirks,mp", CASES)
def test_inflow_outflow_rhs_and_steps(dims, S, n, bc, inviscid, quirks, mp):
    inflow = (1.0, 2.5, 0.0, 0.0, 1.0)
    # Prandtl 0.9: the isothermal ghosts' heat-conduction factor (Prandtl 0.72) differs from the interior's
    o, g, met, coords = _setup(dims, S, n, bc, inflow, inviscid, quirks, mapping=_mapping(dims, mp), prandtl=PRANDTL)
    tag = f"bc{bc[:2 * dims]} dims={dims} S={S} inviscid={inviscid} quirks={quirks}"
    with g:
        gm = g.download_metric()
        merr = [float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) for a, b in zip(gm, met)]
        _report("metric " + tag, merr)
        assert max(merr) <= 1e-14
        prim = _stream(o, dims, coords, inflow)
        q_ref, _, _ = o.rhs(met, [a.copy() for a in prim])
        g.upload_metric(met)
        g.upload_state(prim)
        q = g.rhs()
        sl = interior(o, dims, n)
        tau = 0.2 / (n[0] * 5.0)
        for steps, fixed in ((0, 1e-11), (1, 1e-12), (10, 1e-11)):
            if steps == 0:
                errs = rel_linf(q, q_ref, sl)
            else:
                g.upload_state(prim)
                g.run(tau, steps)
                errs = rel_linf(conserved(g.download_state()), conserved(o.run(tau, met, prim, steps)), sl)
            noise = _noise(o, dims, n, met, prim, steps, tau)
            what = "rhs" if steps == 0 else f"{steps} step(s)"
            _report(f"{what} {tag}", [e for i, e in enumerate(errs) if dims == 3 or i != 3])
            _report(f"{what} oracle 1-ulp spread {tag}", [e for i, e in enumerate(noise) if dims == 3 or i != 3])
            assert _bar(errs, fixed, noise, dims, steps, tag), what


@pytest.mark.parametrize("dims,S,inviscid,quirks", [(2, 5, False, orc.QUIRKS_REFERENCE),
                                                    (2, 9, True, orc.QUIRKS_REFERENCE),
                                                    (3, 5, False, orc.QUIRK_METRIC_DIAG3D),
                                                    (3, 5, False, orc.QUIRKS_REFERENCE)])
def test_manufactured_solution_path(dims, S, inviscid, quirks):
    n = (20, 20, 20) if dims == 3 else (28, 24, 1)
    bc = (orc.BC_MMS,) * 6
    o, g, met, coords = _setup(dims, S, n, bc, (1.0, 1.0, 0.0, 0.0, 1.0), inviscid, quirks, mms_on=True,
                               reynolds=100.0)
    x, y, z = cases.interior_coords(dims, S, n, *coords)
    P = mms.reference_params(dims, False, inviscid, mach=MACH)
    ex = mms.exact_state(P, x, y, z, GAMMA, MACH)
    src = mms.source(P, x, y, z, GAMMA, MACH, 100.0, 0.72, inviscid)
    tag = f"mms dims={dims} S={S} inviscid={inviscid} quirks={quirks}"
    sl = interior(o, dims, n)
    with g:
        with pytest.raises(shockx.ShxError):
            g.upload_state(ex[:5])
            g.rhs()                     # no boundary state yet
        g.upload_boundary_state(ex)
        g.upload_source(src)
        q = g.rhs()
        q_ref, _, _ = o.rhs(met, [a.copy() for a in ex[:5]], src, ex)
        errs = [float(np.max(np.abs(a[sl] - b[sl]))) / max(float(np.max(np.abs(b[sl]))), 1e-300)
                for a, b in zip(q, q_ref)]
        if dims == 2:
            errs = errs[:3] + errs[4:]
        _report("rhs " + tag, errs)
        # the residual of the exact field is small (truncation error): compare on the scale of
        # the individual flux contributions, i.e. the source
        scale = [max(float(np.max(np.abs(s[sl] * met[9][sl]))), 1e-300) for s in src]
        aerr = [float(np.max(np.abs(a[sl] - b[sl]))) / s for a, b, s in zip(q, q_ref, scale)]
        _report("rhs/source-scale " + tag, aerr)
        assert max(aerr) <= 1e-12
        tau = 0.1 / (n[0] * 4.0)
        g.upload_state(ex[:5])
        g.mms_error()                   # from here on the library keeps the stored pressure
        g.run(tau, 20)
        got = g.download_state()
        ref = o.run(tau, met, ex[:5], 20, src=src, bstate=ex)
        errs = rel_linf(conserved(got), conserved(ref), sl)
        if dims == 2:
            errs = errs[:3] + errs[4:]
        _report("20 steps " + tag, errs)
        assert max(errs) <= 1e-11
        # ErrorManufacturedSolution's norms (stored p one stage behind, last k plane)
        e_got = g.mms_error()
        e_want = o.mms_error(ref, o.last_p, ex[0], ex[5])
        print(f"[parity] MMS error {tag}: GPU {e_got} oracle {e_want}", flush=True)
        # the norms are of differences from the exact field, so the states' 20-step bar (1e-11
        # relative) carries over as an absolute bar on the scale of rho and p
        scale = [float(np.max(np.abs(ex[0][sl])))] * 2 + [float(np.max(np.abs(ex[5][sl])))] * 2
        assert all(abs(a - b) <= 1e-11 * s for a, b, s in zip(e_got, e_want, scale)), (e_got, e_want)


def test_create_rejects_unpaired_periodic_faces():
    with pytest.raises(shockx.ShxError):
        shockx.Case(dims=2, n=(16, 16, 1), bc=(0, orc.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0))
    with pytest.raises(shockx.ShxError):
        shockx.Case(dims=2, n=(16, 16, 1), bc=INOUT + (0, 0, 0, 0), inflow=(0.0, 1.0, 0.0, 0.0, 1.0))
