"""GPU parity: libshockx.so (HIP, gfx950) against the CPU oracle on identical inputs.

Bar (FP64 form): relative L-infinity per conserved variable <= 1e-12 after one RK step and
<= 1e-10 after 100 steps; metric and stage residuals within a few ulps of the oracle.
"""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, setup_2d_mms, setup_3d_turb, smooth_periodic_state
from shock_amd import cases, shockx

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-12
TOL_100 = 1e-10


def _case(o, dims, S, time_order, nn, inviscid=False, quirks=orc.QUIRKS_REFERENCE, eps_override=0.0,
          reynolds=None, mach=None):
    c = o.c
    return shockx.Case(dims=dims, space_order=S, time_order=time_order, n=nn, gamma=c.gamma,
                       mach=c.mach if mach is None else mach, reynolds=c.reynolds if reynolds is None else reynolds,
                       prandtl=c.prandtl, T0=c.T0, inviscid=inviscid, quirks=quirks,
                       eps_override=eps_override)


def _report(tag, errs):
    print(f"[parity] {tag}: " + " ".join(f"{e:.2e}" for e in errs), flush=True)


@pytest.mark.parametrize("S", [5, 9])
def test_metric_2d_and_3d(S):
    o, (xe, ye, ze), met, prim, nn = setup_2d_mms(n=24, S=S)
    with _case(o, 2, S, 4, nn) as g:
        g.build_metric(xe, ye, ze)
        gm = g.download_metric()
    errs = [float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) for a, b in zip(gm, met)]
    _report(f"metric2d S={S}", errs)
    assert max(errs) <= 1e-14
    o3, (xe, ye, ze), met3, prim3, nn3 = setup_3d_turb(n=12, S=S)
    with _case(o3, 3, S, 4, nn3) as g:
        g.build_metric(xe, ye, ze)
        gm = g.download_metric()
    errs = [float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) for a, b in zip(gm, met3)]
    _report(f"metric3d S={S}", errs)
    assert max(errs) <= 1e-14


@pytest.mark.parametrize("S,inviscid", [(5, False), (5, True), (9, False)])
def test_rhs_2d(S, inviscid):
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=S, inviscid=inviscid)
    q_ref, _, _ = o.rhs(met, [a.copy() for a in prim])
    with _case(o, 2, S, 4, nn, inviscid=inviscid) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        q = g.rhs()
    sl = interior(o, 2, nn)
    errs = rel_linf(q, q_ref, sl)
    _report(f"rhs2d S={S} inviscid={inviscid}", errs)
    assert max(errs[:3] + errs[4:]) <= 1e-12


@pytest.mark.parametrize("S,quirks", [(5, orc.QUIRKS_REFERENCE), (5, 0), (9, orc.QUIRKS_REFERENCE), (9, 0)])
def test_rhs_3d(S, quirks):
    o, coords, met, prim, nn = setup_3d_turb(n=12, S=S, quirks=quirks)
    q_ref, _, _ = o.rhs(met, [a.copy() for a in prim])
    with _case(o, 3, S, 4, nn, quirks=quirks) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        q = g.rhs()
    sl = interior(o, 3, nn)
    errs = rel_linf(q, q_ref, sl)
    _report(f"rhs3d S={S} quirks={quirks}", errs)
    assert max(errs) <= 1e-12


@pytest.mark.parametrize("S,time_order", [(5, 4), (5, 3), (9, 4)])
def test_one_step_2d(S, time_order):
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=S, time_order=time_order)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 1)
    with _case(o, 2, S, time_order, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.calc_runge_kutta(tau)
        out = g.download_state()
    sl = interior(o, 2, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"step2d S={S} rk={time_order}", errs)
    assert max(errs[:3] + errs[4:]) <= TOL_STEP


# 3D viscous cases with the reference's diagonal metric run the zeta direction as one fused
# kernel (WENO, viscous terms and the stage update), for both RK orders
@pytest.mark.parametrize("S,time_order", [(5, 4), (9, 4), (5, 3), (9, 3)])
def test_one_step_3d(S, time_order):
    o, coords, met, prim, nn = setup_3d_turb(n=12, S=S, time_order=time_order)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 1)
    with _case(o, 3, S, time_order, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.calc_runge_kutta(tau)
        out = g.download_state()
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"step3d S={S} rk={time_order}", errs)
    assert max(errs) <= TOL_STEP


@pytest.mark.parametrize("S", [5, 9])
def test_ten_steps_3d(S):
    o, coords, met, prim, nn = setup_3d_turb(n=14, S=S, seed=19)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 10)
    with _case(o, 3, S, 4, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.run(tau, 10)
        out = g.download_state()
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"10 steps 3d S={S}", errs)
    assert max(errs) <= 1e-11


def test_hundred_steps_2d_64_mms():
    """2D 64x64 periodic manufactured field, WENO5 + viscous, RK4, fixed tau, 100 steps."""
    o, coords, met, prim, nn = setup_2d_mms(n=64, S=5)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 100)
    with _case(o, 2, 5, 4, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.run(tau, 100)
        out = g.download_state()
    sl = interior(o, 2, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report("2D 64x64 periodic MMS field, 100 steps", errs)
    assert max(errs[:3] + errs[4:]) <= TOL_100


@pytest.mark.parametrize("S", [5, 9])
def test_hundred_steps_3d_turbulence(S):
    """16^3 periodic curvilinear box with random-phase turbulence, viscous, RK4, 100 steps (the
    oracle's own spread under one ulp of input noise: 1e-15 for W5, 1.3e-13 for W9)."""
    o, coords, met, prim, nn = setup_3d_turb(n=16, S=S, seed=23)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 100)
    with _case(o, 3, S, 4, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.run(tau, 100)
        out = g.download_state()
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"100 steps 3d turbulence S={S}", errs)
    assert max(errs) <= TOL_100


def test_hundred_steps_3d_shock():
    """16^3 Mach-3 normal shock meeting a vorticity field on a perturbed curvilinear box, WENO9,
    Re 5000, supersonic inflow and Riemann outflow in x, periodic in y and z, 100 steps (the
    oracle's own spread under one ulp of input noise: up to 3e-11)."""
    nn = (16, 16, 16)
    coords, prim, inflow = cases.shock_3d(9, nn, 5)
    bc = (orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0)
    o = orc.Oracle(dims=3, S=9, n=nn, bc=bc, inflow=inflow, mach=3.0, reynolds=5000.0)
    met = o.metric_from_coords(*coords)
    o.ghost_periodic(met)
    o.ghost_periodic(prim)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 100)
    with shockx.Case(dims=3, space_order=9, time_order=4, n=nn, gamma=1.4, mach=3.0, reynolds=5000.0,
                     bc=bc, inflow=inflow) as g:
        g.build_metric(*coords)
        g.upload_state(prim)
        g.run(tau, 100)
        out = g.download_state()
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report("100 steps 3d shock S=9", errs)
    assert max(errs) <= TOL_100


def test_check_metric_on_a_folded_grid():
    # x = xi + a sin(2 pi xi) with a > 1/(2 pi) folds the grid: the Jacobian is negative on a band
    nn = (24, 20, 1)
    o = orc.Oracle(dims=2, S=5, n=nn, bc=(orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0))
    fold = lambda xi, eta, zeta: (xi + 0.25 * np.sin(2 * np.pi * xi), eta, zeta)
    xe, ye, ze = cases.ext_coords(2, 5, nn, fold, periodic=(False, True, False))
    met = o.metric_from_coords(xe, ye, ze)
    want = o.check_metric(met)
    assert 0 < want < o.c.ntot
    with shockx.Case(dims=2, space_order=5, n=nn, bc=(shockx.BC_INFLOW_SUPERSONIC, shockx.BC_OUTFLOW_RIEMANN, 0, 0, 0, 0)) as g:
        g.build_metric(xe, ye, ze)
        got = g.check_metric()
    print(f"[parity] negative Jacobians: GPU {got}, oracle {want}", flush=True)
    assert got == want


def test_cfl_and_nan():
    o, coords, met, prim, nn = setup_3d_turb(n=12, S=5)
    with _case(o, 3, 5, 4, nn) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        dt = g.cfl_dt(0.5)
        assert g.check_nan() == 0
        bad = [a.copy() for a in prim]
        bad[0][o.g + 2, o.g + 3, o.g + 4] = np.nan
        bad[0][o.g + 5, o.g + 1, o.g + 7] = np.inf
        g.upload_state(bad)
        cnt = g.check_nan()
        fixed = g.download_state()
    ref_dt = o.cfl_dt(0.5, met, prim)
    assert abs(dt - ref_dt) <= 1e-15 * ref_dt      # max is order independent: exact up to sqrt rounding
    ref_bad = [a.copy() for a in bad]
    assert cnt == o.check_nan(ref_bad) == 2
    assert fixed[0][o.g + 2, o.g + 3, o.g + 4] == 1e6 and fixed[1][o.g + 5, o.g + 1, o.g + 7] == 1e6


@pytest.mark.parametrize("S,quirks", [(5, orc.QUIRKS_REFERENCE), (5, 0), (9, orc.QUIRKS_REFERENCE), (9, 0)])
def test_one_step_across_march_chunks_and_columns(S, quirks):
    """40 x 12 x 64 block: the derivative march splits i into two chunks of planes and k into two
    columns, so the chunk-seam window priming, the cross-column halo and (W5) the prefetch at a
    seam all enter the step; with the reference's diagonal metric and with the full metric (T
    derivatives marched too)."""
    nn = (40, 12, 64)
    o = orc.Oracle(dims=3, S=S, n=nn, quirks=quirks, mach=0.5, reynolds=2000.0)
    xe, ye, ze = cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(0.05))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    prim = smooth_periodic_state(o, nn, 31 + S, 0.5)
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 1)
    with _case(o, 3, S, 4, nn, quirks=quirks) as g:
        g.build_metric(xe, ye, ze)
        g.upload_state(prim)
        g.calc_runge_kutta(tau)
        out = g.download_state()
    sl = interior(o, 3, nn)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"step3d 40x12x64 S={S} quirks={quirks}", errs)
    assert max(errs) <= TOL_STEP


@pytest.mark.parametrize("S", [5, 9])
def test_zeta_quirk_with_eta_longer_than_zeta(S):
    """14 x 20 x 12 block with the reference's quirks: the zeta viscous stencil takes its k offset
    from j (src/ZD.c:547-548), and with n_eta > n_zeta the offsets run past the end of the k row, so
    the reference reads the next row's first nodes (its unpadded k-fastest index); the oracle indexes
    the same way.  Stage residual and one step against the oracle."""
    nn = (14, 20, 12)
    o = orc.Oracle(dims=3, S=S, n=nn, quirks=orc.QUIRKS_REFERENCE, mach=0.5, reynolds=2000.0)
    assert o.shape[1] > o.shape[2]
    xe, ye, ze = cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(0.05))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    prim = smooth_periodic_state(o, nn, 17 + S, 0.5)
    q_ref, _, _ = o.rhs(met, [a.copy() for a in prim])
    tau = o.cfl_dt(0.5, met, prim)
    ref = o.run(tau, met, prim, 1)
    sl = interior(o, 3, nn)
    with _case(o, 3, S, 4, nn) as g:
        g.build_metric(xe, ye, ze)
        g.upload_state(prim)
        q = g.rhs()
        g.calc_runge_kutta(tau)
        out = g.download_state()
    errs = rel_linf(q, q_ref, sl)
    _report(f"rhs3d 14x20x12 zeta quirk S={S}", errs)
    # the residual is a difference of fluxes: WENO9 with the reference's epsilon moves it by up
    # to ~5e-12 under rounding; a wrong row mapping would move it by O(1)
    assert max(errs) <= (1e-12 if S == 5 else 1e-10)
    errs = rel_linf(conserved(out), conserved(ref), sl)
    _report(f"step3d 14x20x12 zeta quirk S={S}", errs)
    assert max(errs) <= TOL_STEP
