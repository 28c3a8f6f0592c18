"""Pin the oracle's assembled operators: invariants, closed forms and orders of accuracy.

Checks (CPU only):
  * affine grids: CreateMetric (src/Functions.c:6388-6509) returns the exact inverse Jacobian
  * periodic ghost fill is the wrap-around permutation
  * free-stream preservation on Cartesian grids (uniform state -> Q = 0)
  * telescoping conservation of mass and momentum on periodic blocks (src/WENO.c:469-485)
  * mirror symmetry of the upwind-biased WENO flux (plus/minus stencils, src/WENO.c:334-385)
  * order of accuracy of the inviscid and viscous operators against analytic derivatives
  * CFL reduction and checkNAN against brute force
"""
import numpy as np
import pytest

import orc
from helpers import interior, setup_2d_mms, setup_3d_turb
from shock_amd import cases


def _uniform(o, rho=1.2, u=0.7, v=-0.3, w=0.2, p=1.1, gamma=1.4, mach=0.5):
    prim = o.zeros(5)
    prim[0][...] = rho
    prim[1][...] = u
    prim[2][...] = v
    prim[3][...] = w if o.c.dims == 3 else 0.0
    prim[4][...] = cases.energy(prim[1], prim[2], prim[3], p, prim[0], gamma, mach)
    return prim


def _metric(o, dims, S, n, mapping):
    xe, ye, ze = cases.periodic_ext_coords(dims, S, n, mapping)
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    return met, (xe, ye, ze)


def test_metric_exact_on_affine_grids():
    # 2D full affine map, 3D diagonal map (the reference's 3D metric keeps only the diagonal)
    A2 = np.array([[1.3, 0.4], [-0.2, 0.9]])
    n = (12, 10, 1)
    o = orc.Oracle(dims=2, S=5, n=n)
    xe, ye, ze = cases.periodic_ext_coords(2, 5, n, lambda a, b, c: (A2[0, 0] * a + A2[0, 1] * b,
                                                                     A2[1, 0] * a + A2[1, 1] * b, 0 * a),
                                           period=(1.0, 1.0, 1.0))
    # exact affine extension instead of the periodic shift (the map is not periodic in this test)
    I, J = np.meshgrid(np.arange(xe.shape[0]) - 1 - o.g, np.arange(xe.shape[1]) - 1 - o.g, indexing="ij")
    xi, eta = I / n[0], J / n[1]
    xe = np.ascontiguousarray((A2[0, 0] * xi + A2[0, 1] * eta)[..., None])
    ye = np.ascontiguousarray((A2[1, 0] * xi + A2[1, 1] * eta)[..., None])
    met = o.metric_from_coords(xe, ye, np.zeros_like(xe))
    Jac = np.array([[A2[0, 0] / n[0], A2[0, 1] / n[1]], [A2[1, 0] / n[0], A2[1, 1] / n[1]]])
    inv = np.linalg.inv(Jac)
    assert np.allclose(met[9], np.linalg.det(Jac), rtol=1e-13)
    for (r, cidx), mi in zip([(0, 0), (0, 1), (1, 0), (1, 1)], [0, 1, 3, 4]):
        assert np.allclose(met[mi], inv[r, cidx], rtol=1e-12, atol=1e-12)
    assert np.allclose(met[8], 1.0)
    # 3D diagonal affine
    n3 = (8, 9, 10)
    o3 = orc.Oracle(dims=3, S=5, n=n3)
    sc = (1.5, 0.8, 2.0)
    xe, ye, ze = cases.periodic_ext_coords(3, 5, n3, lambda a, b, c: (sc[0] * a, sc[1] * b, sc[2] * c), period=sc)
    met3 = o3.metric_from_coords(xe, ye, ze)
    for d, mi in enumerate([0, 4, 8]):
        assert np.allclose(met3[mi], n3[d] / sc[d], rtol=1e-12)
    assert np.allclose(met3[9], np.prod(sc) / np.prod(n3), rtol=1e-12)
    for mi in [1, 2, 3, 5, 6, 7]:
        assert np.all(met3[mi] == 0.0)


def test_periodic_ghost_fill_is_wraparound():
    n = (7, 8, 9)
    o = orc.Oracle(dims=3, S=5, n=n)
    f = o.zeros(1)
    rng = np.random.default_rng(1)
    g = o.g
    f[0][g:g + 7, g:g + 8, g:g + 9] = rng.normal(size=(7, 8, 9))
    o.ghost_periodic(f)
    core = f[0][g:g + 7, g:g + 8, g:g + 9]
    tiled = np.tile(core, (3, 3, 3))
    ref = tiled[7 - g:7 - g + o.shape[0], 8 - g:8 - g + o.shape[1], 9 - g:9 - g + o.shape[2]]
    assert np.array_equal(f[0], ref)


@pytest.mark.parametrize("dims,S", [(2, 5), (2, 9), (3, 5), (3, 9)])
def test_free_stream_preserved_on_cartesian_grids(dims, S):
    n = (12, 11, 1) if dims == 2 else (11, 11, 12)
    o = orc.Oracle(dims=dims, S=S, n=n)
    met, _ = _metric(o, dims, S, n, cases.cartesian_unit(dims))
    prim = _uniform(o)
    q, _, _ = o.rhs(met, prim)
    sl = interior(o, dims, n)
    scale = 1.0 / float(np.max(met[9]))          # fluxes carry J
    for m in range(5):
        assert np.max(np.abs(q[m][sl])) * scale < 1e-12


@pytest.mark.parametrize("dims,S,quirks", [(2, 5, orc.QUIRKS_REFERENCE), (2, 9, orc.QUIRKS_REFERENCE),
                                           (3, 5, orc.QUIRK_METRIC_DIAG3D), (3, 5, 0)])
def test_mass_and_momentum_telescope_on_periodic_blocks(dims, S, quirks):
    if dims == 2:
        n = (16, 16, 1)
        o = orc.Oracle(dims=2, S=S, n=n, quirks=quirks)
        tp = 2 * np.pi
        met, coords = _metric(o, 2, S, n, lambda a, b, c: (a + 0.03 * np.sin(tp * b), b + 0.02 * np.sin(tp * a), 0 * a))
        x, y, z = cases.interior_coords(2, S, n, *coords)
        prim = cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y))
    else:
        o, coords, met, prim, n = setup_3d_turb(n=12, S=S, quirks=quirks)
    q, _, _ = o.rhs(met, prim)
    sl = interior(o, dims, n)
    for m in range(dims + 1 if dims == 3 else 3):
        tot = np.sum(q[m][sl])
        mag = np.sum(np.abs(q[m][sl]))
        assert abs(tot) <= 1e-12 * mag, (m, tot, mag)


def test_inviscid_energy_telescopes():
    o, coords, met, prim, n = setup_2d_mms(n=16, S=5, inviscid=True)
    q, _, _ = o.rhs(met, prim)
    sl = interior(o, 2, n)
    assert abs(np.sum(q[4][sl])) <= 1e-12 * np.sum(np.abs(q[4][sl]))


def test_weno_flux_mirror_symmetry():
    """Reflecting the data in x (and u -> -u) reflects Q (and Q_x -> -Q_x)."""
    o, coords, met, prim, n = setup_2d_mms(n=24, S=5, inviscid=True)
    q, _, _ = o.rhs(met, [a.copy() for a in prim])
    g = o.g
    sl = interior(o, 2, n)
    refl = []
    for m, a in enumerate(prim):
        b = a.copy()
        b[sl] = a[sl][::-1, :, :]
        if m == 1:
            b[sl] = -b[sl]
        refl.append(b)
    q2, _, _ = o.rhs(met, refl)
    for m in (0, 1, 2, 4):
        expect = q[m][sl][::-1, :, :] * (-1.0 if m == 1 else 1.0)
        err = np.max(np.abs(q2[m][sl] - expect)) / np.max(np.abs(expect))
        assert err < 1e-12, (m, err)


def _wave_error(S, N):
    n = (N, 8, 1)
    o = orc.Oracle(dims=2, S=S, n=n, inviscid=True)
    met, coords = _metric(o, 2, S, n, cases.cartesian_unit(2))
    x, y, z = cases.interior_coords(2, S, n, *coords)
    tp = 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(tp * x)
    u = np.full_like(x, 0.8)
    v = np.zeros_like(x)
    p = np.ones_like(x)
    prim = [np.ascontiguousarray(a) for a in (rho, u, v, v, cases.energy(u, v, v, p, rho, 1.4, 0.5))]
    q, _, _ = o.rhs(met, prim)
    sl = interior(o, 2, n)
    exact = -met[9] * 0.8 * 0.2 * tp * np.cos(tp * x)
    return np.mean(np.abs(q[0][sl] - exact[sl])) / np.mean(np.abs(exact[sl]))


@pytest.mark.parametrize("S,min_order", [(5, 4.5), (9, 7.5)])
def test_weno_order_of_accuracy(S, min_order):
    e1, e2 = _wave_error(S, 24), _wave_error(S, 48)
    order = np.log2(e1 / e2)
    assert order > min_order, (e1, e2, order)


def _shear_error(S, N, dims=2, quirks=0):
    """viscous x-momentum term of u = 0.1 sin(2 pi y) (+ 0.1 sin(2 pi z) in 3D) with mu = 1:
    Q_x = J Psi mu (u_yy + u_zz)."""
    n = (8, N, 1) if dims == 2 else (10, N, N)
    o = orc.Oracle(dims=dims, S=S, n=n, mms=True, quirks=quirks, reynolds=100.0)
    oi = orc.Oracle(dims=dims, S=S, n=n, mms=True, quirks=quirks, reynolds=100.0, inviscid=True)
    met, coords = _metric(o, dims, S, n, cases.cartesian_unit(dims))
    x, y, z = cases.interior_coords(dims, S, n, *coords)
    tp = 2 * np.pi
    u = 0.1 * np.sin(tp * y) + (0.1 * np.sin(tp * z) if dims == 3 else 0.0)
    rho = np.ones_like(x)
    zero = np.zeros_like(x)
    prim = [np.ascontiguousarray(a) for a in (rho, u, zero, zero, cases.energy(u, zero, zero, 1.0, rho, 1.4, 0.5))]
    qv, _, _ = o.rhs(met, [a.copy() for a in prim])
    qi, _, _ = oi.rhs(met, [a.copy() for a in prim])
    sl = interior(o, dims, n)
    lap = -(tp ** 2) * (0.1 * np.sin(tp * y) + (0.1 * np.sin(tp * z) if dims == 3 else 0.0))
    exact = met[9] * o.c.psi * lap
    num = qv[1] - qi[1]
    return np.max(np.abs(num[sl] - exact[sl])) / np.max(np.abs(exact[sl]))


@pytest.mark.parametrize("S,min_order", [(5, 5.5), (9, 9.0)])
def test_viscous_order_of_accuracy(S, min_order):
    e1, e2 = _shear_error(S, 16), _shear_error(S, 32)
    order = np.log2(e1 / e2)
    assert order > min_order, (e1, e2, order)


def test_viscous_zeta_direction_consistent_without_quirk():
    e1, e2 = _shear_error(5, 16, dims=3), _shear_error(5, 32, dims=3)
    assert np.log2(e1 / e2) > 5.5


def test_zeta_index_quirk_changes_only_the_zeta_viscous_term():
    """With the quirk the zeta stencil samples the (i, j) line at k' around j (src/ZD.c:547-548):
    a field varying only in z then gives a wrong (k-independent face value) viscous term."""
    e_fixed = _shear_error(5, 16, dims=3, quirks=0)
    e_ref = _shear_error(5, 16, dims=3, quirks=orc.QUIRK_ZETA_VISC_J)
    assert e_fixed < 1e-3 and e_ref > 1e-2


def test_cfl_and_nan_against_brute_force():
    o, coords, met, prim, n = setup_3d_turb(n=10, S=5)
    dt = o.cfl_dt(0.5, met, prim)
    sl = interior(o, 3, n)
    rho, u, v, w, e = [a[sl] for a in prim]
    c = o.c
    p = np.abs(rho * (c.gamma - 1) / c.upsilon * (e - 0.5 * (u * u + v * v + w * w)))
    cs = np.sqrt(c.upsilon * c.gamma * p / rho)
    tot = np.zeros_like(rho)
    for d in range(3):
        a, b, z = met[3 * d][sl], met[3 * d + 1][sl], met[3 * d + 2][sl]
        tot += np.abs(u * a + v * b + w * z) + cs * np.sqrt(a * a + b * b + z * z)
    assert abs(dt - 0.5 / tot.max()) <= 1e-15 * dt
    bad = [a.copy() for a in prim]
    g = o.g
    bad[0][g + 1, g + 2, g + 3] = np.nan
    bad[0][g + 4, g + 4, g + 4] = -np.inf
    bad[0][0, 0, 0] = np.nan          # ghost nodes are not checked
    assert o.check_nan(bad) == 2
    assert bad[0][g + 1, g + 2, g + 3] == 1e6 and bad[3][g + 4, g + 4, g + 4] == 1e6


@pytest.mark.parametrize("dims", [2, 3])
def test_check_metric_counts_negative_jacobians(dims):
    # a right-handed grid has none; mirroring one axis makes every Jacobian negative
    n = (10, 9, 8) if dims == 3 else (10, 9, 1)
    o = orc.Oracle(dims=dims, S=5, n=n)
    ident = cases.cartesian_unit(dims)
    for mp, want in ((ident, 0), (lambda a, b, c: (-ident(a, b, c)[0],) + tuple(ident(a, b, c)[1:]), o.c.ntot)):
        xe, ye, ze = cases.ext_coords(dims, 5, n, mp, periodic=(False, False, False))
        met = o.metric_from_coords(xe, ye, ze)
        assert o.check_metric(met) == want


def test_rk_step_keeps_uniform_state():
    o = orc.Oracle(dims=2, S=5, n=(12, 12, 1))
    met, _ = _metric(o, 2, 5, (12, 12, 1), cases.cartesian_unit(2))
    prim = _uniform(o)
    out = o.run(1e-3, met, prim, 2)
    sl = interior(o, 2, (12, 12, 1))
    for m in (0, 1, 2, 4):
        assert np.allclose(out[m][sl], prim[m][sl], rtol=1e-13, atol=1e-13)
