"""Pin the oracle's CalcValuesForPost (src/Functions.c:3687-3761) and CalcLambda2
(src/Functions.c:5846-6070).

  * Mach = |u| / sqrt(p / rho) * Ma on every interior node (closed form)
  * gradRho of a linear density on an affine grid with a different spacing per direction: the
    central differences and the metric are exact, so it equals |grad rho| (a metric entry paired
    with the wrong difference fails this)
  * lambda2 of a linear velocity field on a sheared affine grid (full metric): the interpolated
    differences are exact, so it equals the middle eigenvalue of S^2 + Omega^2 of the field's
    gradient, capped at 0, computed here with numpy's symmetric eigensolver.  The reference's S
    carries half the velocity derivative on its diagonal (s11 = 0.5 u_x, src/Functions.c:5952-5957),
    which the expected value reproduces; a rotation-free field (pure strain) gives 0 because S^2
    is positive semidefinite
  * ErrorManufacturedSolution's norms (src/ManufacturedSolution.c:745-790) against a numpy brute
    force over the same node set: the interior of the last k plane, squares divided by the full
    interior node count
"""
import numpy as np
import pytest

import orc
from helpers import interior
from shock_amd import cases

GAMMA, MACH = 1.4, 0.5


def _affine_case(dims, S, n, A):
    o = orc.Oracle(dims=dims, S=S, n=n, bc=(orc.BC_INFLOW_SUPERSONIC, orc.BC_OUTFLOW_RIEMANN) * dims + (0, 0) * (3 - dims),
                   mach=MACH, quirks=0)
    if dims == 2:
        mapping = lambda xi, eta, zeta: (A[0][0] * xi + A[0][1] * eta, A[1][0] * xi + A[1][1] * eta, np.zeros_like(xi))
    else:
        mapping = lambda xi, eta, zeta: tuple(A[r][0] * xi + A[r][1] * eta + A[r][2] * zeta for r in range(3))
    xe, ye, ze = cases.ext_coords(dims, S, n, mapping, periodic=(False, False, False))
    met = o.metric_from_coords(xe, ye, ze)
    x, y, z = cases.interior_coords(dims, S, n, xe, ye, ze)
    return o, met, np.ascontiguousarray(x), np.ascontiguousarray(y), np.ascontiguousarray(z)


@pytest.mark.parametrize("dims", [2, 3])
def test_mach_and_density_gradient(dims):
    n = (9, 8, 7) if dims == 3 else (9, 8, 1)
    A = np.diag([1.3, 0.7, 1.9])
    o, met, x, y, z = _affine_case(dims, 5, n, A)
    rng = np.random.default_rng(3)
    gx, gy, gz = 0.3, -0.2, (0.45 if dims == 3 else 0.0)
    rho = 1.0 + gx * x + gy * y + gz * z
    u, v = 0.3 + 0.1 * rng.random(o.shape), 0.1 * rng.random(o.shape)
    w = 0.1 * rng.random(o.shape) if dims == 3 else np.zeros(o.shape)
    p = 0.8 + 0.3 * rng.random(o.shape)
    prim = [np.ascontiguousarray(a) for a in (rho, u, v, w, cases.energy(u, v, w, p, rho, GAMMA, MACH))]
    out = o.calc_post(met, prim, p)
    sl = interior(o, dims, n)
    mach = np.sqrt(u * u + v * v + w * w) / np.sqrt(p / rho) * MACH
    assert np.allclose(out[0][sl], mach[sl], rtol=1e-15, atol=0)
    assert np.allclose(out[1][sl], np.sqrt(gx * gx + gy * gy + gz * gz), rtol=1e-12, atol=0)


@pytest.mark.parametrize("S", [5, 9])
def test_lambda2_of_a_linear_field(S):
    n = (10, 9, 8)
    A = np.array([[1.3, 0.2, 0.1], [0.15, 0.9, -0.05], [0.1, 0.05, 1.1]])
    o, met, x, y, z = _affine_case(3, S, n, A)
    G = np.array([[0.4, -1.1, 0.3], [0.9, -0.2, 0.5], [-0.4, 0.7, -0.3]])   # du_i/dx_j
    X = np.stack([x, y, z])
    vel = [0.2 + np.tensordot(G[q], X, axes=1) for q in range(3)]
    rho = np.ones(o.shape)
    p = np.ones(o.shape)
    prim = [rho] + [np.ascontiguousarray(a) for a in vel] + [cases.energy(*vel, p, rho, GAMMA, MACH)]
    out = o.calc_post(met, prim, p)
    # the reference's strain tensor halves its diagonal (s11 = 0.5 u_x, src/Functions.c:5952-5957)
    Sm, Om = 0.5 * (G + G.T), 0.5 * (G - G.T)
    Sm[np.diag_indices(3)] *= 0.5
    lam = np.linalg.eigvalsh(Sm @ Sm + Om @ Om)
    want = min(lam[1], 0.0)
    assert want < 0.0
    sl = interior(o, 3, n)
    assert np.allclose(out[2][sl], want, rtol=1e-10, atol=0)
    # pure strain: S^2 is positive semidefinite, lambda2 is capped at 0
    Gs = 0.5 * (G + G.T)
    vel = [0.2 + np.tensordot(Gs[q], X, axes=1) for q in range(3)]
    prim = [rho] + [np.ascontiguousarray(a) for a in vel] + [cases.energy(*vel, p, rho, GAMMA, MACH)]
    out = o.calc_post(met, prim, p)
    assert np.all(out[2][sl] == 0.0)


@pytest.mark.parametrize("dims", [2, 3])
def test_mms_error_norms_brute_force(dims):
    # ErrorManufacturedSolution: the last k plane only, squares divided by the full interior count
    n = (7, 6, 5) if dims == 3 else (7, 6, 1)
    o = orc.Oracle(dims=dims, S=5, n=n, mach=MACH)
    rng = np.random.default_rng(4)
    prim = [1.0 + rng.random(o.shape) for _ in range(5)]
    p, ex_r, ex_p = (1.0 + rng.random(o.shape) for _ in range(3))
    got = o.mms_error(prim, p, ex_r, ex_p)
    g = o.g
    k = g + n[2] - 1 if dims == 3 else 0
    sl = (slice(g, g + n[0]), slice(g, g + n[1]), k)
    N = n[0] * n[1] * n[2]
    dr, dp = ex_r[sl] - prim[0][sl], ex_p[sl] - p[sl]
    want = (np.sqrt(np.sum(dr * dr / N)), np.max(np.abs(dr)), np.sqrt(np.sum(dp * dp / N)), np.max(np.abs(dp)))
    assert np.allclose(got, want, rtol=1e-14, atol=0)
