"""Pin the oracle's scheme building blocks against what the mathematics fixes.

The reference has no tests or stored vectors and cannot be compiled here (it needs mpi.h and
cgnslib.h), so each oracle function is checked against exact rational solutions of textbook
accuracy conditions (tests/golden/linear_schemes.json, made by tests/golden/make_golden.py),
exact algebraic identities, or closed forms.  Each check is chosen so that a plausible slip
(a dropped term, a wrong sign or index, a mirrored stencil) fails.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import orc

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "linear_schemes.json")))


def frac(v):
    return np.array([float(Fraction(x)) for x in v])


def cfg(S):
    return orc.Oracle(dims=2, S=S, n=(S + 3, S + 3, 1)).c


# ---------------------------------------------------------------- DefineParameters constants
@pytest.mark.parametrize("S,w", [(5, 3), (9, 5)])
def test_nodal_derivative_is_central_difference(S, w):
    """src/Functions.c:1887-1894 / 1936-1942: the nodal derivative is the (2w)th-order central difference."""
    c = cfg(S)
    got = np.array(c.zd_der[: 2 * w + 1])
    assert np.allclose(got, frac(GOLD["central_d1"][str(w)]), rtol=0, atol=1e-16)


@pytest.mark.parametrize("S,w", [(5, 3), (9, 5)])
def test_midpoint_interpolation_differences_give_central_first_derivative(S, w):
    """I(i+1/2) - I(i-1/2) with the midpoint-interpolation weights equals the central first derivative."""
    c = cfg(S)
    I = np.array(c.zd_int[: 2 * w])
    comb = np.zeros(2 * w + 1)
    comb[1:] += I        # plus-half stencil: offsets -w+1 .. w
    comb[:-1] -= I       # minus-half stencil: offsets -w .. w-1
    assert np.allclose(comb, frac(GOLD["central_d1"][str(w)]), rtol=0, atol=1e-16)
    assert abs(I.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("S,w", [(5, 3), (9, 5)])
def test_midpoint_derivative_differences_give_central_second_derivative(S, w):
    """D(i+1/2) - D(i-1/2) with the midpoint-derivative weights is the central second derivative
    (6th order for S=5, the [2,-27,270,-490,270,-27,2]/180 of the reference's notes; 10th for S=9)."""
    c = cfg(S)
    D = np.array(c.zd_mid[: 2 * w])
    comb = np.zeros(2 * w + 1)
    comb[1:] += D
    comb[:-1] -= D
    assert np.allclose(comb, frac(GOLD["central_d2"][str(w)]), rtol=0, atol=2e-15)


def test_rk_tableaux_reproduce_stability_polynomials():
    """RK factors (src/Functions.c:1961-2009) applied to dU/dt = -k U give the textbook RK3/RK4 polynomials."""
    for order in (3, 4):
        o = orc.Oracle(dims=2, S=5, n=(8, 8, 1), time_order=order)
        met = o.zeros(10)
        met[9][...] = 1.0
        un = o.zeros(5)
        un[0][...] = 2.0
        un[4][...] = 1.0
        urk = [a.copy() for a in un]
        qsum = o.zeros(5)
        tau, kap = 0.3, 1.7
        for s in range(order):
            q = o.zeros(5)
            q[0][...] = -kap * urk[0]
            o.rk_update(s, tau, met, un, urk, q, qsum)
        z = kap * tau
        poly = 1 - z + z ** 2 / 2 - z ** 3 / 6 + (z ** 4 / 24 if order == 4 else 0.0)
        g = o.g
        assert abs(urk[0][g, g, 0] - 2.0 * poly) < 1e-14


# ---------------------------------------------------------------- WENO building blocks
@pytest.mark.parametrize("S,r", [(5, 2), (9, 4)])
def test_linear_weights_reduce_to_upwind_flux(S, r):
    """With eps -> inf the nonlinear weights become the optimal ones, and the reference's split
    form -Phi(f+) + Theta(f) must equal the (2r+1)-point upwind-biased linear flux, and
    +Phi(f-) + Theta(f) its mirror image (Phi_Function_W5/W9 + Theta_Function_W5/W9,
    src/WENO.c:11-204, 1754-1816)."""
    rng = np.random.default_rng(3)
    up = frac(GOLD["upwind"][str(r)])
    theta5 = np.array([-1, 7, 7, -1]) / 12.0
    theta9 = np.array([-3, 29, -139, 533, 533, -139, 29, -3]) / 840.0
    for _ in range(20):
        f = rng.normal(size=2 * r + 2)            # nodes m-r .. m+r+1
        fp = f[: 2 * r + 1]
        fm = f[1:][::-1]                          # mirrored stencil m+r+1 .. m-r+1
        dp = np.diff(fp)
        dm = fm[:-1] - fm[1:]
        phi = orc.phi_w5 if S == 5 else orc.phi_w9
        th = theta5 @ f[r - 1: r + 3] if S == 5 else theta9 @ f[r - 3: r + 5]
        plus_only = -phi(1e100, fp, dp) + th
        assert abs(plus_only - up @ fp) < 1e-13
        minus_only = phi(1e100, fm, dm) + th
        assert abs(minus_only - up @ fm) < 1e-13


@pytest.mark.parametrize("S,r", [(5, 2), (9, 4)])
def test_smoothness_indicators_are_jiang_shu_integrals(S, r):
    """IS_k of the oracle equal a common constant times sum_l int (p_k^(l))^2 of the sub-stencil
    reconstruction (12x for the difference form of WENO5, 5040x for the integer forms of WENO9)."""
    forms = [np.array([[float(Fraction(x)) for x in row] for row in A]) for A in GOLD["is_forms"][str(r)]]
    scale = 12.0 if S == 5 else 5040.0
    rng = np.random.default_rng(4)
    for _ in range(20):
        f = rng.normal(size=2 * r + 1)
        ref = np.array([f @ A @ f for A in forms])
        got = orc.is_w5(np.diff(f)) if S == 5 else orc.is_w9(f)
        assert np.allclose(got, scale * ref, rtol=1e-12, atol=1e-12)


def test_weno5_weights_pick_smooth_stencil_at_a_jump():
    """ENO property: a jump inside the upwind stencil leaves only the smooth sub-stencil."""
    f = np.array([0.0, 0.0, 0.0, 1.0, 1.0])       # jump between nodes 2 and 3
    d = np.diff(f)
    is_ = orc.is_w5(d)
    assert is_[0] == 0.0 and is_[1] > 0 and is_[2] > 0
    # with the reference epsilon the flux equals the one-sided stencil-0 reconstruction
    # (2 f0 - 7 f1 + 11 f2)/6 up to round-off
    th = np.array([-1, 7, 7, -1]) / 12.0 @ np.array([f[1], f[2], f[3], f[4]])
    val = -orc.phi_w5(orc.Oracle().c.eps, f, d) + th
    assert abs(val - (2 * f[0] - 7 * f[1] + 11 * f[2]) / 6) < 1e-12


def euler_flux(U, n, gamma):
    rho, mx, my, mz, E = U
    u = np.array([mx, my, mz]) / rho
    P = (gamma - 1) * (E - 0.5 * rho * (u @ u))
    un = u @ n
    return np.array([rho * un, mx * un + P * n[0], my * un + P * n[1], mz * un + P * n[2], (E + P) * un])


def test_eigenvectors_diagonalise_the_flux_jacobian():
    """CalcEigenVectorsIn*Direction (src/WENO.c:1085-1548): L R = I and R diag(un, un, un, un+c,
    un-c) L = dF/dU for the Euler flux in direction n (pressure Upsilon*p, c^2 = gamma P/rho)."""
    rng = np.random.default_rng(5)
    gamma = 1.4
    for _ in range(10):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        rho = rng.uniform(0.5, 2.0)
        vel = rng.normal(size=3)
        P = rng.uniform(0.5, 2.0)
        c = np.sqrt(gamma * P / rho)
        L, R = orc.eigvec(n, vel[0], vel[1], vel[2], rho, c, gamma)
        assert np.allclose(L @ R, np.eye(5), atol=1e-12)
        E = P / (gamma - 1) + 0.5 * rho * vel @ vel
        U = np.array([rho, *(rho * vel), E])
        Jm = np.zeros((5, 5))
        for m in range(5):
            h = 1e-6 * max(1.0, abs(U[m]))
            e = np.zeros(5)
            e[m] = h
            Jm[:, m] = (euler_flux(U + e, n, gamma) - euler_flux(U - e, n, gamma)) / (2 * h)
        un = vel @ n
        lam = np.diag([un, un, un, un + c, un - c])
        assert np.allclose(R @ lam @ L, Jm, atol=1e-7)


def test_viscous_factors_are_the_newtonian_stress_contraction():
    """CreateViscidMetric (src/Functions.c:6245-6386): factor of d(u_k)/d(xi_b) in component i of
    the direction-a flux = J sum_j a_j (d_ik r_bj + d_jk r_bi - 2/3 d_ij r_bk); heat = J r_b . a."""
    rng = np.random.default_rng(6)
    for d in range(3):
        m = rng.normal(size=10)
        fac = orc.visc_fac(d, m)
        rows = m[:9].reshape(3, 3)
        a = rows[d]
        J = m[9]
        for i in range(3):
            for k in range(3):
                for b in range(3):
                    r = rows[b]
                    ref = J * sum(a[j] * ((i == k) * r[j] + (j == k) * r[i] - 2.0 / 3.0 * (i == j) * r[k])
                                  for j in range(3))
                    assert abs(fac[i * 9 + k * 3 + b] - ref) < 1e-12 * max(1, abs(ref))
        for b in range(3):
            assert abs(fac[27 + b] - J * rows[b] @ a) < 1e-12
