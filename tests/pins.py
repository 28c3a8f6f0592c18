"""Pins of three result-changing choices of the reference's operators (checked on the oracle).

Each function returns one figure of merit; tests/test_oracle_pins.py asserts it against a bar
and, on altered copies of the oracle (the plausible alternative choice planted in the C source),
that the same figure fails the bar.

  free_stream_stretched  The Lax-Friedrichs split adds lambda * J_{i+1/2} * U_s on every stencil
                         node (src/WENO.c:314, 353-383): one J for the whole stencil.  A uniform
                         state on a separably stretched grid then gives Q = 0 to round-off.
  char_flux_1d           The eigen-system of an interface is taken at the arithmetic means of
                         u, v, w, rho and c of its two nodes (src/WENO.c:1104-1109).  For 1D data
                         (v = w = 0, Cartesian metric) the shear fields carry nothing, the three
                         remaining characteristic fields have distinct speeds, and WENO's Phi is
                         scale-equivariant, so the interface flux equals characteristic LF-WENO
                         built from numpy.linalg.eig of the 1D Euler Jacobian at that mean state
                         whatever the eigenvector normalisation.
  mu_locality            The viscous face coefficients are K = (mu_i fac_i + mu_{i+1} fac_{i+1})/2
                         (src/ZD.c:122-132): a change of mu at node m enters every flux through
                         mu_m fac_m only, so the response of Q to it does not depend on the metric
                         of any other node.
The density of the interface state (src/WENO.c:1107) is not observable: it scales the two shear
eigenvectors and their left rows inversely, and WENO's Phi is scale-equivariant.
"""
import numpy as np

import orc
from shock_amd import cases


def _uniform(o, rho=1.2, u=0.7, v=-0.3, w=0.2, p=1.1):
    prim = o.zeros(5)
    prim[0][...] = rho
    prim[1][...] = u
    prim[2][...] = v
    prim[3][...] = w if o.c.dims == 3 else 0.0
    prim[4][...] = cases.energy(prim[1], prim[2], prim[3], p, prim[0], o.c.gamma, o.c.mach)
    return prim


def _interior(o, n):
    g = o.g
    if o.c.dims == 2:
        return (slice(g, g + n[0]), slice(g, g + n[1]), slice(0, 1))
    return (slice(g, g + n[0]), slice(g, g + n[1]), slice(g, g + n[2]))


def free_stream_stretched(dims, S):
    """max over the interior and the five equations of |Q| / J for a uniform stream on the grid
    x = xi + 0.08 sin(2 pi xi), y = eta + 0.06 sin(2 pi eta) (, z = zeta + 0.05 sin(2 pi zeta))."""
    n = (12, 11, 1) if dims == 2 else (10, 11, 12)
    o = orc.Oracle(dims=dims, S=S, n=n, inviscid=True)
    tp = 2.0 * np.pi
    if dims == 2:
        mapping = lambda a, b, c: (a + 0.08 * np.sin(tp * a), b + 0.06 * np.sin(tp * b), 0 * a)
    else:
        mapping = lambda a, b, c: (a + 0.08 * np.sin(tp * a), b + 0.06 * np.sin(tp * b), c + 0.05 * np.sin(tp * c))
    xe, ye, ze = cases.periodic_ext_coords(dims, S, n, mapping)
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    prim = _uniform(o)
    q, _, _ = o.rhs(met, prim)
    sl = _interior(o, n)
    return max(float(np.max(np.abs(q[m][sl] / met[9][sl]))) for m in range(5))


def _char_flux_numpy(o, prim, der, i0, S):
    """Characteristic LF-WENO interface flux of 1D Euler data (rows i of the padded block, j = 0
    line of a Cartesian unit-metric block) at interface i0 + 1/2, in the variables
    (rho, rho u, rho e); eigenvectors of the flux Jacobian from numpy."""
    gam, ups = o.c.gamma, o.c.upsilon
    rho, u, e = prim[0][:, 0, 0], prim[1][:, 0, 0], prim[4][:, 0, 0]
    p, c = der[0][:, 0, 0], der[2][:, 0, 0]
    U = np.stack([rho, rho * u, rho * e], axis=1)
    F = np.stack([rho * u, rho * u * u + ups * p, rho * u * (e + ups * p / rho)], axis=1)
    r = (S - 1) // 2
    lam = max(abs(u[s]) + c[s] for s in range(i0 - r, i0 + r + 2))
    ub = 0.5 * (u[i0] + u[i0 + 1])
    cb = 0.5 * (c[i0] + c[i0 + 1])
    H = cb * cb / (gam - 1.0) + 0.5 * ub * ub
    A = np.array([[0.0, 1.0, 0.0],
                  [0.5 * (gam - 3.0) * ub * ub, (3.0 - gam) * ub, gam - 1.0],
                  [ub * (0.5 * (gam - 1.0) * ub * ub - H), H - (gam - 1.0) * ub * ub, gam * ub]])
    ev, R = np.linalg.eig(A)
    assert np.all(np.isreal(ev)) and len(set(np.round(ev.real, 12))) == 3
    R = R.real
    L = np.linalg.inv(R)
    phi = orc.phi_w9 if S == 9 else orc.phi_w5
    ph = np.zeros(3)
    for m in range(3):
        fp = np.array([0.5 * L[m] @ (F[i0 + s - r] + lam * U[i0 + s - r]) for s in range(S)])
        fm = np.array([0.5 * L[m] @ (F[i0 + r + 1 - s] - lam * U[i0 + r + 1 - s]) for s in range(S)])
        ph[m] = -phi(o.c.eps, fp, np.diff(fp)) + phi(o.c.eps, fm, -np.diff(fm))
    if S == 9:
        w = {0: 533.0, 1: -139.0, 2: 29.0, 3: -3.0}
        th = sum(w[a] * (F[i0 + 1 + a] + F[i0 - a]) for a in range(4)) / 840.0
    else:
        th = (7.0 * (F[i0] + F[i0 + 1]) - (F[i0 - 1] + F[i0 + 2])) / 12.0
    return R @ ph + th


def char_flux_1d(S, seed=3):
    """max relative deviation of the oracle's interface flux (mass, x momentum, energy) from the
    numpy characteristic flux, over interfaces with a jump between their two nodes and smooth
    ones; the y/z momentum fluxes must vanish."""
    r = (S - 1) // 2
    n = (4 * S, 1, 1)
    o = orc.Oracle(dims=2, S=S, n=n, inviscid=True)
    rng = np.random.default_rng(seed)
    met = o.zeros(10)
    met[0][...] = 1.0
    met[4][...] = 1.0
    met[8][...] = 1.0
    met[9][...] = 1.0
    worst = 0.0
    ng0 = o.shape[0]
    for i0 in (o.g + r, o.g + r + 3, o.g + 2 * r + 1):
        x = np.arange(ng0, dtype=np.float64)
        rho = 1.0 + 0.2 * np.sin(0.7 * x + rng.uniform(0, 6))
        u = 0.6 + 0.3 * np.cos(0.5 * x + rng.uniform(0, 6))
        p = 1.0 + 0.1 * np.sin(0.9 * x)
        jump = x > i0
        rho = np.where(jump, 0.35 * rho, rho)
        p = np.where(jump, 0.3 * p, p)
        u = np.where(jump, u - 0.8, u)
        prim = o.zeros(5)
        prim[0][...] = rho[:, None, None]
        prim[1][...] = u[:, None, None]
        prim[4][...] = cases.energy(u, 0 * u, 0 * u, p, rho, o.c.gamma, o.c.mach)[:, None, None]
        der = o.calc_values(met, prim)
        out = o.weno_interface(0, met, prim, der, i0, o.g, 0)
        ref = _char_flux_numpy(o, prim, der, i0, S)
        scale = np.max(np.abs(ref))
        worst = max(worst, float(np.max(np.abs(out[[0, 1, 4]] - ref)) / scale),
                    float(np.max(np.abs(out[[2, 3]]))) / scale)
    return worst


def mu_locality(dims, dr, seed=5):
    """Response of the viscous Q of direction dr to a change of mu at one interior node m, on a
    curvilinear grid A and on grid B = A with the metric (and J) of every node but m perturbed:
    max |dQ_A - dQ_B| / max |dQ_A|.  Also returns max |dQ_A| relative to the viscous Q (the
    response must not be trivially zero)."""
    n = (9, 8, 1) if dims == 2 else (8, 9, 10)
    o = orc.Oracle(dims=dims, S=5, n=n, quirks=0, reynolds=50.0)
    tp = 2.0 * np.pi
    if dims == 2:
        mapping = lambda a, b, c: (a + 0.04 * np.sin(tp * b), b + 0.05 * np.sin(tp * a), 0 * a)
    else:
        mapping = lambda a, b, c: (a + 0.04 * np.sin(tp * b), b + 0.05 * np.sin(tp * c), c + 0.03 * np.sin(tp * a))
    xe, ye, ze = cases.periodic_ext_coords(dims, 5, n, mapping)
    metA = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(metA)
    rng = np.random.default_rng(seed)
    prim = _uniform(o)
    sl = _interior(o, n)
    for f, a in ((1, 0.3), (2, 0.2), (3, 0.25), (0, 0.1)):
        if f == 3 and dims == 2:
            continue
        prim[f][sl] += a * rng.standard_normal(prim[f][sl].shape)
    prim[4][...] = cases.energy(prim[1], prim[2], prim[3], 1.1, prim[0], o.c.gamma, o.c.mach)
    o.ghost_periodic(prim)
    der = o.calc_values(metA, prim)
    grad = o.visc_derivs(prim, der)
    g = o.g
    m = (g + n[0] // 2, g + n[1] // 2, (g + n[2] // 2) if dims == 3 else 0)

    def response(met):
        q0 = o.zeros(5)
        o.visc_dir(dr, met, prim, der, grad, q0)
        d2 = [a.copy() for a in der]
        d2[3][m] *= 1.3
        q1 = o.zeros(5)
        o.visc_dir(dr, met, prim, d2, grad, q1)
        return np.stack([b - a for a, b in zip(q0, q1)]), np.stack(q0)

    rA, q0 = response(metA)
    metB = [a * (1.0 + 0.15 * rng.uniform(-1, 1, size=a.shape)) for a in metA]
    for a, b in zip(metA, metB):
        b[m] = a[m]
    rB, _ = response(metB)
    scale = float(np.max(np.abs(rA)))
    return float(np.max(np.abs(rA - rB))) / scale, scale / float(np.max(np.abs(q0)))
