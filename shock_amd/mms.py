"""Manufactured-solution verification inputs for the RHS / RK path (host side, numpy).

The reference verifies SHOCK with a steady manufactured solution
(src/ManufacturedSolution.c): every primitive is a constant plus one sinusoid per coordinate,

    rho = rho0 + rho_x sin(a pi x) + rho_y cos(a pi y) + rho_z sin(a pi z)
    u   = u0   + u_x   sin(..x)    + u_y   cos(..y)    + u_z   cos(..z)
    v   = v0   + v_x   cos(..x)    + v_y   sin(..y)    + v_z   sin(..z)
    w   = w0   + w_x   sin(..x)    + w_y   sin(..y)    + w_z   cos(..z)
    p   = p0   + p_x   cos(..x)    + p_y   sin(..y)    + p_z   cos(..z)

(WriteManufacturedSolution, src/ManufacturedSolution.c:566-648), evaluated in long double at
the physical node coordinates (L = 1).  The source that makes this field steady is added to
every stage's Q as J * source (AddManufacturedSolutionSource, :25-175), the viscosity is mu = 1,
and the ghost layers of every face without a neighbour hold the exact state
(WriteBCManufacturedSolution*, :202-338).

The reference ships the source as expanded closed forms.  Here it is derived independently:
the steady residual of SHOCK's non-dimensional equations (momentum flux rho u u + Upsilon p,
energy flux (rho e + Upsilon p) u, stress Psi mu (grad u + grad u^T - 2/3 div u), heat flux
Gamma mu grad T with T = p / rho) is assembled from the analytic first and second derivatives
of the fields above, which separate per coordinate (all mixed second derivatives vanish).
"""
import numpy as np

LD = np.longdouble
PI_L = LD("3.14159265358979323846264338327950288419716939937510")

# Function kinds per field and coordinate (x, y, z): 's' = sin, 'c' = cos.
KINDS = {
    "rho": ("s", "c", "s"),
    "u": ("s", "c", "c"),
    "v": ("c", "s", "s"),
    "w": ("s", "s", "c"),
    "p": ("c", "s", "c"),
}
FIELDS = ("rho", "u", "v", "w", "p")


def _set(f0, fx, fy, fz, ax, ay, az):
    return [f0, fx, fy, fz, ax, ay, az]


def reference_params(dims, supersonic, inviscid, gamma=1.4, mach=0.5, R=287.0, T0=300.0):
    """Parameter sets of ConfigureManufacturedSolution (src/ManufacturedSolution.c:340-556),
    in SHOCK units: velocities over u_ref = Ma sqrt(gamma R T0), pressures over p_ref = 1e5,
    densities over rho_ref = 1.  Returns {field: [f0, fx, fy, fz, ax, ay, az]} (long double).

    The 3D subsonic sets are taken as the reference executes them: w_y and p_y are assigned
    twice there (:509-510, :517-518), so w_z = p_z = 0 and the second values of w_y and p_y hold.
    """
    u_ref = LD(mach) * np.sqrt(LD(gamma * R * T0))     # gamma R T0 formed in double, as in C
    p_ref = LD(100000)
    rho_ref = LD(1)

    def r(s):
        return LD(s) / rho_ref

    def p(s):
        return LD(s) / u_ref

    def q(s):
        return LD(s) * LD(100000) / p_ref

    t3 = LD(2) / LD(3)
    z = LD(0)
    if dims == 2:
        if supersonic:
            P = {
                "rho": _set(r("1"), r("0.15"), r("-0.1"), z, LD(1), LD("0.5"), z),
                "u": _set(p("800"), p("50"), p("-30"), z, LD("1.5"), LD("0.6"), z),
                "v": _set(p("800"), p("-75"), p("40"), z, LD("0.5"), t3 if inviscid else LD("1.5"), z),
                "p": _set(q("1"), q("0.2"), q("0.5"), z,
                          LD(2) if inviscid else t3, LD(1), z),
            }
        elif inviscid:
            P = {
                "rho": _set(r("1"), r("0.15"), r("-0.1"), z, LD(1), LD("0.5"), z),
                "u": _set(p("70"), p("5"), p("-7"), z, LD("1.5"), LD("0.6"), z),
                "v": _set(p("90"), p("-15"), p("8.5"), z, LD("0.5"), t3, z),
                "p": _set(q("1"), q("0.2"), q("0.5"), z, LD(2), LD(1), z),
            }
        else:
            P = {
                "rho": _set(r("1"), r("0.1"), r("0.15"), z, LD("0.75"), LD(1), z),
                "u": _set(p("70"), p("4"), p("-12"), z, LD(5) / LD(3), LD("1.5"), z),
                "v": _set(p("90"), p("-20"), p("4"), z, LD("1.5"), LD(1), z),
                "p": _set(q("1"), q("-0.3"), q("0.2"), z, LD(1), LD("1.25"), z),
            }
        P["w"] = _set(z, z, z, z, z, z, z)
        return P
    if supersonic:      # identical for Euler and Navier-Stokes (:455-470)
        return {
            "rho": _set(r("1"), r("0.15"), r("-0.1"), r("-0.12"), LD(1), LD("0.5"), LD("1.5")),
            "u": _set(p("800"), p("50"), p("-30"), p("-18"), LD("1.5"), LD("0.6"), LD("0.5")),
            "v": _set(p("800"), p("-75"), p("40"), p("-30"), LD("0.5"), t3, LD("1.25")),
            "w": _set(p("800"), p("15"), p("-25"), p("35"), LD(1) / LD(3), LD("1.5"), LD(1)),
            "p": _set(q("1"), q("0.2"), q("0.5"), q("-0.35"),
                      LD(2), LD(1), LD(1) / LD(3)),
        }
    return {            # subsonic, identical for Euler and Navier-Stokes (:505-519)
        "rho": _set(r("1"), r("0.1"), r("0.15"), r("-0.12"), LD("0.75"), LD(1), LD("1.5")),
        "u": _set(p("70"), p("4"), p("-12"), p("-8"), LD(5) / LD(3), LD("1.5"), LD("0.5")),
        "v": _set(p("90"), p("-20"), p("4"), p("-3"), LD("1.5"), LD(1), LD("1.25")),
        "w": _set(p("90"), p("-20"), p("6"), z, LD("1.5"), LD(1), LD("1.5")),
        "p": _set(q("1"), q("-0.3"), q("-0.3"), z, LD(1), LD("1.25"), LD(1) / LD(3)),
    }


def _field(P, name, X):
    """value, gradient (3) and diagonal second derivatives (3) of one field, long double."""
    prm = P[name]
    val = LD(prm[0]) + np.zeros(X[0].shape, dtype=LD)
    grad, d2 = [], []
    for d in range(3):
        amp, a = LD(prm[1 + d]), LD(prm[4 + d])
        k = a * PI_L
        arg = k * X[d]
        s, c = np.sin(arg), np.cos(arg)
        if KINDS[name][d] == "s":
            val = val + amp * s
            grad.append(amp * k * c)
            d2.append(-amp * k * k * s)
        else:
            val = val + amp * c
            grad.append(-amp * k * s)
            d2.append(-amp * k * k * c)
    return val, grad, d2


def _coords(x, y, z):
    return [np.asarray(a, dtype=np.float64).astype(LD) for a in (x, y, z)]


def exact_state(P, x, y, z, gamma, mach):
    """rho, u, v, w, e, p (float64) at the given coordinates, evaluated as
    WriteManufacturedSolution does: primitives in long double stored as double, then e from the
    stored doubles with long double arithmetic."""
    X = _coords(x, y, z)
    vals = {f: _field(P, f, X)[0].astype(np.float64) for f in FIELDS}
    rho, u, v, w, p = (vals[f] for f in FIELDS)
    ups = 1.0 / (gamma * (mach * mach))
    kin = LD("0.5") * (u * u + v * v + w * w).astype(LD)
    e = (kin + (p / rho).astype(LD) / (LD(gamma) - LD(1)) * LD(ups)).astype(np.float64)
    return [np.ascontiguousarray(a) for a in (rho, u, v, w, e, p)]


def source(P, x, y, z, gamma, mach, reynolds, prandtl, inviscid):
    """Steady residual of the exact field: (mass, x-, y-, z-momentum, energy) source per node,
    float64, long double arithmetic, mu = 1."""
    X = _coords(x, y, z)
    g = LD(gamma)
    ups = LD(1.0 / (gamma * (mach * mach)))
    psi = LD(1.0 / reynolds)                            # Psi = 1/Re stored in double
    gam_heat = LD(1) / ((g - LD(1)) * LD(mach) ** 2 * LD(reynolds) * LD(prandtl))
    rho, drho, d2rho = _field(P, "rho", X)
    p, dp, d2p = _field(P, "p", X)
    U, dU, d2U = [], [], []
    for f in ("u", "v", "w"):
        a, b, c = _field(P, f, X)
        U.append(a); dU.append(b); d2U.append(c)
    div = dU[0][0] + dU[1][1] + dU[2][2]
    adv_rho = U[0] * drho[0] + U[1] * drho[1] + U[2] * drho[2]
    mass = adv_rho + rho * div
    mom = []
    for j in range(3):
        adv = U[0] * dU[j][0] + U[1] * dU[j][1] + U[2] * dU[j][2]
        m = U[j] * mass + rho * adv + ups * dp[j]
        if not inviscid:
            lap = d2U[j][0] + d2U[j][1] + d2U[j][2]
            m = m - psi * (lap + d2U[j][j] / LD(3))   # div tau_j = lap u_j + 1/3 d_j div u
        mom.append(m)
    q2 = U[0] * U[0] + U[1] * U[1] + U[2] * U[2]
    H = LD("0.5") * rho * q2 + ups * g / (g - LD(1)) * p
    dH = [LD("0.5") * drho[i] * q2 + rho * (U[0] * dU[0][i] + U[1] * dU[1][i] + U[2] * dU[2][i])
          + ups * g / (g - LD(1)) * dp[i] for i in range(3)]
    energy = U[0] * dH[0] + U[1] * dH[1] + U[2] * dH[2] + H * div
    if not inviscid:
        work = LD(0)
        for j in range(3):
            lap = d2U[j][0] + d2U[j][1] + d2U[j][2]
            work = work + U[j] * (lap + d2U[j][j] / LD(3))
            for i in range(3):
                tau = dU[j][i] + dU[i][j] - (LD(2) / LD(3) * div if i == j else LD(0))
                work = work + tau * dU[j][i]
        T = p / rho
        lapT = LD(0)
        for i in range(3):
            Ti = (dp[i] - T * drho[i]) / rho
            lapT = lapT + (d2p[i] - LD(2) * Ti * drho[i] - T * d2rho[i]) / rho
        energy = energy - psi * work - gam_heat * lapT
    return [np.ascontiguousarray(a.astype(np.float64)) for a in (mass, mom[0], mom[1], mom[2], energy)]


def error_norms(P, x, y, z, rho, p, last_plane_only=False):
    """L2 and L-infinity errors of density and pressure against the exact field over the
    interior arrays given (shape n0 x n1 x n2).  The reference (ErrorManufacturedSolution,
    src/ManufacturedSolution.c:708-780) sums only the last k-plane but divides by the full node
    count N; last_plane_only=True reproduces that, False measures the whole block."""
    X = _coords(x, y, z)
    rex = _field(P, "rho", X)[0]
    pex = _field(P, "p", X)[0]
    N = LD(rho.size)
    sl = (slice(None), slice(None), slice(-1, None)) if last_plane_only else (slice(None),) * 3
    dr = rex[sl] - rho[sl].astype(LD)
    dp = pex[sl] - p[sl].astype(LD)
    return {
        "L2_rho": float(np.sqrt(np.sum(dr * dr) / N)), "Linf_rho": float(np.max(np.abs(dr))),
        "L2_p": float(np.sqrt(np.sum(dp * dp) / N)), "Linf_p": float(np.max(np.abs(dp))),
    }
