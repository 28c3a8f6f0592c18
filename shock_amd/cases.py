"""Case builders: grids and initial states for the SHOCK RHS / RK path.

Host-side setup only (numpy); the time-stepping path runs on the GPU through
libshockx.so.  Arrays use the reference's ghost-padded k-fastest layout
(src/Functions.c:56-103): shape (ng0, ng1, ng2), ng = n + S + 1, g = (S+1)/2
ghost layers, interior indices [g, g+n-1].  In 2D ng2 = 1.

Non-dimensionalisation follows SHOCK (docs/goveq.tex:143-163): Upsilon =
1/(gamma Ma^2), e = |u|^2/2 + Upsilon p / ((gamma-1) rho), c^2 = Upsilon gamma p / rho.
Inputs quoted in acoustic units (c = 1 at p = 1/gamma, rho = 1) map to SHOCK units
by u_S = u / Ma and p_S = gamma p.
"""
import numpy as np


def padded_shape(dims, S, n):
    g = (S + 1) // 2
    ng = [n[0] + S + 1, n[1] + S + 1, (n[2] + S + 1) if dims == 3 else 1]
    return tuple(ng), g


def _extend_axis(a, axis, gx, periodic, shift):
    """Add gx layers on both sides of `axis`.  Periodic: wrapped copy plus `shift` times the wrap
    count (the periodic mesh transfer, src/Functions.c:2657-2661).  Otherwise each layer mirrors
    the spacing of the interior node at the same distance from the face
    (ExtrapolateGhostCells, src/Functions.c:6526-6670): x[-1-l] = x[-l] + x[l] - x[l+1]."""
    n = a.shape[axis]
    if periodic:
        e = np.arange(-gx, n + gx)
        wraps = np.floor_divide(e, n)
        out = np.take(a, e - wraps * n, axis=axis)
        sh = [1, 1, 1]
        sh[axis] = e.size
        return out + (wraps.astype(np.float64) * shift).reshape(sh)
    out = np.moveaxis(np.zeros(a.shape[:axis] + (n + 2 * gx,) + a.shape[axis + 1:]), axis, 0)
    src = np.moveaxis(a, axis, 0)
    out[gx:gx + n] = src
    for l in range(gx):
        lo = gx - 1 - l
        out[lo] = out[lo + 1] + out[gx + l] - out[gx + l + 1]
        hi = gx + n + l
        out[hi] = out[hi - 1] + out[gx + n - 1 - l] - out[gx + n - 2 - l]
    return np.ascontiguousarray(np.moveaxis(out, 0, axis))


def ext_coords(dims, S, n, mapping, periodic=(True, True, True), period=(1.0, 1.0, 1.0)):
    """Extended coordinate arrays (g+1 layers beyond the interior on each side, the reference's
    x_extrapolate, src/Functions.c:6388-6420 and 6672-6790) of a block.

    mapping(xi, eta, zeta) -> (x, y, z) gives the interior coordinates; xi = i/n along periodic
    directions (the end point is the next period's first node) and xi = i/(n-1) along
    non-periodic ones (both faces are grid nodes).  Directions are extended in the order
    i, j, k, each over the already extended earlier directions, as the reference does.
    """
    ng, g = padded_shape(dims, S, n)
    ii = []
    for d in range(3):
        if d >= dims:
            ii.append(np.zeros(1))
        else:
            ii.append(np.arange(n[d]) / (n[d] if periodic[d] else (n[d] - 1)))
    XI, ETA, ZETA = np.meshgrid(ii[0], ii[1], ii[2], indexing="ij")
    arrs = [np.broadcast_to(a, XI.shape).astype(np.float64) for a in mapping(XI, ETA, ZETA)]
    for d in range(dims):
        arrs = [_extend_axis(a, d, g + 1, periodic[d], period[d] if comp == d else 0.0)
                for comp, a in enumerate(arrs)]
    xe, ye, ze = [np.ascontiguousarray(a) for a in arrs]
    if dims == 2:
        ze = np.zeros(xe.shape)
    return xe, ye, ze


def ext_from_nodes(dims, S, xyz, periodic=(False, False, False), period=(1.0, 1.0, 1.0)):
    """Extended coordinate arrays of a block from its node coordinates (x, y, z of shape
    (ni, nj, nk), as a mesh file gives them): g+1 layers added along each direction, i, j, k in
    turn, as ext_coords does."""
    g = (S + 1) // 2
    arrs = [np.asarray(a, dtype=np.float64) for a in xyz]
    for d in range(dims):
        arrs = [_extend_axis(a, d, g + 1, periodic[d], period[d] if comp == d else 0.0)
                for comp, a in enumerate(arrs)]
    xe, ye, ze = [np.ascontiguousarray(a) for a in arrs]
    if dims == 2:
        ze = np.zeros(xe.shape)
    return xe, ye, ze


def periodic_ext_coords(dims, S, n, mapping, period=(1.0, 1.0, 1.0)):
    """Extended coordinates of a block periodic in every direction (see ext_coords)."""
    return ext_coords(dims, S, n, mapping, (True, True, True), period)


def interior_coords(dims, S, n, xe, ye, ze):
    """Padded node coordinates (without the extra layer) from the extended arrays."""
    if dims == 3:
        return xe[1:-1, 1:-1, 1:-1], ye[1:-1, 1:-1, 1:-1], ze[1:-1, 1:-1, 1:-1]
    return xe[1:-1, 1:-1, :], ye[1:-1, 1:-1, :], ze[1:-1, 1:-1, :]


def cartesian_unit(dims):
    if dims == 2:
        return lambda xi, eta, zeta: (xi, eta, np.zeros_like(xi))
    return lambda xi, eta, zeta: (xi, eta, zeta)


def sine_bump_3d(amp=0.05):
    """x = xi + amp sin(2 pi xi) sin(2 pi eta) sin(2 pi zeta), y = eta, z = zeta."""
    tp = 2.0 * np.pi
    return lambda xi, eta, zeta: (xi + amp * np.sin(tp * xi) * np.sin(tp * eta) * np.sin(tp * zeta), eta, zeta)


def energy(u, v, w, p, rho, gamma, mach):
    ups = 1.0 / (gamma * (mach * mach))
    return 0.5 * (u * u + v * v + w * w) + ups * p / ((gamma - 1.0) * rho)


def mms_periodic_2d_state(x, y, gamma=1.4, mach=0.5):
    """Manufactured field on the unit square (acoustic units), mapped to SHOCK units:
    rho = 1 + 0.1 sin(2pi x) sin(2pi y), u = 0.5 + 0.1 cos(2pi x) sin(2pi y),
    v = 0.1 sin(2pi x) cos(2pi y), p = 1/gamma + 0.05 cos(2pi (x + y))."""
    tp = 2.0 * np.pi
    rho = 1.0 + 0.1 * np.sin(tp * x) * np.sin(tp * y)
    u = (0.5 + 0.1 * np.cos(tp * x) * np.sin(tp * y)) / mach
    v = (0.1 * np.sin(tp * x) * np.cos(tp * y)) / mach
    w = np.zeros_like(x)
    p = gamma * (1.0 / gamma + 0.05 * np.cos(tp * (x + y)))
    e = energy(u, v, w, p, rho, gamma, mach)
    return [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w, e)]


def random_phase_velocity(n, seed, kpeak=4.0, mt=0.5, c=1.0):
    """Solenoidal random-phase velocity field on an n^3 periodic box with energy spectrum
    E(k) ~ k^4 exp(-2 (k/kpeak)^2), scaled so that sqrt(<|u|^2>) = mt * c."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kz1 = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k1, k1, kz1, indexing="ij")
    kk = np.sqrt(KX * KX + KY * KY + KZ * KZ)
    kk_safe = np.where(kk == 0.0, 1.0, kk)
    spec = kk_safe ** 4 * np.exp(-2.0 * (kk_safe / kpeak) ** 2)
    amp = np.sqrt(spec / (4.0 * np.pi * kk_safe * kk_safe))
    amp[kk == 0.0] = 0.0
    # Nyquist modes have no representable -k partner: drop them so the field stays solenoidal
    nyq = (np.abs(KX) == n // 2) | (np.abs(KY) == n // 2) | (np.abs(KZ) == n // 2)
    amp[nyq] = 0.0
    uh = []
    for _ in range(3):
        ph = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
        uh.append(amp * np.exp(1j * ph))
    # solenoidal projection: u_hat -= k (k . u_hat) / |k|^2
    kdotu = (KX * uh[0] + KY * uh[1] + KZ * uh[2]) / (kk_safe * kk_safe)
    uh = [uh[0] - KX * kdotu, uh[1] - KY * kdotu, uh[2] - KZ * kdotu]
    vel = [np.fft.irfftn(a, s=(n, n, n), axes=(0, 1, 2)) for a in uh]
    rms = np.sqrt(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2))
    scale = mt * c / rms
    return [v * scale for v in vel]


def turbulence_3d_state(dims, S, n, seed, gamma=1.4, mach=0.5, mt=0.5):
    """Decaying compressible turbulence: uniform rho = 1, p = 1 (SHOCK units), random-phase
    solenoidal velocity with turbulent Mach number mt (c = 1/mach in SHOCK units)."""
    ng, g = padded_shape(dims, S, n)
    c = 1.0 / mach
    u, v, w = random_phase_velocity(n[0], seed, mt=mt, c=c)
    prim = [np.zeros(ng) for _ in range(5)]
    sl = (slice(g, g + n[0]), slice(g, g + n[1]), slice(g, g + n[2]))
    prim[0][...] = 1.0
    prim[1][sl] = u
    prim[2][sl] = v
    prim[3][sl] = w
    p = 1.0
    prim[4][...] = energy(prim[1], prim[2], prim[3], p, prim[0], gamma, mach)
    return prim


def perturbed_box(seed, amp_h, n, periodic=(False, True, True), modes=3):
    """Mapping of the unit cube with a seeded smooth perturbation of each coordinate, at most
    amp_h grid spacings: x_d = xi_d + amp_h h_d sum_m a_m sin(2 pi k_m . xi + phi_m) with the
    amplitudes normalised so the sum stays within [-1, 1].  Wave numbers are integers along
    periodic directions, so the perturbation is periodic there."""
    rng = np.random.default_rng(seed)
    tp = 2.0 * np.pi
    terms = []
    for d in range(3):
        h = 1.0 / (n[d] if periodic[d] else (n[d] - 1))
        ks = rng.integers(1, 3, size=(modes, 3)).astype(np.float64)
        ph = rng.uniform(0.0, tp, size=(modes, 3))
        a = rng.uniform(0.5, 1.0, size=modes)
        a = a / a.sum()
        terms.append((amp_h * h, ks, ph, a))

    def mapping(xi, eta, zeta):
        out = []
        for d, base in enumerate((xi, eta, zeta)):
            amp, ks, ph, a = terms[d]
            pert = np.zeros_like(base, dtype=np.float64)
            for m in range(len(a)):
                pert = pert + a[m] * np.sin(tp * ks[m, 0] * xi + ph[m, 0]) * np.sin(tp * ks[m, 1] * eta + ph[m, 1]) \
                    * np.sin(tp * ks[m, 2] * zeta + ph[m, 2])
            out.append(base + amp * pert)
        return tuple(out)
    return mapping


def normal_shock_states(mach1, gamma=1.4):
    """Rankine-Hugoniot jump of a normal shock with upstream Mach mach1:
    (rho2/rho1, u2/u1, p2/p1)."""
    m2 = mach1 * mach1
    r = (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    p = 1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0)
    return r, 1.0 / r, p


def shock_3d(S, n, seed, mach=3.0, gamma=1.4, x_shock=0.3, vort_rms=0.1, amp_h=0.05):
    """Mach-`mach` planar normal shock at x = x_shock meeting a seeded random-phase vorticity field
    (rms velocity vort_rms of the upstream speed) on a perturbed curvilinear box, inflow at x = 0,
    outflow at x = 1, periodic in y and z.  SHOCK units with the reference Mach equal to the
    upstream Mach: upstream rho = p = u = 1, c = 1/mach.  Returns the extended coordinates, the
    primitive state (padded) and the inflow tuple (rho, u, v, w, p)."""
    xe, ye, ze = ext_coords(3, S, n, perturbed_box(seed, amp_h, n), periodic=(False, True, True))
    ng, g = padded_shape(3, S, n)
    x = xe[1:-1, 1:-1, 1:-1]
    rr, ur, pr_ = normal_shock_states(mach, gamma)
    post = x > x_shock
    rho = np.where(post, rr, 1.0)
    u = np.where(post, ur, 1.0)
    p = np.where(post, pr_, 1.0)
    # vortical fluctuations on the index grid (a smooth periodic field), upstream speed 1
    fu, fv, fw = random_phase_velocity(n[0], seed + 1, mt=vort_rms, c=1.0)
    v = np.zeros(ng)
    w = np.zeros(ng)
    sl = (slice(g, g + n[0]), slice(g, g + n[1]), slice(g, g + n[2]))
    uf = np.zeros(ng)
    uf[sl] = fu
    v[sl] = fv
    w[sl] = fw
    u = u + uf
    e = energy(u, v, w, p, rho, gamma, mach)
    prim = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w, e)]
    return (xe, ye, ze), prim, (1.0, 1.0, 0.0, 0.0, 1.0)
