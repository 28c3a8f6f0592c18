"""Shared setup for the tests: grids, states and comparison metrics.

The oracle (oracle/orc.py) is the checker; shock_amd.cases builds the inputs both sides get.
"""
import numpy as np

import orc
from shock_amd import cases


def interior(o_or_g, dims, n):
    g = o_or_g if isinstance(o_or_g, int) else o_or_g.g
    if dims == 2:
        return (slice(g, g + n[0]), slice(g, g + n[1]), slice(0, 1))
    return (slice(g, g + n[0]), slice(g, g + n[1]), slice(g, g + n[2]))


def conserved(prim):
    rho, u, v, w, e = prim
    return [rho, rho * u, rho * v, rho * w, rho * e]


def rel_linf(a, b, sl, floor=1e-300):
    """max |a-b| / max |b| over the interior (per field)."""
    out = []
    for x, y in zip(a, b):
        den = max(np.max(np.abs(y[sl])), floor)
        out.append(float(np.max(np.abs(x[sl] - y[sl])) / den))
    return out


def setup_2d_mms(n=32, S=5, time_order=4, inviscid=False, mach=0.5, reynolds=1000.0, eps_override=0.0):
    """2D periodic unit square with the manufactured field (acoustic -> SHOCK units)."""
    nn = (n, n, 1)
    o = orc.Oracle(dims=2, S=S, n=nn, time_order=time_order, mach=mach, reynolds=reynolds,
                   inviscid=inviscid, eps_override=eps_override)
    xe, ye, ze = cases.periodic_ext_coords(2, S, nn, cases.cartesian_unit(2))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    x, y, z = cases.interior_coords(2, S, nn, xe, ye, ze)
    prim = cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y), mach=mach)
    o.ghost_periodic(prim)
    return o, (xe, ye, ze), met, prim, nn


def setup_3d_turb(n=16, S=5, time_order=4, inviscid=False, quirks=orc.QUIRKS_REFERENCE, seed=7,
                  reynolds=2000.0, mach=0.5, amp=0.05):
    """3D periodic curvilinear box (x = xi + amp sin sin sin) with random-phase turbulence."""
    nn = (n, n, n)
    o = orc.Oracle(dims=3, S=S, n=nn, time_order=time_order, mach=mach, reynolds=reynolds,
                   inviscid=inviscid, quirks=quirks)
    xe, ye, ze = cases.periodic_ext_coords(3, S, nn, cases.sine_bump_3d(amp))
    met = o.metric_from_coords(xe, ye, ze)
    o.ghost_periodic(met)
    prim = cases.turbulence_3d_state(3, S, nn, seed, mach=mach)
    o.ghost_periodic(prim)
    return o, (xe, ye, ze), met, prim, nn


def smooth_periodic_state(o, n, seed, mach):
    """A smooth periodic field on a non-cubic block: three integer-wave-number modes per variable
    with seeded phases (SHOCK units, c = 1/mach)."""
    rng = np.random.default_rng(seed)
    g = o.g
    I, J, K = np.meshgrid(*[np.arange(n[d]) / n[d] for d in range(3)], indexing="ij")
    tp = 2.0 * np.pi

    def field(amp):
        f = np.zeros(I.shape)
        for _ in range(3):
            k = rng.integers(1, 4, size=3)
            f += amp / 3.0 * np.sin(tp * (k[0] * I + k[1] * J + k[2] * K) + rng.uniform(0, tp))
        return f
    prim = o.zeros(5)
    sl = interior(o, 3, n)
    prim[0][...] = 1.0
    prim[0][sl] += field(0.1)
    for m in (1, 2, 3):
        prim[m][sl] = field(0.4 / mach)
    p = np.ones(o.shape)
    p[sl] += field(0.1)
    prim[4][...] = cases.energy(prim[1], prim[2], prim[3], p, prim[0], o.c.gamma, mach)
    o.ghost_periodic(prim)
    return prim
