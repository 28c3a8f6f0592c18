"""Manufactured-solution convergence ladder on the GPU path (FP64 and mixed forms).

Each rung is a 3D block of n^3 nodes on [0, 1]^3 (uniform, or smoothly stretched and separable so
that the diagonal 3D metric of the reference stays exact) with the reference's subsonic Navier-Stokes
parameter set (src/ManufacturedSolution.c:505-519), manufactured-solution faces on all six sides
(src/ManufacturedSolution.c:202-338), the source added every stage (:25-175) and mu = 1.  The
state starts from the exact field and is marched to a steady state with the CFL time step; the
error against the exact field is the discretisation error of the scheme.  Reported per rung:
L2 and L-infinity of density and pressure over the whole block and, for comparison with the
reference's own output, its norm over the last k-plane divided by the full node count
(ErrorManufacturedSolution, src/ManufacturedSolution.c:708-780).

The zeta-index quirk of the reference's viscous operator is switched off here: with it the 3D
viscous residual does not converge (the manufactured residual stays O(1) under refinement).
On the stretched grid the reference's metric (second-order central differences of the
coordinates, src/Functions.c:6388-6509) limits the observed order to two; the uniform grid shows
the order of the flux discretisation.
"""
import math
import sys
import time

import numpy as np

from . import cases, mms, shockx

GAMMA, MACH, PRANDTL = 1.4, 0.5, 0.72


def stretched(amp=0.05):
    """x = xi + amp/(2 pi) sin(2 pi xi) (likewise y, z): non-uniform spacing, diagonal metric."""
    tp = 2.0 * math.pi
    return lambda xi, eta, zeta: (xi + amp / tp * np.sin(tp * xi), eta + amp / tp * np.sin(tp * eta),
                                  zeta + amp / tp * np.sin(tp * zeta))


def setup(n, S=5, precision=shockx.FP64, reynolds=1000.0, quirks=shockx.QUIRK_METRIC_DIAG3D, amp=0.0):
    """GPU case, exact state (rho,u,v,w,e,p on the padded grid) and interior coordinates."""
    nn = (n, n, n)
    per = (False, False, False)
    xe, ye, ze = cases.ext_coords(3, S, nn, stretched(amp), periodic=per)
    x, y, z = cases.interior_coords(3, S, nn, xe, ye, ze)
    P = mms.reference_params(3, False, False, gamma=GAMMA, mach=MACH)
    # the exact fields and the source are pointwise: evaluated in i slabs (long double arithmetic
    # takes minutes at 256^3), with a progress line per slab for large rungs
    nslab = max(1, x.shape[0] // 32)
    cuts = np.linspace(0, x.shape[0], nslab + 1).astype(int)
    exact = [np.empty_like(x) for _ in range(6)]
    src = [np.empty_like(x) for _ in range(5)]
    for q in range(nslab):
        sl = slice(cuts[q], cuts[q + 1])
        for a, b in zip(exact, mms.exact_state(P, x[sl], y[sl], z[sl], GAMMA, MACH)):
            a[sl] = b
        for a, b in zip(src, mms.source(P, x[sl], y[sl], z[sl], GAMMA, MACH, reynolds, PRANDTL, False)):
            a[sl] = b
        if nslab > 2:
            print(f"[ladder] n={n}: exact state and source, slab {q + 1} of {nslab}", file=sys.stderr, flush=True)
    c = shockx.Case(dims=3, space_order=S, time_order=4, n=nn, gamma=GAMMA, mach=MACH, reynolds=reynolds,
                    prandtl=PRANDTL, manufactured=True, quirks=quirks, bc=(shockx.BC_MANUFACTURED,) * 6,
                    precision=precision)
    c.build_metric(xe, ye, ze)
    c.upload_state(exact[:5])
    c.upload_boundary_state(exact)
    c.upload_source(src)
    return c, P, (x, y, z), exact


def rung(n, S=5, precision=shockx.FP64, t_end=2.0, cfl=0.5, reynolds=1000.0, chunk=50, amp=0.0):
    """March one rung to t_end; returns a dict of errors, steps and wall time.  amp: grid stretching
    (0: uniform)."""
    c, P, (x, y, z), exact = setup(n, S, precision, reynolds, amp=amp)
    g = c.g
    tau = c.cfl_dt(cfl)
    steps = int(math.ceil(t_end / tau))
    t0 = time.perf_counter()
    done = 0
    prev = None
    change = float("nan")
    last_note = t0
    while done < steps:
        k = min(chunk, steps - done)
        c.run(tau, k)
        done += k
        if steps - done < chunk and prev is None:
            prev = c.download_state()[0][g:-g, g:-g, g:-g].copy()
        if time.perf_counter() - last_note > 30.0:           # progress of a long rung
            c.synchronize()
            last_note = time.perf_counter()
            print(f"[ladder] n={n} S={S} {'fp64' if precision == shockx.FP64 else 'mixed'}: step {done} of "
                  f"{steps}", file=sys.stderr, flush=True)
    c.synchronize()
    wall = time.perf_counter() - t0
    st = c.download_state()
    sl = (slice(g, g + n),) * 3
    rho = st[0][sl]
    u, v, w, e = st[1][sl], st[2][sl], st[3][sl], st[4][sl]
    ups = 1.0 / (GAMMA * MACH * MACH)
    p = np.abs(rho * (GAMMA - 1.0) / ups * (e - 0.5 * (u * u + v * v + w * w)))
    if prev is not None:
        change = float(np.max(np.abs(rho - prev)))
    xi, yi, zi = x[sl], y[sl], z[sl]
    full = mms.error_norms(P, xi, yi, zi, rho, p)
    plane = mms.error_norms(P, xi, yi, zi, rho, p, last_plane_only=True)
    c.close()
    return {"n": n, "S": S, "precision": "fp64" if precision == shockx.FP64 else "mixed", "steps": steps,
            "tau": tau, "wall_s": round(wall, 3), "cell_updates_per_s": n ** 3 * 4 * steps / wall,
            "full": full, "last_plane": plane, "last_chunk_drho": change}


def orders(rows, key="L2_rho", norm="full"):
    out = []
    for a, b in zip(rows, rows[1:]):
        ea, eb = a[norm][key], b[norm][key]
        out.append(math.log(ea / eb) / math.log((b["n"] - 1) / (a["n"] - 1)) if ea > 0 and eb > 0 else float("nan"))
    return out
