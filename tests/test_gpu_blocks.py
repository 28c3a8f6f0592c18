"""Multi-block cases on one device (shx_blocks): the reference's zone interfaces with their halo
transfer (src/Functions.c:3580-3684) and CGNS transforms (src/Functions.c:3902-4287), as an
on-device gather before every stage.

A periodic box cut into two blocks must evolve as the uncut box: every interior node sees the same
stencil values, so the blocks' states equal the single block's.
  * identity cut (3D along i and, for SSP-RK3, along k; 2D along j): bit for bit against the
    single block on the GPU, and against the oracle's single block after one step (1e-12); the
    mixed FP32/FP64 form likewise (along i and j, WENO5 and WENO9), bit for bit against its own
    single block and within its one-step bar (1e-7) of the FP64 oracle;
  * rotated cut: block B stores its axes turned a quarter turn about k (its i runs along -j of the
    box, its j along i), so the interface transform is (2, -1, 3) / (-2, 1, 3) and the ghost metric
    takes the donor's rows permuted and signed; the general metric (no diagonal-metric quirk).  The
    sweeps of B run along other box directions than the single block's, so the sums round
    differently: 1e-12 after three steps.
  * the ghost metric of the interface faces, zeroed and transferred again, equals the metric the
    blocks had (bit for bit for the identity cut; within 1e-14 for the rotated one).
"""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, smooth_periodic_state
from shock_amd import cases, shockx

pytestmark = pytest.mark.gpu

P = shockx.BC_PERIODIC
I = shockx.BC_INTERFACE


def _case(n, dims, quirks, bc, S=5, inviscid=False, time_order=4, precision=shockx.FP64):
    return shockx.Case(dims=dims, space_order=S, time_order=time_order, n=n, mach=0.5, reynolds=2000.0,
                       quirks=quirks, bc=bc, inviscid=inviscid, precision=precision)


def _box(dims, n, S, seed, quirks=orc.QUIRKS_REFERENCE, inviscid=False, time_order=4):
    """Whole periodic box: the oracle's configuration, extended coordinates (sine bump in 3D) and a
    smooth periodic state with its ghosts filled."""
    o = orc.Oracle(dims=dims, S=S, n=n, quirks=quirks, mach=0.5, reynolds=2000.0, inviscid=inviscid,
                   time_order=time_order)
    if dims == 3:
        coords = cases.periodic_ext_coords(3, S, n, cases.sine_bump_3d(0.05))
        prim = smooth_periodic_state(o, n, seed, 0.5)
    else:
        coords = cases.periodic_ext_coords(2, S, n, cases.cartesian_unit(2))
        x, y, _ = cases.interior_coords(2, S, n, *coords)
        prim = cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y), mach=0.5)
        o.ghost_periodic(prim)
    return o, coords, prim


def _slab(a, axis, start, length):
    sl = [slice(None)] * a.ndim
    sl[axis] = slice(start, start + length)
    return np.ascontiguousarray(a[tuple(sl)])


@pytest.mark.parametrize("dims,n,axis,quirks,S,inviscid,rk,prec", [
    (3, (24, 16, 16), 0, orc.QUIRKS_REFERENCE, 5, False, 4, shockx.FP64),
    (2, (24, 24, 1), 1, orc.QUIRKS_REFERENCE, 9, False, 4, shockx.FP64),
    # SSP-RK3, cut along k; without the zeta-index quirk, whose reads depend on the block's extents
    (3, (16, 16, 24), 2, orc.QUIRK_METRIC_DIAG3D, 5, False, 3, shockx.FP64),
    # the mixed FP32/FP64 form: bit-equal to its own whole box, the mixed bar against the oracle
    (3, (24, 16, 16), 0, orc.QUIRKS_REFERENCE, 5, False, 4, shockx.MIXED),
    # (a cut along j without the zeta-index quirk: the quirk's zeta stencil sits at k = j + l - g,
    # so with it the operator depends on where a block starts along j)
    (3, (16, 24, 16), 1, orc.QUIRK_METRIC_DIAG3D, 9, False, 4, shockx.MIXED),
])
def test_identity_cut_equals_whole_box(dims, n, axis, quirks, S, inviscid, rk, prec):
    g = (S + 1) // 2
    o, coords, prim = _box(dims, n, S, 11, quirks, inviscid, rk)
    half = n[axis] // 2
    nb = list(n)
    nb[axis] = half
    nb = tuple(nb)
    steps = 3
    met = o.metric_from_coords(*coords)
    o.ghost_periodic(met)
    tau = o.cfl_dt(0.5, met, prim)
    with _case(n, dims, quirks, (P,) * 6, S, inviscid, rk, prec) as whole:
        whole.build_metric(*coords)
        whole.upload_state(prim)
        whole.run(tau, 1)
        one = whole.download_state()
        whole.run(tau, steps - 1)
        ref = whole.download_state()
    bc = [P] * 6
    bc[2 * axis] = bc[2 * axis + 1] = I
    blocks = []
    try:
        for b in range(2):
            c = _case(nb, dims, quirks, tuple(bc), S, inviscid, rk, prec)
            blocks.append(c)
            c.build_metric(*[_slab(x, axis, b * half, nb[axis] + 2 * g + 2) for x in coords])
            c.upload_state([_slab(f, axis, b * half, nb[axis] + 2 * g) for f in prim])
        lo, hi = 2 * axis, 2 * axis + 1
        ident = (1, 2, 3)
        conn = [(0, hi, 1, lo, ident), (1, lo, 0, hi, ident), (0, lo, 1, hi, ident), (1, hi, 0, lo, ident)]
        with shockx.Blocks(blocks, conn) as grp:
            grp.transfer_metric()
            grp.run(tau, 1)
            got1 = [np.concatenate([_slab(blocks[b].download_state()[m], axis, g, half) for b in range(2)], axis=axis)
                    for m in range(5)]
            grp.run(tau, steps - 1)
            got = [np.concatenate([_slab(blocks[b].download_state()[m], axis, g, half) for b in range(2)], axis=axis)
                   for m in range(5)]
    finally:
        for c in blocks:
            c.close()
    sl = interior(o, dims, n)
    cut = list(sl)
    cut[axis] = slice(None)
    cut = tuple(cut)
    same = all(np.array_equal(a[cut], _slab(b, axis, g, n[axis])[cut]) for a, b in zip(got, ref))
    diff = rel_linf(conserved(got), conserved([_slab(b, axis, g, n[axis]) for b in ref]), cut)
    print(f"[blocks] identity cut dims={dims} S={S} RK{rk} {'mixed' if prec == shockx.MIXED else 'fp64'}: bit-equal to the whole box: {same}; rel Linf "
          + " ".join(f"{e:.2e}" for e in diff), flush=True)
    assert same
    want1 = o.run(tau, met, prim, 1)
    errs = rel_linf(conserved(got1), conserved([_slab(b, axis, g, n[axis]) for b in want1]), cut)
    print(f"[blocks] identity cut one step vs oracle: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    # the mixed form's one-step bar against the FP64 oracle is its FP32 rounding of the deviation
    assert max(errs[:3] + errs[4:] if dims == 2 else errs) <= (1e-7 if prec == shockx.MIXED else 1e-12)
    assert all(np.array_equal(a[cut], b[cut]) for a, b in
               zip(got1, [_slab(x, axis, g, n[axis]) for x in one]))


def _rot(a):
    """Box array (padded or extended; block B holds box I in [12, 24) of the interior) -> block B
    storage: B's i runs along -J of the box, its j along I: B[i', j', k] = box[12 + j', last - i', k]."""
    return np.ascontiguousarray(a[12:, ::-1, :].transpose(1, 0, 2))


def test_rotated_cut_equals_whole_box():
    S, g = 5, 3
    n = (24, 16, 16)
    quirks = 0
    _, coords, prim = _box(3, n, S, 13, quirks)
    steps = 3
    with _case(n, 3, quirks, (P,) * 6, S) as whole:
        whole.build_metric(*coords)
        whole.upload_state(prim)
        tau = whole.cfl_dt(0.5)
        whole.run(tau, steps)
        ref = whole.download_state()
    # block A: box I in [0, 12), unrotated; block B: box I in [12, 24) turned (n = 16 x 12 x 16)
    A = _case((12, 16, 16), 3, quirks, (I, I, P, P, P, P), S)
    B = _case((16, 12, 16), 3, quirks, (P, P, I, I, P, P), S)
    try:
        A.build_metric(*[np.ascontiguousarray(x[:12 + 2 * g + 2]) for x in coords])
        A.upload_state([np.ascontiguousarray(f[:12 + 2 * g]) for f in prim])
        B.build_metric(*[_rot(x) for x in coords])
        B.upload_state([_rot(f) for f in prim])
        conn = [(0, 1, 1, 2, (2, -1, 3)), (1, 2, 0, 1, (-2, 1, 3)),
                (0, 0, 1, 3, (2, -1, 3)), (1, 3, 0, 0, (-2, 1, 3))]
        with shockx.Blocks([A, B], conn) as grp:
            ma = A.download_metric()
            mb = B.download_metric()
            # zero the interface ghost layers of the metric; the transfer restores them
            for blk, m0, ax in ((A, ma, 0), (B, mb, 1)):
                z = [x.copy() for x in m0]
                for x in z:
                    s1 = [slice(None)] * 3
                    s1[ax] = slice(0, g)
                    x[tuple(s1)] = 0.0
                    s1[ax] = slice(x.shape[ax] - g, None)
                    x[tuple(s1)] = 0.0
                blk.upload_metric(z)
            grp.transfer_metric()
            for blk, m0 in ((A, ma), (B, mb)):
                m1 = blk.download_metric()
                err = max(float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300) for a, b in zip(m1, m0))
                print(f"[blocks] rotated cut: transferred ghost metric vs built {err:.2e}", flush=True)
                assert err <= 1e-14
            grp.run(tau, steps)
            sa, sb = A.download_state(), B.download_state()
    finally:
        A.close()
        B.close()
    # B back to box order: box[12 + j', 21 - i', k] = B[i', j', k]
    got = []
    for m in range(5):
        a = sa[m][g:g + 12, g:g + 16, g:g + 16]
        b = sb[m][g:g + 16, g:g + 12, g:g + 16].transpose(1, 0, 2)[:, ::-1, :]
        got.append(np.concatenate([a, b], axis=0))
    want = [x[g:g + 24, g:g + 16, g:g + 16] for x in ref]
    full = (slice(None),) * 3
    errs = rel_linf(conserved(got), conserved(want), full)
    print(f"[blocks] rotated cut {steps} steps vs whole box: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-12


def test_group_refusals():
    n = (12, 12, 12)
    _, coords, prim = _box(3, (24, 12, 12), 5, 3, 0)
    A = _case(n, 3, 0, (I, I, P, P, P, P))
    try:
        A.build_metric(*[np.ascontiguousarray(x[:12 + 8]) for x in coords])
        A.upload_state([np.ascontiguousarray(f[:12 + 6]) for f in prim])
        with pytest.raises(shockx.ShxError):
            A.run(1e-4, 1)                       # interface faces need the group
        with pytest.raises(shockx.ShxError):
            shockx.Blocks([A], [(0, 1, 0, 0, (1, 2, 3))])          # face 0 not listed
        with pytest.raises(shockx.ShxError):
            shockx.Blocks([A], [(0, 1, 0, 0, (2, 1, 3)), (0, 0, 0, 1, (2, 1, 3))])   # normal not mapped to normal
        with shockx.Blocks([A], [(0, 1, 0, 0, (1, 2, 3)), (0, 0, 0, 1, (1, 2, 3))]) as grp:
            with pytest.raises(shockx.ShxError):
                A.close()                        # a block of a live group
            grp.transfer_metric()
            grp.run(1e-4, 1)                     # a block joined to itself: a periodic pair
    finally:
        A.close()


DIVISIONS = ["2d_cube_16x16.sddc", "3d_cube_16x16x16.sddc"]


@pytest.mark.parametrize("fname", DIVISIONS)
def test_reference_division_equals_single_block(fname):
    """The reference's testcase divisions (tests/golden/division, 4 zones in 2D, 8 in 3D) run as
    blocks on one device, joined by compat_division_connections, on a stretched Cartesian box with
    supersonic-inflow writers on the outer faces: three steps equal the undivided block bit for bit.
    The zeta-index quirk is off, since with it the reference's operator depends on the division (the
    k offset is taken from j)."""
    import ctypes
    import test_division as TD
    path = [f for f in TD.FILES if f.endswith(fname)][0]
    D, nz, zones = TD._read_all(path)
    L = TD._lib()
    L.compat_division_connections.argtypes = [ctypes.POINTER(TD.Division), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(TD.Connection)]
    arr = (TD.Division * nz)(*zones)
    out = (TD.Connection * (2 * D * nz))()
    nc = L.compat_division_connections(arr, nz, D, out)
    conn = [(c.block, c.face, c.donor, c.donor_face, tuple(c.transform)) for c in out[:nc]]
    lo = [[zones[z].corners[0][k] - 1 for k in range(D)] for z in range(nz)]
    ext = [[zones[z].corners[1][k] - zones[z].corners[0][k] + 1 for k in range(D)] for z in range(nz)]
    n = tuple(max(lo[z][k] + ext[z][k] for z in range(nz)) for k in range(D)) + ((1,) if D == 2 else ())
    S, g = 5, 3
    quirks = orc.QUIRK_METRIC_DIAG3D
    inflow = (1.0, 2.5, 0.0, 0.0, 1.0)
    amp = 0.05
    tp = 2.0 * np.pi
    if D == 3:
        mapping = lambda xi, eta, zeta: (xi + amp / tp * np.sin(tp * xi), eta + amp / tp * np.sin(tp * eta),
                                         zeta + amp / tp * np.sin(tp * zeta))
    else:
        mapping = lambda xi, eta, zeta: (xi + amp / tp * np.sin(tp * xi), eta + amp / tp * np.sin(tp * eta), zeta)
    coords = cases.ext_coords(D, S, n, mapping, periodic=(False, False, False))
    o = orc.Oracle(dims=D, S=S, n=n, quirks=quirks, mach=0.5, reynolds=2000.0)
    if D == 3:
        prim = smooth_periodic_state(o, n, 5, 0.5)
    else:
        x, y, _ = cases.interior_coords(2, S, n, *coords)
        prim = cases.mms_periodic_2d_state(np.ascontiguousarray(x), np.ascontiguousarray(y), mach=0.5)
    W = shockx.BC_INFLOW_SUPERSONIC

    def case(nn, bc):
        return shockx.Case(dims=D, space_order=S, time_order=4, n=nn, mach=0.5, reynolds=2000.0,
                           quirks=quirks, bc=bc, inflow=inflow)

    steps = 3
    with case(n, (W,) * 6) as whole:
        whole.build_metric(*coords)
        whole.upload_state(prim)
        tau = 0.5 * whole.cfl_dt(0.5)
        whole.run(tau, steps)
        ref = whole.download_state()
    blocks = []
    try:
        for z in range(nz):
            nn = tuple(ext[z]) + ((1,) if D == 2 else ())
            bc = [W] * 6
            for f in range(2 * D):
                if zones[z].face[f].neighbour:
                    bc[f] = shockx.BC_INTERFACE
            c = case(nn, tuple(bc))
            blocks.append(c)
            es = tuple(slice(lo[z][k], lo[z][k] + ext[z][k] + 2 * g + 2) for k in range(D)) + ((slice(None),) if D == 2 else ())
            ps = tuple(slice(lo[z][k], lo[z][k] + ext[z][k] + 2 * g) for k in range(D)) + ((slice(None),) if D == 2 else ())
            c.build_metric(*[np.ascontiguousarray(x[es]) for x in coords])
            c.upload_state([np.ascontiguousarray(f[ps]) for f in prim])
        with shockx.Blocks(blocks, conn) as grp:
            grp.transfer_metric()
            grp.run(tau, steps)
            got = [c.download_state() for c in blocks]
    finally:
        for c in blocks:
            c.close()
    same = True
    for z in range(nz):
        zi = tuple(slice(g, g + ext[z][k]) for k in range(D)) + ((slice(None),) if D == 2 else ())
        wi = tuple(slice(g + lo[z][k], g + lo[z][k] + ext[z][k]) for k in range(D)) + ((slice(None),) if D == 2 else ())
        for m in range(5):
            same = same and np.array_equal(got[z][m][zi], ref[m][wi])
    print(f"[blocks] {fname}: {nz} zones, {nc} connections, bit-equal to the single block: {same}", flush=True)
    assert same
