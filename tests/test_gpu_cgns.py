"""The reference's own testcases on the GPU: meshes read from their CGNS files (compat_load_mesh,
the mesh part of loadFile, src/Load.c:252-744), boundary conditions from the files' BC families,
divisions from their .sddc files (tests/golden/cgns, tests/golden/division).

  * one zone: the whole testcase zone as one block (the reference run on one process), ghost
    coordinates extrapolated (ExtrapolateGhostCells), BCInflow / BCOutflow / BCWallInviscid faces
    as supersonic-inflow, Riemann-outflow and slip-wall writers: one step against the oracle on
    the same inputs, <= 1e-12 per conserved variable.  The 3D testcase is 64 x 64 x 16, so with
    the reference's quirks the zeta-index quirk reads across rows (n_eta > n_zeta);
  * divided: the testcase's division (4 zones in 2D, 8 in 3D: the 64 x 64 x 16 zone cut along i
    and j, the 32^3 zone along all three axes) as blocks on one device, each block
    read with its own record (its coordinates and its faces' BCs), joined by
    compat_division_connections: three inviscid steps equal the one-zone run bit for bit.  The
    zeta-index quirk is off here, since with it the reference's operator depends on the division.
    Viscous runs are not expected to: where a wall or outflow face meets an interface, the ghost
    nodes beyond both hold the donor's writer output of the previous stage (the reference's
    transfers copy ghost-inclusive planes before its writers run), and the viscous cross
    derivatives read them; the one-zone run has this stage's writer values there (measured: 6e-7
    relative after three steps of the 2D testcase).
  * the C driver: shock_amd/host/testcase_run (plain C over both libraries: division, mesh, BCs,
    ghost extrapolation, metric, blocks, shx_blocks_run) on every zone of the 2D and 3D testcases'
    divisions, viscous, three steps, equals bit for bit the same run assembled here from the
    Python readers and shockx.Blocks from the program's own initial states (written to its output
    file).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf
from shock_amd import cases, meshio, shockx

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
GAMMA, MACH, RE = 1.4, 0.5, 1000.0
INFLOW = (1.0, 0.8, 0.0, 0.0, 1.0)
OUTFLOW = (0.95, 3.0)


def _zone(name, dims, nproc):
    mesh = os.path.join(GOLD, "cgns", name + ".cgns")
    div = os.path.join(GOLD, "division", name + ".sddc")
    recs = [meshio.load_division(div, dims, nproc, z) for z in range(1, nproc + 1)]
    size = [max(r.corners[1][k] for r in recs) for k in range(dims)]
    return mesh, recs, size


def _state(dims, S, n, coords, amp=0.2):
    """the inflow state with smooth perturbations on every padded node"""
    x, y, z = cases.interior_coords(dims, S, n, *coords)
    tp = 2 * np.pi
    rho = INFLOW[0] * (1 + amp * np.sin(tp * x) * np.cos(tp * y))
    u = INFLOW[1] * (1 + amp * np.cos(tp * y + 0.3) * np.sin(tp * z + 0.1))
    # v of the same size in 2D (z = 0) as in 3D: a field far smaller than the others makes its own
    # relative error ill-conditioned (the oracle's one-ulp spread reaches the bar)
    v = amp * np.sin(tp * x) * (np.sin(tp * z + 0.2) if dims == 3 else np.cos(tp * y + 0.2))
    w = amp * np.cos(tp * x) * np.sin(tp * y) if dims == 3 else np.zeros_like(x)
    p = INFLOW[4] * (1 + amp * np.cos(tp * (x + y)))
    e = cases.energy(u, v, w, p, rho, GAMMA, MACH)
    return [np.ascontiguousarray(a) for a in (rho, u, v, w, e)]


def _case(dims, S, n, bc, quirks, inviscid=False):
    return shockx.Case(dims=dims, space_order=S, time_order=4, n=n, gamma=GAMMA, mach=MACH, reynolds=RE,
                       quirks=quirks, bc=bc, inflow=INFLOW, outflow=OUTFLOW, inviscid=inviscid)


# WENO5: on these states (inflow ghosts against a perturbed interior) WENO9 with the reference's
# epsilon, 10 sqrt(DBL_MIN), is ill-conditioned -- the oracle's own one-ulp spread after one step
# reaches 8e-13 (2D) and 3e-10 (3D) -- while WENO5's stays near 2e-15
@pytest.mark.parametrize("name,dims,nproc,S", [("2d_cube_64x64", 2, 4, 5), ("3d_cube_16x16x16", 3, 8, 5),
                                                ("3d_cube_32x32x32", 3, 8, 5)])
def test_testcase_zone_one_step_against_oracle(name, dims, nproc, S):
    mesh, recs, size = _zone(name, dims, nproc)
    whole = meshio.whole_zone(recs[0].zonename, size)
    xyz, info = meshio.load_mesh(mesh, dims, whole)
    bc = meshio.face_kinds(info, dims)
    print(f"[cgns] {name}: zone {size}, BCs {meshio.bc_strings(info, dims)} -> kinds {bc}", flush=True)
    n = tuple(size) + ((1,) if dims == 2 else ())
    coords = cases.ext_from_nodes(dims, S, xyz)
    quirks = orc.QUIRKS_REFERENCE
    o = orc.Oracle(dims=dims, S=S, n=n, bc=bc, inflow=INFLOW, mach=MACH, reynolds=RE, quirks=quirks,
                   outflow=OUTFLOW)
    met = o.metric_from_coords(*coords)
    o.ghost_periodic(met)
    o.mirror_metric(met)
    prim = _state(dims, S, n, coords)
    tau = 0.5 * o.cfl_dt(0.5, met, prim)
    want = o.run(tau, met, prim, 1)
    sl = interior(o, dims, n)
    rng = np.random.default_rng(7)
    spread = np.zeros(5)
    for _ in range(2):
        pert = [a * (1.0 + 2.2e-16 * rng.standard_normal(a.shape)) for a in prim]
        spread = np.maximum(spread, rel_linf(conserved(o.run(tau, met, pert, 1)), conserved(want), sl))
    with _case(dims, S, n, bc, quirks) as g:
        g.build_metric(*coords)
        g.upload_state(prim)
        g.run(tau, 1)
        got = g.download_state()
    errs = rel_linf(conserved(got), conserved(want), sl)
    if dims == 2:
        errs = errs[:3] + errs[4:]
        spread = np.concatenate([spread[:3], spread[4:]])
    print(f"[cgns] {name} one step vs oracle: " + " ".join(f"{e:.2e}" for e in errs)
          + "; the oracle's one-ulp spread " + " ".join(f"{e:.2e}" for e in spread), flush=True)
    assert max(errs) <= 1e-12


@pytest.mark.parametrize("name,dims,nproc", [("2d_cube_64x64", 2, 4), ("3d_cube_16x16x16", 3, 8),
                                             ("3d_cube_32x32x32", 3, 8)])
def test_testcase_division_blocks_equal_one_zone(name, dims, nproc):
    S, g = 5, 3
    quirks = orc.QUIRK_METRIC_DIAG3D
    mesh, recs, size = _zone(name, dims, nproc)
    n = tuple(size) + ((1,) if dims == 2 else ())
    # every block read with its own record, placed in the zone
    zone_xyz = [np.zeros(n) for _ in range(3)]
    infos = []
    for r in recs:
        xyz, info = meshio.load_mesh(mesh, dims, r)
        infos.append(info)
        sl = tuple(slice(r.corners[0][k] - 1, r.corners[1][k]) for k in range(dims))
        for a, b in zip(zone_xyz, xyz):
            a[sl] = b if dims == 3 else b
    coords = cases.ext_from_nodes(dims, S, zone_xyz)
    _, winfo = meshio.load_mesh(mesh, dims, meshio.whole_zone(recs[0].zonename, size))
    prim = _state(dims, S, n, coords)
    steps = 3
    with _case(dims, S, n, meshio.face_kinds(winfo, dims), quirks, True) as whole:
        whole.build_metric(*coords)
        whole.upload_state(prim)
        tau = 0.5 * whole.cfl_dt(0.5)
        whole.run(tau, steps)
        ref = whole.download_state()
    L = meshio.lib()
    L.compat_division_connections.argtypes = [ctypes.POINTER(meshio.Division), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(shockx.Interface)]
    arr = (meshio.Division * nproc)(*recs)
    out = (shockx.Interface * (2 * dims * nproc))()
    nc = L.compat_division_connections(arr, nproc, dims, out)
    assert nc > 0
    conn = [(c.block, c.face, c.donor, c.donor_face, tuple(c.transform)) for c in out[:nc]]
    lo = [[r.corners[0][k] - 1 for k in range(dims)] for r in recs]
    ext = [[r.corners[1][k] - r.corners[0][k] + 1 for k in range(dims)] for r in recs]
    blocks = []
    try:
        for z, r in enumerate(recs):
            nn = tuple(ext[z]) + ((1,) if dims == 2 else ())
            c = _case(dims, S, nn, meshio.face_kinds(infos[z], dims, r), quirks, True)
            blocks.append(c)
            es = tuple(slice(lo[z][k], lo[z][k] + ext[z][k] + 2 * g + 2) for k in range(dims)) + \
                ((slice(None),) if dims == 2 else ())
            ps = tuple(slice(lo[z][k], lo[z][k] + ext[z][k] + 2 * g) for k in range(dims)) + \
                ((slice(None),) if dims == 2 else ())
            c.build_metric(*[np.ascontiguousarray(x[es]) for x in coords])
            c.upload_state([np.ascontiguousarray(f[ps]) for f in prim])
        with shockx.Blocks(blocks, conn) as grp:
            grp.transfer_metric()
            grp.run(tau, steps)
            got = [c.download_state() for c in blocks]
    finally:
        for c in blocks:
            c.close()
    worst = 0.0
    same = True
    for z in range(nproc):
        zi = tuple(slice(g, g + ext[z][k]) for k in range(dims)) + ((slice(None),) if dims == 2 else ())
        wi = tuple(slice(g + lo[z][k], g + lo[z][k] + ext[z][k]) for k in range(dims)) + \
            ((slice(None),) if dims == 2 else ())
        for m in range(5):
            same = same and np.array_equal(got[z][m][zi], ref[m][wi])
            scale = max(float(np.max(np.abs(ref[m][wi]))), 1e-300)
            worst = max(worst, float(np.max(np.abs(got[z][m][zi] - ref[m][wi]))) / scale)
    print(f"[cgns] {name}: {nproc} blocks, {nc} connections, {steps} inviscid steps: bit-equal to the one-zone run: "
          f"{same} (largest relative difference {worst:.2e})", flush=True)
    assert same


def _connections(recs, dims):
    L = meshio.lib()
    L.compat_division_connections.argtypes = [ctypes.POINTER(meshio.Division), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(shockx.Interface)]
    arr = (meshio.Division * len(recs))(*recs)
    out = (shockx.Interface * (2 * dims * len(recs)))()
    nc = L.compat_division_connections(arr, len(recs), dims, out)
    assert nc > 0
    return [(c.block, c.face, c.donor, c.donor_face, tuple(c.transform)) for c in out[:nc]]


@pytest.mark.parametrize("name,dims,nproc", [("2d_cube_64x64", 2, 4), ("3d_cube_16x16x16", 3, 8),
                                             ("3d_cube_32x32x32", 3, 8)])
def test_c_driver_equals_the_python_assembly(tmp_path, name, dims, nproc):
    exe = os.path.join(ROOT, "shock_amd", "host", "testcase_run")
    if not os.path.exists(exe):
        pytest.skip("testcase_run not built")
    S, steps = 5, 3
    mesh, recs, size = _zone(name, dims, nproc)
    div = os.path.join(GOLD, "division", name + ".sddc")
    out = str(tmp_path / "states.bin")
    print(f"[cgns] C driver on {name}: {nproc} zones, {steps} steps", flush=True)
    r = subprocess.run([exe, mesh, div, str(dims), str(nproc), str(steps), out], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout + r.stderr, flush=True)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.float64)
    blocks, shapes = [], []
    try:
        for z, rec in enumerate(recs):
            xyz, info = meshio.load_mesh(mesh, dims, rec)
            n = meshio.block_size(dims, rec)
            c = shockx.Case(dims=dims, space_order=S, time_order=4, n=n, gamma=GAMMA, mach=MACH, reynolds=RE,
                            quirks=orc.QUIRK_METRIC_DIAG3D, bc=meshio.face_kinds(info, dims, rec), inflow=INFLOW,
                            outflow=OUTFLOW, T_wall=0.0)
            blocks.append(c)
            c.build_metric(*meshio.extend_coords(dims, S, xyz))
            shapes.append(cases.padded_shape(dims, S, n)[0])
        sizes = [int(np.prod(sh)) for sh in shapes]
        assert raw.size == 2 * 5 * sum(sizes)
        halves, off = [], 0
        for _ in range(2):
            states = []
            for sh, sz in zip(shapes, sizes):
                states.append([raw[off + m * sz: off + (m + 1) * sz].reshape(sh) for m in range(5)])
                off += 5 * sz
            halves.append(states)
        initial, final = halves
        for c, st in zip(blocks, initial):
            c.upload_state(st)
        with shockx.Blocks(blocks, _connections(recs, dims)) as grp:
            grp.transfer_metric()
            tau = 0.5 * min(c.cfl_dt(0.5) for c in blocks)
            grp.run(tau, steps)
            got = [c.download_state() for c in blocks]
    finally:
        for c in blocks:
            c.close()
    same = all(np.array_equal(g, f) for gz, fz in zip(got, final) for g, f in zip(gz, fz))
    g = (S + 1) // 2
    inner = (slice(g, -g),) * dims + ((slice(None),) if dims == 2 else ())
    moved = max(float(np.max(np.abs(f[inner] - i[inner]))) for fz, iz in zip(final, initial) for f, i in zip(fz, iz))
    print(f"[cgns] C driver final states bit-equal to the Python assembly: {same} (largest interior change "
          f"over the run {moved:.3e})", flush=True)
    assert same and np.isfinite(moved) and moved > 1e-6
