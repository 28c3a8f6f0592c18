"""The CGNS mesh reader (compat_load_mesh: the mesh part of loadFile, src/Load.c:252-744) on the
reference's testcase meshes (tests/golden/cgns: the HDF5 files as the reference ships them, with
their division files in tests/golden/division), CPU only:

  * every division record's block reads, and the blocks tile the zone: the zone size in the file
    equals the division's bounding box, and each block equals the same range of the whole zone
    read at once;
  * the coordinates are the unit cube's with uniform spacing, x = (i-1)/(NI-1), y and z likewise,
    to float32 rounding, x depending on i alone: the file's i-fastest order lands on the library's
    i axis (the 3D testcase is 64 x 64 x 16, so a transposed read fails);
  * boundary conditions: the testcases' families resolved through FamilySpecified, BCInflow at
    i = 1, BCOutflow at i = NI and BCWallInviscid on the other faces, on exactly the faces without
    a neighbour;
  * refusals with the reference's messages: a cell dimension that is not the file's, a zone the
    file does not hold; and a block outside the zone, a file that is not HDF5, a truncated file.
"""
import os

import numpy as np
import pytest

from shock_amd import meshio, shockx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
TESTCASES = [("2d_cube_16x16", 2, 4), ("2d_cube_64x64", 2, 4), ("3d_cube_16x16x16", 3, 8),
             ("3d_cube_32x32x32", 3, 8)]


def _paths(name):
    return os.path.join(GOLD, "cgns", name + ".cgns"), os.path.join(GOLD, "division", name + ".sddc")


def _need_lib():
    if not os.path.exists(meshio.COMPAT_LIB):
        pytest.skip("libshock_compat.so not built")


@pytest.mark.parametrize("name,dims,nproc", TESTCASES, ids=[t[0] for t in TESTCASES])
def test_blocks_tile_the_zone_and_match_the_whole_read(name, dims, nproc):
    _need_lib()
    mesh, div = _paths(name)
    recs = [meshio.load_division(div, dims, nproc, z) for z in range(1, nproc + 1)]
    hi = [max(r.corners[1][k] for r in recs) for k in range(dims)]
    (x, y, z), info = meshio.load_mesh(mesh, dims, recs[0])
    assert [info.zone_size[k] for k in range(dims)] == hi
    assert info.celldim == dims and info.start_iteration == 0 and info.coord_real8 == 0
    whole = meshio.whole_zone(recs[0].zonename, hi)
    (X, Y, Z), winfo = meshio.load_mesh(mesh, dims, whole)
    for r in recs:
        (bx, by, bz), _ = meshio.load_mesh(mesh, dims, r)
        sl = tuple(slice(r.corners[0][k] - 1, r.corners[1][k]) for k in range(dims))
        for a, A in ((bx, X), (by, Y), (bz, Z)):
            assert np.array_equal(a, A[sl] if dims == 3 else A[sl + (slice(None),)])


@pytest.mark.parametrize("name,dims,nproc", TESTCASES, ids=[t[0] for t in TESTCASES])
def test_coordinates_are_the_uniform_unit_cube(name, dims, nproc):
    _need_lib()
    mesh, div = _paths(name)
    rec = meshio.load_division(div, dims, nproc, 1)
    (_, _, _), info = meshio.load_mesh(mesh, dims, rec)
    size = [info.zone_size[k] for k in range(dims)]
    (X, Y, Z), _ = meshio.load_mesh(mesh, dims, meshio.whole_zone(rec.zonename, size))
    axes = np.meshgrid(*[np.arange(s) / (s - 1) for s in size], indexing="ij")
    if dims == 2:
        axes = [a[:, :, None] for a in axes] + [np.zeros(X.shape)]
    for got, want in zip((X, Y, Z), axes):
        assert got.shape == want.shape
        assert float(np.max(np.abs(got - want))) <= 2e-7          # float32 storage
        assert np.array_equal(got, got.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("name,dims,nproc", TESTCASES, ids=[t[0] for t in TESTCASES])
def test_boundary_conditions_of_the_faces_without_neighbour(name, dims, nproc):
    _need_lib()
    mesh, div = _paths(name)
    for zrec in range(1, nproc + 1):
        r = meshio.load_division(div, dims, nproc, zrec)
        _, info = meshio.load_mesh(mesh, dims, r)
        bcs = meshio.bc_strings(info, dims)
        for f in range(2 * dims):
            if r.face[f].neighbour:
                assert bcs[f] == ""
                continue
            want = {0: "BCInflow", 1: "BCOutflow"}.get(f, "BCWallInviscid")
            assert bcs[f] == want, (zrec, f, bcs)
        kinds = meshio.face_kinds(info, dims, r)
        for f in range(2 * dims):
            if r.face[f].neighbour:
                assert kinds[f] == shockx.BC_INTERFACE
        assert shockx.BC_WALL_SLIP in kinds


@pytest.mark.parametrize("dims,S,n", [(3, 5, (7, 6, 5)), (3, 9, (12, 11, 10)), (2, 5, (9, 8, 1)), (2, 9, (14, 13, 1))])
def test_c_ghost_extrapolation_equals_the_tests_extension(dims, S, n):
    """compat_extend_coords (the C host's ExtrapolateGhostCells) gives, bit for bit, the extended
    coordinates of cases.ext_from_nodes, which every oracle parity test builds its metric from; on
    an affine grid the extrapolated layers are the grid's own continuation."""
    _need_lib()
    from shock_amd import cases
    rng = np.random.default_rng(sum(n) + S)
    xyz = [np.cumsum(rng.uniform(0.5, 1.5, size=n), axis=a) + 0.1 * rng.standard_normal(n) for a in range(3)]
    if dims == 2:
        xyz[2] = np.zeros(n)
    got = meshio.extend_coords(dims, S, xyz)
    want = cases.ext_from_nodes(dims, S, xyz)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    g = (S + 1) // 2 + 1
    ii = np.arange(-g, n[0] + g, dtype=np.float64)
    aff = [np.broadcast_to((0.25 * np.arange(n[0]))[:, None, None], n).copy()] + [np.zeros(n), np.zeros(n)]
    ext = meshio.extend_coords(dims, S, aff)
    assert np.array_equal(ext[0][:, 0, 0], 0.25 * ii)


def test_refusals(tmp_path):
    _need_lib()
    mesh, div = _paths("3d_cube_16x16x16")
    rec = meshio.load_division(div, 3, 8, 1)
    with pytest.raises(meshio.MeshError, match="Mesh file dimension 3 does not match compile time dimension 2"):
        meshio.load_mesh(mesh, 2, meshio.whole_zone(rec.zonename, (64, 64)))
    with pytest.raises(meshio.MeshError, match="Could not find associated zone 'nowhere'"):
        meshio.load_mesh(mesh, 3, meshio.whole_zone("nowhere", (64, 64, 16)))
    with pytest.raises(meshio.MeshError, match="outside the zone"):
        meshio.load_mesh(mesh, 3, meshio.whole_zone(rec.zonename, (64, 64, 17)))
    with pytest.raises(meshio.MeshError, match="not an HDF5 file"):
        meshio.load_mesh(div, 3, rec)
    raw = open(mesh, "rb").read()
    cut = tmp_path / "cut.cgns"
    cut.write_bytes(raw[:len(raw) // 2])           # the coordinates lie beyond the cut
    with pytest.raises(meshio.MeshError):
        meshio.load_mesh(str(cut), 3, rec)
    short = tmp_path / "short.cgns"
    short.write_bytes(raw[:200])                   # the root header is cut
    with pytest.raises(meshio.MeshError):
        meshio.load_mesh(str(short), 3, rec)
