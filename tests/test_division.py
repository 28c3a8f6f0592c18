"""The division-file reader (compat_load_division, loadDivision of src/Load.c:191-250) on the
reference's testcase division files (tests/golden/division/*.sddc), CPU only:

  * every record decodes to what an independent decode of the documented layout gives
  * the zones tile their bounding box exactly (volumes add up, no two overlap)
  * interfaces are mutual: zone A's face (d, side) names zone B, whose opposite face names A, with
    range and donor range exchanged, and the range lies on A's face in A's local indices
  * a process count that does not match the file size is rejected
"""
import ctypes
import glob
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "shock_amd", "lib", "libshock_compat.so")
FILES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "division", "*.sddc")))


class Interface(ctypes.Structure):
    _fields_ = [("neighbour", ctypes.c_int), ("transform", ctypes.c_int * 3),
                ("range", (ctypes.c_longlong * 3) * 2), ("drange", (ctypes.c_longlong * 3) * 2),
                ("rc", ctypes.c_float * 3), ("ra", ctypes.c_float * 3), ("tr", ctypes.c_float * 3)]


class Division(ctypes.Structure):
    _fields_ = [("zonename", ctypes.c_char * 33), ("corners", (ctypes.c_longlong * 3) * 2),
                ("face", Interface * 6), ("ni", ctypes.c_int)]


def _lib():
    if not os.path.exists(LIB):
        pytest.skip("libshock_compat.so not built")
    L = ctypes.CDLL(LIB)
    L.compat_load_division.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_int,
                                       ctypes.POINTER(Division)]
    return L


def _layout(D):
    ifb = 4 + 4 * D + 2 * D * 8 * 2 + 3 * 4 * 3
    return ifb, 32 + 8 * D * 2 + ifb * 2 * D


def _decode(buf, D):
    """independent decode of one record (little-endian)"""
    ifb, _ = _layout(D)
    lo_hi = struct.unpack_from("<%dq" % (2 * D), buf, 32)
    faces = []
    for f in range(2 * D):
        off = 32 + 16 * D + f * ifb
        nb = struct.unpack_from("<i", buf, off)[0]
        tr = struct.unpack_from("<%di" % D, buf, off + 4)
        rg = struct.unpack_from("<%dq" % (4 * D), buf, off + 4 + 4 * D)
        faces.append((nb, tr, rg))
    return buf[:32].rstrip(b"\0"), lo_hi, faces


def _read_all(path):
    D = 3 if os.path.basename(path).startswith("3d") else 2
    _, dl = _layout(D)
    size = os.path.getsize(path)
    nproc = size // dl
    L = _lib()
    out = []
    for z in range(1, nproc + 1):
        d = Division()
        assert L.compat_load_division(path.encode(), D, nproc, z, 0, ctypes.byref(d)) == 0
        out.append(d)
    return D, nproc, out


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_records_match_an_independent_decode(path):
    D, nproc, zones = _read_all(path)
    raw = open(path, "rb").read()
    _, dl = _layout(D)
    for z, d in enumerate(zones):
        name, lo_hi, faces = _decode(raw[z * dl:(z + 1) * dl], D)
        assert d.zonename == name
        assert [d.corners[h][k] for h in range(2) for k in range(D)] == list(lo_hi)
        assert d.ni == sum(1 for nb, _, _ in faces if nb)
        for f, (nb, tr, rg) in enumerate(faces):
            it = d.face[f]
            assert it.neighbour == nb
            if nb:
                assert tuple(it.transform[:D]) == tr
                got = [it.range[0][k] for k in range(D)] + [it.range[1][k] for k in range(D)] + \
                      [it.drange[0][k] for k in range(D)] + [it.drange[1][k] for k in range(D)]
                assert got == list(rg)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_zones_tile_the_block_and_interfaces_are_mutual(path):
    D, nproc, zones = _read_all(path)
    lo = np.array([[d.corners[0][k] for k in range(D)] for d in zones])
    hi = np.array([[d.corners[1][k] for k in range(D)] for d in zones])
    vol = np.prod(hi - lo + 1, axis=1)
    box = np.prod(hi.max(0) - lo.min(0) + 1)
    assert vol.sum() == box
    for a in range(nproc):
        for b in range(a + 1, nproc):
            overlap = np.all(np.minimum(hi[a], hi[b]) >= np.maximum(lo[a], lo[b]))
            assert not overlap
    for a, d in enumerate(zones):
        n_local = hi[a] - lo[a] + 1
        for f in range(2 * D):
            it = d.face[f]
            if not it.neighbour:
                continue
            b = it.neighbour - 1
            dirn, side = f // 2, f % 2
            back = zones[b].face[2 * dirn + (1 - side)]
            assert back.neighbour == a + 1
            for k in range(D):
                assert back.range[0][k] == it.drange[0][k] and back.range[1][k] == it.drange[1][k]
                assert back.drange[0][k] == it.range[0][k] and back.drange[1][k] == it.range[1][k]
            plane = n_local[dirn] if side else 1
            assert it.range[0][dirn] == plane and it.range[1][dirn] == plane


def test_wrong_process_count_is_rejected():
    L = _lib()
    d = Division()
    assert L.compat_load_division(FILES[0].encode(), 2, 3, 1, 0, ctypes.byref(d)) == -1


class Connection(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int), ("face", ctypes.c_int), ("donor", ctypes.c_int),
                ("donor_face", ctypes.c_int), ("transform", ctypes.c_int * 3)]


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_division_connections_for_blocks_on_one_device(path):
    """compat_division_connections turns every zone's records into the shx_blocks connections: one
    per face with a neighbour, joined to the neighbour's face that holds the donor range, with a
    transform that takes the face's normal to the donor face's, is the inverse of the donor side's,
    and maps the tangential extents onto equal ones (whole faces)."""
    D, nproc, zones = _read_all(path)
    L = _lib()
    L.compat_division_connections.argtypes = [ctypes.POINTER(Division), ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(Connection)]
    arr = (Division * nproc)(*zones)
    out = (Connection * (2 * D * nproc))()
    n = L.compat_division_connections(arr, nproc, D, out)
    assert n == sum(d.ni for d in zones)
    ext = [[zones[z].corners[1][k] - zones[z].corners[0][k] + 1 for k in range(D)] for z in range(nproc)]
    conns = {(c.block, c.face): c for c in out[:n]}
    assert len(conns) == n
    for (b, f), c in conns.items():
        a, ad = f // 2, c.donor_face // 2
        t = list(c.transform[:D])
        assert sorted(abs(x) for x in t) == list(range(1, D + 1))
        assert abs(t[a]) - 1 == ad
        back = conns[(c.donor, c.donor_face)]
        assert back.donor == b and back.donor_face == f
        tb = list(back.transform[:D])
        for q in range(D):
            aq = abs(t[q]) - 1
            assert abs(tb[aq]) - 1 == q
            if q != a:
                assert (t[q] > 0) == (tb[aq] > 0)
                assert ext[b][q] == ext[c.donor][aq]
