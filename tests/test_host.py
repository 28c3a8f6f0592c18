"""Host-side checks that need no GPU: the C ABI library loads and exports every symbol the
header declares; the case builders produce what the workloads specify."""
import os
import re
import subprocess

import numpy as np
import pytest

from shock_amd import cases, shockx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shockx.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(shx_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_python_mirror_agree():
    assert declared_symbols() == sorted(shockx.EXPORTS)


@pytest.mark.skipif(not os.path.exists(shockx.LIB_PATH), reason="libshockx.so not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", shockx.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (shx_[a-z_0-9]+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    lib = shockx.lib()          # loads without a GPU; no compute call is made here
    for s in declared_symbols():
        assert hasattr(lib, s)


COMPAT_NAMES = ["CalcRungeKutta", "checkNAN", "WriteValuesFromU1ToU2", "compat_define", "compat_release"]


def test_compat_header_declares_the_reference_names():
    txt = open(os.path.join(ROOT, "shock_amd", "compat", "shock_compat.h")).read()
    for name in COMPAT_NAMES:
        assert re.search(r"\b" + name + r"\s*\(", txt), name
    # the reference's argument order of CalcRungeKutta (src/Functions.h:206-214)
    sig = re.search(r"void CalcRungeKutta\(([^;]*)\);", txt).group(1)
    kinds = re.findall(r"struct (strct_[a-zA-Z]+)", sig)
    assert kinds == ["strct_configuration", "strct_mesh", "strct_U", "strct_U", "strct_Flux", "strct_Flux",
                     "strct_Flux", "strct_Flux"]


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "shock_amd", "lib", "libshock_compat.so")),
                    reason="libshock_compat.so not built")
def test_compat_library_exports_the_reference_names():
    path = os.path.join(ROOT, "shock_amd", "lib", "libshock_compat.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (\w+)", out))
    assert all(n in exported for n in COMPAT_NAMES), exported


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(shockx, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(shockx, "_lib", None)
    with pytest.raises(shockx.ShxError):
        shockx.lib()


def test_padded_shape_and_ghost_width():
    assert cases.padded_shape(3, 5, (256, 256, 256)) == ((262, 262, 262), 3)
    assert cases.padded_shape(3, 9, (384, 384, 384)) == ((394, 394, 394), 5)
    assert cases.padded_shape(2, 5, (64, 64, 1)) == ((70, 70, 1), 3)


def test_periodic_coordinates_wrap_with_translation():
    n = (8, 6, 5)
    xe, ye, ze = cases.periodic_ext_coords(3, 5, n, cases.sine_bump_3d(0.05))
    g = 3
    # extended index e <-> padded e-1; interior padded g..g+n-1; ghost padded g-1 = interior last - 1.0
    assert np.allclose(xe[g, 2:-2, 2:-2], xe[g + n[0], 2:-2, 2:-2] - 1.0, atol=0)
    assert np.array_equal(ye[:, g + 1, :], ye[:, g + 1 + n[1], :] - 1.0)
    assert np.array_equal(ze[:, :, 1], ze[:, :, 1 + n[2]] - 1.0)
    # interior nodes follow the mapping exactly
    xi = np.arange(n[0]) / n[0]
    assert np.allclose(xe[g + 1:g + 1 + n[0], g + 1, g + 1], xi + 0.0)   # eta = zeta = 0 row: bump vanishes


def test_turbulence_field_is_solenoidal_with_requested_mach():
    n = 32
    u, v, w = cases.random_phase_velocity(n, 11, mt=0.5, c=2.0)
    rms = np.sqrt(np.mean(u * u + v * v + w * w))
    assert abs(rms - 1.0) < 1e-12
    k = np.fft.fftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    div = KX * np.fft.fftn(u) + KY * np.fft.fftn(v) + KZ * np.fft.fftn(w)
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(np.fft.fftn(u)))
    # same seed, same field
    u2, _, _ = cases.random_phase_velocity(n, 11, mt=0.5, c=2.0)
    assert np.array_equal(u, u2)


def test_manufactured_state_in_shock_units():
    x = np.linspace(0, 1, 9)[:, None, None] * np.ones((1, 9, 1))
    y = np.transpose(x, (1, 0, 2))
    rho, u, v, w, e = cases.mms_periodic_2d_state(x, y, gamma=1.4, mach=0.5)
    ups = 1.0 / (1.4 * 0.25)
    p = (e - 0.5 * (u * u + v * v)) * (0.4 * rho) / ups
    # p_S = gamma * (1/gamma + 0.05 cos(2 pi (x+y)))
    assert np.allclose(p, 1.0 + 1.4 * 0.05 * np.cos(2 * np.pi * (x + y)))
    assert np.allclose(u * 0.5, 0.5 + 0.1 * np.cos(2 * np.pi * x) * np.sin(2 * np.pi * y))


def test_c_driver_compiles_against_the_header(tmp_path):
    from shock_amd import build
    out = tmp_path / "drive.o"
    subprocess.run(["gcc", "-c", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    build.DRIVER_SRC, "-o", str(out)], check=True)
    assert out.exists()


def _compat_case(path, faces, n=(8, 8, 1), dims=2):
    """A compat_run case file (shock_amd/compat/compat_run.c) with zero metric and state: only the
    configuration is read before the checks under test end the run."""
    ng = [n[0] + 6, n[1] + 6, 1 if dims == 2 else n[2] + 6]
    hdr = np.array([dims, 5, n[0], n[1], n[2], 4, 0, 1] + list(faces), dtype=np.int32)
    par = np.array([1.4, 0.5, 1000.0, 0.72, 300.0, 1e-4, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0])
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(par.tobytes())
        f.write(np.zeros(15 * int(np.prod(ng))).tobytes())


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "shock_amd", "compat", "compat_run")),
                    reason="compat_run not built")
@pytest.mark.parametrize("faces,donor,msg", [
    ((0, 0, 0, 0, 1, 1), "1", "interface to rank 1"),        # a neighbour that is another rank
    ((0, 2, 0, 0, 1, 1), None, "face right is not"),         # one face of the pair is a boundary
])
def test_compat_rejects_interfaces_that_are_not_periodic_self_interfaces(tmp_path, faces, donor, msg):
    # the configuration check runs before any device call, so this needs no GPU
    from shock_amd import build
    case = tmp_path / "case.bin"
    _compat_case(case, faces)
    cmd = [build.COMPAT_RUN, str(case), str(tmp_path / "out.bin")] + ([donor] if donor else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 4242 % 256, (r.returncode, r.stdout, r.stderr)
    assert msg in r.stdout
