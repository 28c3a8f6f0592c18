"""Pins of the reference choices that the invariants of the other oracle tests leave open.

Three choices of the reference change results but pass every conservation, symmetry, order and
Cartesian free-stream check (tests/pins.py states why each check below decides them):
  * the Lax-Friedrichs weight J_{i+1/2} on every stencil node (src/WENO.c:314, 353-383),
  * the arithmetic-mean eigen-state of an interface (src/WENO.c:1104-1109),
  * the face coefficient (mu_i fac_i + mu_{i+1} fac_{i+1}) / 2 of the viscous fluxes
    (src/ZD.c:122-132).
The mutation tests compile altered copies of oracle/shx_oracle.c, each with the plausible
alternative choice planted (node J, RMS sound speed, density-weighted velocity, averaged mu times
averaged factor), and check that the figure of the matching pin fails its bar on that copy.
"""
import os
import re
import shutil
import subprocess
import sys

import pytest

import pins

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

FREE_STREAM_BAR = 1e-11     # |Q| / J of a uniform stream (the real oracle: <= 3e-13)
CHAR_FLUX_BAR = 1e-12       # relative to the interface flux (the real oracle: <= 2e-14)
MU_LOCAL_BAR = 1e-12        # relative to the response (the real oracle: <= 8e-15)


@pytest.mark.parametrize("dims,S", [(2, 5), (2, 9), (3, 5), (3, 9)])
def test_free_stream_on_stretched_grids(dims, S):
    assert pins.free_stream_stretched(dims, S) < FREE_STREAM_BAR


@pytest.mark.parametrize("S", [5, 9])
def test_interface_flux_is_characteristic_wrt_arithmetic_mean_state(S):
    assert pins.char_flux_1d(S) < CHAR_FLUX_BAR


@pytest.mark.parametrize("dims,dr", [(2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])
def test_viscous_mu_enters_through_its_own_node_factor(dims, dr):
    dev, rel = pins.mu_locality(dims, dr)
    assert rel > 1e-3          # the change of mu does move Q
    assert dev < MU_LOCAL_BAR


# ---- mutation tests ---------------------------------------------------------------------------
# (pattern, replacement, expected number of replacements) applied to shx_oracle.c
MUTANTS = {
    "lf_node_J": [(r"\[npl\] \* Jh\)", "[npl] * met[9][npl])", 5),
                  (r"\[nmi\] \* Jh\)", "[nmi] * met[9][nmi])", 5)],
    "rms_sound_speed": [(r"cs = 0\.5 \* \(der\[2\]\[n\] \+ der\[2\]\[np1\]\);",
                         "cs = sqrt(0.5 * (der[2][n] * der[2][n] + der[2][np1] * der[2][np1]));", 1)],
    "density_weighted_u": [(r"u = 0\.5 \* \(prim\[1\]\[n\] \+ prim\[1\]\[np1\]\);",
                            "u = (prim[0][n] * prim[1][n] + prim[0][np1] * prim[1][np1])"
                            " / (prim[0][n] + prim[0][np1]);", 1)],
    "mu_fac_of_means": [(r"Km\[l\] = 0\.5 \* \(der\[3\]\[nm\] \* fm\[l\] \+ der\[3\]\[n\] \* f0\[l\]\);",
                         "Km[l] = 0.25 * (der[3][nm] + der[3][n]) * (fm[l] + f0[l]);", 1),
                        (r"Kp\[l\] = 0\.5 \* \(der\[3\]\[n\] \* f0\[l\] \+ der\[3\]\[np\] \* fp\[l\]\);",
                         "Kp[l] = 0.25 * (der[3][n] + der[3][np]) * (f0[l] + fp[l]);", 1)],
}

# the pin each mutant must fail, as a Python expression over tests/pins.py, and its bar
CHECKS = {
    "lf_node_J": ("pins.free_stream_stretched(2, 5)", FREE_STREAM_BAR),
    "rms_sound_speed": ("pins.char_flux_1d(5)", CHAR_FLUX_BAR),
    "density_weighted_u": ("pins.char_flux_1d(5)", CHAR_FLUX_BAR),
    "mu_fac_of_means": ("pins.mu_locality(3, 0)[0]", MU_LOCAL_BAR),
}


def _build_mutant(name, tmp):
    src = open(os.path.join(ORACLE, "shx_oracle.c")).read()
    for pat, rep, count in MUTANTS[name]:
        src, k = re.subn(pat, rep, src)
        assert k == count, (name, pat, k)
    shutil.copy(os.path.join(ORACLE, "shx_oracle.h"), tmp)
    with open(os.path.join(tmp, "shx_oracle.c"), "w") as f:
        f.write(src)
    lib = os.path.join(tmp, "liboracle_%s.so" % name)
    subprocess.run(["gcc", "-O2", "-std=gnu99", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                    "-o", lib, os.path.join(tmp, "shx_oracle.c"), "-lm"], check=True)
    return lib


def _figure(lib, expr):
    env = dict(os.environ, SHX_ORACLE_LIB=lib, PYTHONPATH=os.pathsep.join(
        [ROOT, ORACLE, os.path.join(ROOT, "tests")]))
    code = "import pins; print(repr(float(%s)))" % expr
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, check=True)
    return float(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_the_alternative_choice(name, tmp_path):
    lib = _build_mutant(name, str(tmp_path))
    expr, bar = CHECKS[name]
    fig = _figure(lib, expr)
    assert fig > 100 * bar, (name, fig)
