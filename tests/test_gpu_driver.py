"""The C host loop (shock_amd/host/shockx_drive.c) against the oracle on the same case."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import interior, setup_2d_mms
from shock_amd import build

pytestmark = pytest.mark.gpu


def test_c_driver_matches_oracle():
    exe = build.DRIVER
    assert os.path.exists(exe), "build the driver with python -m shock_amd.build"
    r = subprocess.run([exe, "32", "10"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    m = re.search(r"step   10  mass (\S+)  energy (\S+)", r.stdout)
    mass, energy = float(m.group(1)), float(m.group(2))
    tau = float(re.search(r"tau=(\S+)", r.stdout).group(1))
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=5)
    assert abs(tau - o.cfl_dt(0.5, met, prim)) <= 1e-15 * tau
    out = o.run(tau, met, prim, 10)
    sl = interior(o, 2, nn)
    ref_mass = np.mean(out[0][sl])
    ref_energy = np.mean((out[0] * out[4])[sl])
    assert abs(mass - ref_mass) <= 1e-13 * abs(ref_mass)
    assert abs(energy - ref_energy) <= 1e-13 * abs(ref_energy)
