"""Manufactured-solution refinement ladders on the GPU path, FP64 and mixed forms (the
convergence part of the precision workload at sizes that run in seconds).

3D blocks of n^3 nodes with the reference's subsonic Navier-Stokes parameter set
(src/ManufacturedSolution.c:505-519), manufactured-solution faces on all six sides, the source
added every stage and mu = 1, marched with the CFL step from the exact field to t = 2 (steady):
the remaining error is the discretisation error.  The zeta-index quirk is off (with it the 3D
viscous residual does not converge, see shock_amd/ladder.py).  Asserted per scheme:
  * the observed order of L2(rho) between the rungs reaches the scheme's (> 4.5 for WENO5,
    > 8 for WENO9), in both forms;
  * the mixed form's error is within 2x of the FP64 form's on every rung.
"""
import pytest

from shock_amd import ladder, shockx

pytestmark = pytest.mark.gpu

RUNGS = {5: (17, 33), 9: (13, 17)}
MIN_ORDER = {5: 4.5, 9: 8.0}


@pytest.mark.parametrize("S", [5, 9])
def test_manufactured_ladder_order_and_mixed_error(S):
    rows = {}
    for prec in (shockx.FP64, shockx.MIXED):
        rows[prec] = []
        for n in RUNGS[S]:
            r = ladder.rung(n, S=S, precision=prec, t_end=2.0)
            print(f"[ladder] W{S} n={n} {r['precision']}: L2(rho) {r['full']['L2_rho']:.3e} "
                  f"Linf(rho) {r['full']['Linf_rho']:.3e} steps {r['steps']}", flush=True)
            rows[prec].append(r)
    for prec in (shockx.FP64, shockx.MIXED):
        order = ladder.orders(rows[prec])
        print(f"[ladder] W{S} {'fp64' if prec == shockx.FP64 else 'mixed'} observed order {order}", flush=True)
        assert min(order) > MIN_ORDER[S]
    for a, b in zip(rows[shockx.FP64], rows[shockx.MIXED]):
        ratio = b["full"]["L2_rho"] / a["full"]["L2_rho"]
        print(f"[ladder] W{S} n={a['n']} mixed / fp64 = {ratio:.4f}", flush=True)
        assert ratio <= 2.0
