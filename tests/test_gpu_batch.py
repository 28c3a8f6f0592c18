"""Batches: many independent cases advanced by shared launches from a captured graph must give
each case exactly the state it reaches alone (same kernels, same arithmetic), and match the oracle
to the single-case bar."""
import numpy as np
import pytest

import orc
from helpers import conserved, interior, rel_linf, setup_2d_mms, setup_3d_turb
from shock_amd import cases, shockx

pytestmark = pytest.mark.gpu


def _make(spec, time_order=4):
    dims, S, n, prec, inviscid, seed = spec
    if dims == 2:
        o, (xe, ye, ze), met, prim, nn = setup_2d_mms(n=n, S=S, inviscid=inviscid, time_order=time_order)
    else:
        o, (xe, ye, ze), met, prim, nn = setup_3d_turb(n=n, S=S, inviscid=inviscid, seed=seed,
                                                       time_order=time_order)
    c = shockx.Case(dims=dims, space_order=S, time_order=time_order, n=nn, gamma=1.4, mach=o.c.mach,
                    reynolds=o.c.reynolds, inviscid=inviscid, precision=prec)
    c.build_metric(xe, ye, ze)
    c.upload_state(prim)
    return o, c, met, prim, nn


SPECS = [(2, 5, 24, shockx.FP64, False, 1), (3, 5, 12, shockx.FP64, False, 2), (2, 9, 20, shockx.FP64, False, 3),
         (3, 5, 16, shockx.FP64, False, 4), (3, 9, 12, shockx.FP64, False, 5), (3, 5, 14, shockx.MIXED, False, 6),
         (2, 5, 28, shockx.FP64, True, 7), (3, 5, 12, shockx.FP64, False, 8)]


def test_batch_equals_cases_alone_and_oracle():
    made = [_make(s) for s in SPECS]
    taus = [c.cfl_dt(0.4) for _, c, _, _, _ in made]
    steps = 3
    alone = []
    for (o, c, met, prim, nn), tau in zip(made, taus):
        c.run(tau, steps)
        alone.append(c.download_state())
        c.upload_state(prim)
    with shockx.Batch([m[1] for m in made], taus) as b:
        print("[batch] tuned factors", b.tune(1), flush=True)
        b.run(steps)
        b.synchronize()
        for k, ((o, c, met, prim, nn), tau, ref) in enumerate(zip(made, taus, alone)):
            got = c.download_state()
            sl = interior(o, o.c.dims, nn)
            diff = max(float(np.max(np.abs(a[sl] - r[sl]))) for a, r in zip(got, ref))
            print(f"[batch] case {k} {SPECS[k]}: max |batch - alone| = {diff:.2e}", flush=True)
            assert diff == 0.0
            if SPECS[k][3] == shockx.FP64:
                want = o.run(tau, met, prim, steps)
                errs = rel_linf(conserved(got), conserved(want), sl)
                if o.c.dims == 2:
                    errs = errs[:3] + errs[4:]
                print(f"[batch] case {k} vs oracle: " + " ".join(f"{e:.1e}" for e in errs), flush=True)
                assert max(errs) <= 1e-12
        # a batched case cannot be stepped alone (its state would leave the batch's buffers)
        with pytest.raises(shockx.ShxError, match="in a batch"):
            made[0][1].run(taus[0], 1)
        # a second run continues from the batch's own state
        b.set_tau([t * 0.5 for t in taus])
        b.run(1)
        b.synchronize()
    for _, c, _, _, _ in made:
        c.close()


def test_batch_with_ssp_rk3_cases():
    """SSP-RK3 cases (the stage update reads the stage input U_s as well as U_n) batched beside RK4
    cases: every case ends bit for bit where it ends alone, the FP64 RK3 cases within 1e-12 of the
    oracle after three steps."""
    specs = [((2, 5, 24, shockx.FP64, False, 11), 3), ((3, 5, 12, shockx.FP64, False, 12), 3),
             ((3, 9, 12, shockx.FP64, False, 13), 3), ((3, 5, 14, shockx.MIXED, False, 14), 3),
             ((2, 5, 20, shockx.FP64, True, 15), 3), ((3, 5, 12, shockx.FP64, False, 16), 4)]
    made = [_make(sp, to) for sp, to in specs]
    taus = [c.cfl_dt(0.4) for _, c, _, _, _ in made]
    steps = 3
    alone = []
    for (o, c, met, prim, nn), tau in zip(made, taus):
        c.run(tau, steps)
        alone.append(c.download_state())
        c.upload_state(prim)
    with shockx.Batch([m[1] for m in made], taus) as b:
        b.run(steps)
        b.synchronize()
        for k, ((o, c, met, prim, nn), tau, ref) in enumerate(zip(made, taus, alone)):
            got = c.download_state()
            sl = interior(o, o.c.dims, nn)
            assert all(np.array_equal(a[sl], r[sl]) for a, r in zip(got, ref)), (k, specs[k])
            if specs[k][0][3] == shockx.FP64:
                want = o.run(tau, met, prim, steps)
                errs = rel_linf(conserved(got), conserved(want), sl)
                if o.c.dims == 2:
                    errs = errs[:3] + errs[4:]
                print(f"[batch] RK{specs[k][1]} case {k} {specs[k][0]} vs oracle: "
                      + " ".join(f"{e:.1e}" for e in errs), flush=True)
                assert max(errs) <= 1e-12
    for _, c, _, _, _ in made:
        c.close()


def test_batch_of_many_uneven_cases_equals_cases_alone():
    """28 cases of uneven shape from a seed, as the batch workload draws them (2D and 3D, WENO5 and
    WENO9, FP64 and mixed, viscous and inviscid, 10^2 .. 26^3 nodes): every case in the batch ends
    bit for bit where it ends alone, with persistent blocks tuned for every kernel class."""
    rng = np.random.default_rng(20251203)
    specs = []
    for k in range(28):
        dims = int(rng.choice([2, 3]))
        S = int(rng.choice([5, 9]))
        n = int(rng.integers(S + 5, 41)) if dims == 2 else int(rng.integers(S + 3, 27))
        prec = shockx.MIXED if rng.uniform() < 0.25 else shockx.FP64
        inviscid = bool(rng.uniform() < 0.2)
        specs.append((dims, S, n, prec, inviscid, 100 + k))
    made = [_make(s) for s in specs]
    taus = [c.cfl_dt(0.4) for _, c, _, _, _ in made]
    steps = 2
    alone = []
    for (o, c, met, prim, nn), tau in zip(made, taus):
        c.run(tau, steps)
        alone.append(c.download_state())
        c.upload_state(prim)
    groups = len({(s[0], s[1], s[3], s[4]) for s in specs})
    with shockx.Batch([m[1] for m in made], taus) as b:
        print(f"[batch] {len(specs)} cases in {groups} template groups; tuned {b.tune(1)}", flush=True)
        b.run(steps)
        b.synchronize()
        for k, ((o, c, met, prim, nn), ref) in enumerate(zip(made, alone)):
            got = c.download_state()
            sl = interior(o, o.c.dims, nn)
            assert all(np.array_equal(a[sl], r[sl]) for a, r in zip(got, ref)), (k, specs[k])
    for _, c, _, _, _ in made:
        c.close()


def test_batch_reports_the_case_whose_state_goes_non_finite():
    """checkNAN fused into every case's last stage: a NaN planted in one case's state is reported by
    synchronize (ERR_NAN naming the case; its count in nan_counts, the others zero), that case ends
    where shx_run leaves its twin (the step with the non-finite nodes set to 1e6), the other cases
    end where they end alone, and a second synchronize reports nothing new."""
    specs = [(2, 5, 24, shockx.FP64, False, 1), (3, 5, 12, shockx.FP64, False, 2), (2, 9, 20, shockx.FP64, False, 3)]
    made = [_make(sp) for sp in specs]
    taus = [c.cfl_dt(0.4) for _, c, _, _, _ in made]
    o1, c1, _, prim1, nn1 = made[1]
    bad = [a.copy() for a in prim1]
    g = o1.g
    bad[0][g + 4, g + 5, g + 6] = np.nan
    o1.ghost_periodic(bad)
    c1.upload_state(bad)
    alone = []
    for k, ((o, c, met, prim, nn), tau) in enumerate(zip(made, taus)):
        if k == 1:
            twin = shockx.Case(dims=3, space_order=5, time_order=4, n=nn1, gamma=1.4, mach=o1.c.mach,
                               reynolds=o1.c.reynolds)
            twin.upload_metric(c1.download_metric())
            twin.upload_state(bad)
            with pytest.raises(shockx.ShxError) as ei:
                twin.run(tau, 1)
            assert ei.value.code == shockx.ERR_NAN
            alone.append(twin.download_state())
            twin.close()
        else:
            c.run(tau, 1)
            alone.append(c.download_state())
            c.upload_state(prim)
    with shockx.Batch([m[1] for m in made], taus) as b:
        b.run(1)
        with pytest.raises(shockx.ShxError) as ei:
            b.synchronize()
        assert ei.value.code == shockx.ERR_NAN and "case 1:" in str(ei.value)
        print(f"[batch] {ei.value}", flush=True)
        counts = b.nan_counts()
        assert counts[1] > 1 and counts[0] == 0 and counts[2] == 0
        b.synchronize()                      # nothing new since the last call
        for k, ((o, c, met, prim, nn), ref) in enumerate(zip(made, alone)):
            got = c.download_state()
            sl = interior(o, o.c.dims, nn)
            assert all(np.array_equal(a[sl], r[sl], equal_nan=True) for a, r in zip(got, ref)), k
    for _, c, _, _, _ in made:
        c.close()
