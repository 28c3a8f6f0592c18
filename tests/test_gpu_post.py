"""Post-processing on the device against the oracle: shx_post (CalcValuesForPost with
CalcLambda2, src/Functions.c:3687-3761, 5846-6070), the device-resident film filled by
shx_simulate's sampling (WriteValuesFromUToFilm, src/Functions.c:2510-2553) and the probe
histories (src/Functions.c:5381-5417), including samples and rows rewound by a NaN rollback."""
import numpy as np
import pytest

from helpers import interior, setup_2d_mms, setup_3d_turb
from shock_amd import shockx
from test_gpu_simulate import _sim

pytestmark = pytest.mark.gpu


def _close(a, b, rtol, what):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    err = float(np.max(np.abs(a - b))) / scale
    print(f"[post] {what}: {err:.2e}", flush=True)
    assert err <= rtol, what


def _case3(o, nn, S):
    c = o.c
    return shockx.Case(dims=3, space_order=S, time_order=4, n=nn, gamma=c.gamma, mach=c.mach,
                       reynolds=c.reynolds, prandtl=c.prandtl, T0=c.T0)


@pytest.mark.parametrize("S", [5, 9])
def test_post_of_the_uploaded_state(S):
    o, coords, met, prim, nn = setup_3d_turb(n=12, S=S, seed=5)
    p, _ = o.eos_pc(met, prim)
    want = o.calc_post(met, prim, p)
    with _case3(o, nn, S) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        got = g.post()
    sl = interior(o, 3, nn)
    for m, name in enumerate(("Mach", "gradRho", "Lambda2")):
        _close(got[m][sl], want[m][sl], 1e-12, f"{name} S={S}")


def test_film_of_a_3d_run():
    o, coords, met, prim, nn = setup_3d_turb(n=12, S=5, seed=9)
    tau = o.cfl_dt(0.5, met, prim)
    s = _sim(tau, end=11, accelerator=0, samples=4, iterations_between_samples=2, start_sampling=3)
    want_state, want = o.simulate(met, prim, s)
    assert want["samples"] == 4
    with _case3(o, nn, 5) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.film_open(4)
        rep = g.simulate(0, 11, tau, accelerator=False, decel_factor=0.5, distance_nan=12, distance_forward=5,
                         distance_backward=4, samples=4, iterations_between_samples=2, start_sampling=3)
        assert rep["samples"] == g.film_count() == 4
        for k, (t_want, f_want) in enumerate(want["film"]):
            t_got, f_got = g.film_download(k)
            assert t_got == t_want
            for m, name in enumerate(("rho", "u", "v", "w", "p", "gradRho", "Mach", "Lambda2")):
                _close(f_got[m], f_want[m], 1e-11, f"sample {k} {name}")


def test_film_rewound_by_a_rollback():
    # an unstable tau: samples taken before a NaN are rewritten after the rollback
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=5)
    probes = [int(np.ravel_multi_index((o.g + i, o.g + j, 0), o.shape)) for i, j in ((0, 0), (13, 21), (31, 5))]
    s = _sim(6.0 * o.cfl_dt(0.5, met, prim), samples=6, iterations_between_samples=3, start_sampling=0,
             probes=probes)
    want_state, want = o.simulate(met, prim, s)
    assert want["resets"] >= 1
    c = o.c
    with shockx.Case(dims=2, space_order=5, time_order=4, n=nn, gamma=c.gamma, mach=c.mach,
                     reynolds=c.reynolds, prandtl=c.prandtl, T0=c.T0) as g:
        g.upload_metric(met)
        g.upload_state(prim)
        g.film_open(6)
        rep = g.simulate(s["start"], s["end"], s["tau"], accelerator=True, accel_factor=1.1, decel_factor=0.5,
                         number_resets=10, distance_nan=12, distance_forward=5, distance_backward=4, samples=6,
                         iterations_between_samples=3, start_sampling=0, probes=s["probes"])
        t_got, h_got = g.history(len(s["probes"]), s["end"] - s["start"])
        t_want, h_want = want["history"]
        assert np.array_equal(t_got, t_want)
        for m, name in enumerate(("p", "u", "v", "w")):
            _close(h_got[m], h_want[m], 1e-10, f"probe history {name}")
        print(f"[post] rollback run: oracle resets {want['resets']} samples {want['samples']}, "
              f"GPU resets {rep['resets']} samples {rep['samples']}", flush=True)
        assert rep["resets"] == want["resets"] and rep["samples"] == want["samples"]
        for k, (t_want, f_want) in enumerate(want["film"]):
            t_got, f_got = g.film_download(k)
            assert t_got == t_want
            for m, name in enumerate(("rho", "u", "v", "w", "p", "gradRho", "Mach")):
                _close(f_got[m], f_want[m], 1e-10, f"sample {k} {name}")
