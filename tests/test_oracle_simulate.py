"""Pin the oracle's restatement of startSimulation's iteration loop (src/SHOCK.c:531-691): the NaN
rollback and the time-step control, against what the loop's rules fix in closed form.

  * without NaN and without the accelerator the loop is plain stepping: the state equals the
    oracle's run() bit for bit, tau never changes and time is the sum of the steps
  * with the accelerator on a stable run, tau grows by the factor after distance_forward steps'
    worth of time, then after four times as many, and so on
  * a NaN in the starting state comes back after every rollback (the first backup is that
    state): the loop takes number_resets rollbacks, each to iteration 1, and stops at the next
    NaN, with tau multiplied by the decelerator once per rollback
  * an unstable tau recovers: rollbacks halve tau until the run is stable, and the state that
    the loop ends with is finite
  * probe histories hold each iteration's state at the probe nodes under iteration - (start + 1)
"""
import math

import numpy as np

import orc
from helpers import interior, setup_2d_mms


def _sim(**kw):
    base = dict(start=0, end=12, tau=1.0, time=0.0, time_last_action=0.0, L0=1.0, u0=1.0, accelerator=0,
                accel_factor=1.1, decel_factor=0.7, number_resets=10, distance_nan=5000, distance_forward=1000,
                distance_backward=200, probes=())
    base.update(kw)
    return base


def test_plain_loop_is_plain_stepping():
    o, coords, met, prim, nn = setup_2d_mms(n=16, S=5)
    tau = o.cfl_dt(0.5, met, prim)
    st, rep = o.simulate(met, prim, _sim(end=7, tau=tau, distance_backward=3))
    ref = o.run(tau, met, prim, 7)
    for a, b in zip(st, ref):
        assert np.array_equal(a, b)
    assert rep["resets"] == rep["nan_steps"] == rep["accelerations"] == rep["tau_resets"] == 0
    assert rep["iteration"] == 7 and rep["tau"] == tau
    t = 0.0
    for _ in range(7):
        t += tau * 1.0 / 1.0
    assert rep["time"] == t


def test_accelerator_schedule():
    o, coords, met, prim, nn = setup_2d_mms(n=16, S=5)
    tau0 = 0.25 * o.cfl_dt(0.5, met, prim)
    d, f = 2, 1.1
    st, rep = o.simulate(met, prim, _sim(end=40, tau=tau0, accelerator=1, accel_factor=f, distance_forward=d,
                                         distance_backward=4))
    # closed form of the rule: accelerate at the first iteration whose elapsed time since the last
    # action exceeds dfwd * tau, then dfwd *= 4; only while iteration + distance_backward < end
    tau, t, last, dfwd, n_acc = tau0, 0.0, 0.0, d, 0
    for it in range(1, 41):
        t += tau
        if t - last > dfwd * tau and it + 4 < 40:
            dfwd, tau, last, n_acc = dfwd * 4, tau * f, t, n_acc + 1
    assert n_acc == 2                              # after 3 steps, then after 9 more
    assert rep["accelerations"] == n_acc and rep["tau"] == tau and rep["resets"] == 0
    assert math.isclose(rep["time"], t, rel_tol=0, abs_tol=0)


def test_nan_in_the_starting_state_exhausts_the_resets():
    o, coords, met, prim, nn = setup_2d_mms(n=16, S=5)
    tau0 = o.cfl_dt(0.5, met, prim)
    bad = [a.copy() for a in prim]
    bad[0][o.g + 3, o.g + 5, 0] = np.nan
    k = 4
    st, rep = o.simulate(met, bad, _sim(end=30, tau=tau0, number_resets=k, decel_factor=0.5))
    assert rep["aborted"] == 1 and rep["resets"] == k and rep["nan_steps"] == k + 1
    # each rollback returns to the first backup, iteration 1, so the loop fails at iteration 2
    assert rep["iteration"] == 2
    assert rep["tau"] == tau0 * 0.5 ** k
    # the stopped state is the failed step's, with checkNAN's marks
    sl = interior(o, 2, nn)
    assert np.any(st[0][sl] == 1e6)


def test_unstable_tau_recovers_by_rollback():
    o, coords, met, prim, nn = setup_2d_mms(n=32, S=5)
    tau0 = 6.0 * o.cfl_dt(0.5, met, prim)
    st, rep = o.simulate(met, prim, _sim(end=40, tau=tau0, decel_factor=0.5, accelerator=1, distance_nan=12,
                                         distance_forward=5, distance_backward=4))
    assert rep["aborted"] == 0 and rep["resets"] >= 1 and rep["nan_steps"] == rep["resets"]
    assert rep["iteration"] == 40
    assert all(np.all(np.isfinite(a)) for a in st)


def test_probe_history_records_each_iteration():
    # history row k holds u, v, w of iteration k + 1's state (plain stepping here) and its time
    o, coords, met, prim, nn = setup_2d_mms(n=16, S=5)
    tau = o.cfl_dt(0.5, met, prim)
    probes = [int(np.ravel_multi_index((o.g + 3, o.g + 7, 0), o.shape)),
              int(np.ravel_multi_index((o.g + 11, o.g + 2, 0), o.shape))]
    st, rep = o.simulate(met, prim, _sim(end=5, tau=tau, probes=probes))
    t, hist = rep["history"]
    for k in range(5):
        ref = o.run(tau, met, prim, k + 1)
        for q, n in enumerate(probes):
            idx = np.unravel_index(n, o.shape)
            for m in range(3):
                assert hist[1 + m, q, k] == ref[1 + m][idx]
    assert np.all(np.diff(t) > 0) and t[-1] == rep["time"]


def test_restart_history_skips_iterations_before_the_start():
    # a restart at iteration 5 with a NaN in the state: every rollback goes to the first backup,
    # iteration 1, below the first history row (the reference writes such rows out of bounds);
    # none is recorded, and a restart without NaN fills rows 0 .. end - start - 1
    o, coords, met, prim, nn = setup_2d_mms(n=16, S=5)
    tau = o.cfl_dt(0.5, met, prim)
    probes = [int(np.ravel_multi_index((o.g + 3, o.g + 7, 0), o.shape))]
    bad = [a.copy() for a in prim]
    bad[0][o.g + 2, o.g + 2, 0] = np.nan
    st, rep = o.simulate(met, bad, _sim(start=5, end=12, tau=tau, number_resets=3, probes=probes))
    assert rep["aborted"] == 1 and rep["resets"] == 3 and rep["iteration"] == 2
    t, hist = rep["history"]
    assert t.shape == (7,) and np.all(t == 0.0) and np.all(hist == 0.0)
    st, rep = o.simulate(met, prim, _sim(start=5, end=9, tau=tau, probes=probes))
    t, hist = rep["history"]
    idx = np.unravel_index(probes[0], o.shape)
    for k in range(4):
        ref = o.run(tau, met, prim, k + 1)
        assert hist[1, 0, k] == ref[1][idx]
    assert np.all(t > 0)
