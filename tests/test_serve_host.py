"""Scheduler of the solve server, on the host with stand-in cases: admission by footprint and
slots, every request takes exactly its own number of steps, retirement frees memory, a request
whose state goes non-finite ends as failed while the others finish."""
import numpy as np

from shock_amd import serve, shockx


class FakeCase:
    def __init__(self, r, log):
        self.r, self.log, self.steps = r, log, 0

    def synchronize(self):
        pass

    def download_state(self):
        return [self.steps]

    def close(self):
        self.log.append(("close", self.r.rid))


class FakeBatch:
    resident_peak = 0

    def __init__(self, cs, taus):
        self.cs = cs
        FakeBatch.resident_peak = max(FakeBatch.resident_peak, len(cs))

    def run(self, k):
        for c in self.cs:
            c.steps += k

    def synchronize(self):
        if any(getattr(c, "blow_at", None) is not None and c.steps >= c.blow_at for c in self.cs):
            raise shockx.ShxError("checkNAN: stand-in", shockx.ERR_NAN)

    def nan_counts(self):
        return np.array([int(getattr(c, "blow_at", None) is not None and c.steps >= c.blow_at) for c in self.cs])

    def close(self):
        pass


def test_every_request_runs_its_steps_within_budget():
    reqs = serve.make_requests(12, rate=200.0, seed=3, n_lo=16, n_hi=48)
    log = []

    def open_fn(r):
        r.steps = 5 + r.rid * 3
        r.tau = r.t_end / r.steps
        return FakeCase(r, log)

    fp = lambda n, S: n ** 3                                    # noqa: E731
    budget = 3 * 48 ** 3
    srv = serve.Server(budget_bytes=budget, max_slots=4, chunk=4, keep_results=True, open_fn=open_fn,
                       batch_cls=FakeBatch, footprint_fn=fp)
    stats, done = srv.run(reqs)
    assert stats["requests"] == 12
    for r in done:
        assert r.result == [r.steps]                           # exactly its own step count
        assert r.done >= r.admitted >= r.arrival
    assert FakeBatch.resident_peak <= 4
    assert sorted(rid for _, rid in log) == list(range(12))
    assert stats["p99_s"] >= stats["p50_s"] >= 0.0


def test_request_larger_than_the_budget_is_refused_not_queued():
    # a request that cannot fit an empty device is refused on arrival; the others still run
    reqs = serve.make_requests(5, rate=500.0, seed=4, n_lo=16, n_hi=24)
    reqs[2].n = 400
    log = []

    def open_fn(r):
        r.steps = 3
        r.tau = r.t_end / r.steps
        return FakeCase(r, log)

    srv = serve.Server(budget_bytes=2 * 24 ** 3, max_slots=4, chunk=2, keep_results=True, open_fn=open_fn,
                       batch_cls=FakeBatch, footprint_fn=lambda n, S: n ** 3)
    stats, done = srv.run(reqs)
    assert stats["requests"] == 4 and stats["rejected"] == 1
    assert sorted(r.rid for r in done) == [0, 1, 3, 4]


def test_request_that_goes_non_finite_fails_alone():
    reqs = serve.make_requests(6, rate=300.0, seed=5, n_lo=16, n_hi=24)
    log = []

    def open_fn(r):
        r.steps = 12
        r.tau = r.t_end / r.steps
        c = FakeCase(r, log)
        if r.rid == 2:
            c.blow_at = 5                   # its state goes non-finite during the second chunk
        return c

    srv = serve.Server(budget_bytes=10 * 24 ** 3, max_slots=8, chunk=4, keep_results=True, open_fn=open_fn,
                       batch_cls=FakeBatch, footprint_fn=lambda n, S: n ** 3)
    stats, done = srv.run(reqs)
    assert stats["failed"] == 1 and stats["requests"] == 5
    assert sorted(r.rid for r in done) == [0, 1, 3, 4, 5]
    for r in done:
        assert r.result == [12]
    assert sorted(rid for _, rid in log) == list(range(6))  # the failed case was closed too
