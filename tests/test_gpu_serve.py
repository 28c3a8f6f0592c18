"""Continuous-batching server on the GPU: requests that overlap on the device give each the
state it reaches when served alone (same kernels on the same data)."""
import numpy as np
import pytest

from shock_amd import serve, shockx

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [shockx.FP64, shockx.MIXED])
def test_served_results_equal_alone(precision):
    reqs = serve.make_requests(6, rate=50.0, seed=11, n_lo=12, n_hi=24, t_lo=0.02, t_hi=0.06, precision=precision,
                               shock_frac=0.5)
    srv = serve.Server(max_slots=3, chunk=2, keep_results=True)
    stats, done = srv.run(reqs)
    print(f"[serve] {stats}", flush=True)
    assert stats["requests"] == len(reqs)
    for r in done:
        alone = serve.run_alone(r)
        g = (r.S + 1) // 2
        sl = (slice(g, g + r.n),) * 3
        diff = max(float(np.max(np.abs(a[sl] - b[sl]))) for a, b in zip(r.result, alone))
        print(f"[serve] request {r.rid} {r.kind} n={r.n} S={r.S} steps={r.steps}: max |mix - alone| = {diff:.1e}",
              flush=True)
        assert diff == 0.0
