"""Serving: requests admitted by HBM budget into a continuously batched loop
give exactly the results they give alone (one-at-a-time runs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_results(vh, gpu_ready):
    S = vh.serve
    reqs = S.make_requests(10, rate=200.0, seed=3, n_min=1000, n_max=40000, h_min=5, h_max=40)
    kinds = {r.kind for r in reqs}
    assert kinds == {"sir", "opinion"}
    # a budget that holds only part of the stream at once forces admission waits
    budget = sum(S.footprint(r) for r in reqs) // 3
    srv = S.Server(vh, hbm_budget=budget, quantum=7)
    done, wall = srv.run(reqs)
    srv.close()
    assert len(done) == len(reqs)
    seq, _ = S.run_sequential(vh, reqs)
    a = {r.id: r.digest for r in done}
    b = {r.id: r.digest for r in seq}
    assert a == b
    s = S.summarize(done, wall)
    print(f"[gpu] served {s}", flush=True)
    assert s["p99_s"] >= s["p50_s"] > 0
