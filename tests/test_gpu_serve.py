"""Serving: requests admitted by HBM budget into a continuously batched loop
give exactly the results they give alone (one-at-a-time runs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_results(vh, ora, gpu_ready):
    S = vh.serve
    reqs = S.make_requests(10, rate=200.0, seed=3, n_min=1000, n_max=40000, h_min=5, h_max=40)
    kinds = {r.kind for r in reqs}
    assert kinds == {"sir", "opinion"}
    # a budget that holds only part of the stream at once forces admission waits
    budget = sum(S.footprint(r) for r in reqs) // 3
    # SIR requests of 20000+ agents run in ensembles of their own, the rest share one
    srv = S.Server(vh, hbm_budget=budget, quantum=7, keep_results=True, solo_agents=20_000)
    assert any(r.kind == "sir" and r.n >= 20_000 for r in reqs)
    assert any(r.kind == "sir" and r.n < 20_000 for r in reqs)
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
    # one served opinion request against the checker (host twin of its graph)
    r = min((r for r in done if r.kind == "opinion"), key=lambda r: r.n)
    M = vh.models
    c = M.build_opinion_gnp(ora, r.n, mean_deg=r.mean_deg, seed=r.seed, graph_seed=r.seed + 1,
                            weight_seed=r.seed + 7, device_graph=False)
    for _ in range(r.horizon):
        M.step_opinion(ora, c, eps=0.2)
    want = np.asarray(c.sim._read_agents(c.T)[0])
    assert np.array_equal(np.ascontiguousarray(r.result).view(np.uint8),
                          np.ascontiguousarray(want).view(np.uint8)), (r.n, r.horizon)
