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


def test_served_config4_mix_equal_solo_results(vh, gpu_ready):
    """Config c4 at its stated sizes: 24 Poisson requests, SIR or opinion,
    N ~ logU[10^4, 10^7], horizons U[50, 500], HBM budget 0.9 x free; every
    served result equals its one-at-a-time run (xxh3-128 digests)."""
    import torch
    S = vh.serve
    reqs = S.make_requests(24, rate=20.0, seed=11)
    assert max(r.n for r in reqs) > 10 ** 6 and min(r.n for r in reqs) < 10 ** 5
    print(f"[gpu] serving c4 mix: {len(reqs)} requests, {sum(r.n * r.horizon for r in reqs):.3g} agent-steps",
          flush=True)
    free, _total = torch.cuda.mem_get_info()
    srv = S.Server(vh, hbm_budget=int(0.9 * free), keep_results=False)
    done, wall = srv.run(reqs)
    srv.close()
    seq, _ = S.run_sequential(vh, reqs)
    assert len(done) == len(reqs)
    assert {r.id: r.digest for r in done} == {r.id: r.digest for r in seq}
    print(f"[gpu] serving c4 mix: {S.summarize(done, wall)}", flush=True)


def test_serving_failures_leave_others_unchanged(vh, gpu_ready):
    """A request larger than the whole budget is rejected, one is cancelled
    after its first quantum, one fails to build (an injected device
    allocation failure); every other request's result equals its
    one-at-a-time run."""
    import types
    S = vh.serve
    lib = vh._abi.lib()
    reqs = S.make_requests(10, rate=0, seed=5, n_min=1000, n_max=30000, h_min=12, h_max=40)
    big = types.SimpleNamespace(id=100, kind="opinion", n=10 ** 9, horizon=10, seed=1, arrival=0.0,
                                mean_deg=10.0, p_inf=0.05)
    cancel_id, fail_id = reqs[1].id, reqs[2].id

    def hook(r):
        if r.id == fail_id:
            lib.debug_fail_allocs(1)

    srv = S.Server(vh, hbm_budget=10 ** 9, quantum=5, solo_agents=20_000)
    try:
        done, wall = srv.run(reqs + [big], cancel={cancel_id: 1}, before_admit=hook)
    finally:
        lib.debug_fail_allocs(0)
        srv.close()
    status = {r.id: r.status for r in done}
    print(f"[gpu] statuses {status}", flush=True)
    assert status[100] == "rejected"
    assert status[cancel_id] == "cancelled"
    assert status[fail_id] == "failed"
    assert len(done) == len(reqs) + 1
    ok = [r for r in reqs if status[r.id] == "done"]
    assert len(ok) == len(reqs) - 2
    seq, _ = S.run_sequential(vh, ok)
    assert {r.id: r.digest for r in done if r.status == "done"} == {r.id: r.digest for r in seq}
    s = S.summarize(done, wall)
    assert s["rejected"] == 1 and s["failed"] == 1 and s["cancelled"] == 1


def test_injected_device_error_stays_with_its_simulation(vh, ora, gpu_ready):
    """A device error flag raised in one simulation (the fault hook) makes its
    next apply! fail with an assertion; another simulation steps on,
    equal to the checker."""
    M = vh.models
    a = M.build_opinion_gnp(vh, 2000, mean_deg=6.0, seed=1)
    b = M.build_opinion_gnp(vh, 2000, mean_deg=6.0, seed=2)
    c = M.build_opinion_gnp(ora, 2000, mean_deg=6.0, seed=2, device_graph=False)
    vh.enable_asserts(True)
    vh._abi.lib().debug_set_device_error(a.sim.handle, 1)
    with pytest.raises(AssertionError):
        M.step_opinion(vh, a, eps=0.2)
    for _ in range(3):
        M.step_opinion(vh, b, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
    x, _ = b.sim._read_agents(b.T)
    y, _ = c.sim._read_agents(c.T)
    assert np.array_equal(np.ascontiguousarray(x).view(np.uint8), np.ascontiguousarray(y).view(np.uint8))
