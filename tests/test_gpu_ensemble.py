"""Ensembles (parameter sweeps): every member of one segmented, graph-captured
launch is bit-identical to the same SIR simulation run alone through apply!
(which itself matches the checker)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = [  # n, mean in-degree, p_inf, seed, graph seed
    (1000, 4.0, 0.05, 11, 101), (3000, 12.0, 0.1, 12, 102), (20000, 20.0, 0.01, 13, 103),
    (150000, 8.0, 0.07, 14, 104), (777, 15.0, 0.09, 15, 105), (64, 6.0, 0.1, 16, 106),
    (65536, 10.0, 0.03, 17, 107), (1, 4.0, 0.05, 18, 108),
]


def _solo(V, n, md, p, seed, gseed, steps, n_init):
    from __graft_entry__ import load_package
    vh = load_package()
    m = vh.models.build_sir(V, n, seed=seed, n_init=n_init,
                            edges=vh.inputs.graph_gnp(n, md, gseed), init="frac")
    for _ in range(steps):
        vh.models.step_sir(V, m, p_inf=p, seed=seed, recovery=7)
    st, _ = m.sim._read_agents(m.T)
    return np.ascontiguousarray(st).view(np.uint8)


@pytest.mark.parametrize("expose", ["1", "0"])
def test_ensemble_members_match_solo(vh, ora, gpu_ready, expose, opts):
    """With (default) and without the exposure pass (option sir_expose), in the
    ensemble and in the solo runs; 2% initially infected, so the pass is on
    at first and turns off where the epidemic grows."""
    opts(sir_expose=int(expose), sir_expose_min=0)   # the solo runs of small members too
    print(f"[gpu] ensemble vs solo expose={expose}", flush=True)
    E = vh.ensemble.Ensemble()
    ids = []
    for n, md, p, seed, gseed in SPECS:
        ids.append(E.add_sir(n, md, p, seed, gseed, recovery=7, n_init=max(1, n // 50)))
    block, remap, ms = E.autotune(2)
    print(f"[gpu] autotune: block {block} xcd_remap {remap} {ms:.3f} ms/step", flush=True)
    E.step(13, use_graph=True)
    E.step(4, use_graph=False)
    steps = 17
    for mid, (n, md, p, seed, gseed) in zip(ids, SPECS):
        got, next_step = E.read(mid)
        assert next_step == steps + 1
        want = _solo(vh, n, md, p, seed, gseed, steps, max(1, n // 50))
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8), want), f"member n={n}"
    # a small member against the checker as well
    n, md, p, seed, gseed = SPECS[1]
    got, _ = E.read(ids[1])
    assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                          _solo(ora, n, md, p, seed, gseed, steps, max(1, n // 50)))
    info = E.info()
    assert info["members"] == len(SPECS) and info["agents"] == sum(s[0] for s in SPECS)


def test_ensemble_incremental_exposure(vh, ora, gpu_ready, opts):
    """Slow epidemics (0.1% initially infected, few infections per step) keep
    the exposure pass on for every launch, so the bitmaps are updated from the
    newly infected agents between full launches (every 8th); 30 steps in
    graph and plain launches of several lengths, each member against its solo
    run without the pass, one against the checker."""
    opts(sir_expose=1, sir_expose_min=0)
    print("[gpu] ensemble incremental exposure", flush=True)
    specs = [(20000, 8.0, 0.02, 31, 301), (50000, 6.0, 0.03, 32, 302), (5000, 10.0, 0.015, 33, 303)]
    E = vh.ensemble.Ensemble()
    ids = [E.add_sir(n, md, p, seed, gseed, recovery=7, n_init=max(1, n // 1000))
           for n, md, p, seed, gseed in specs]
    E.step(11, use_graph=True)
    E.step(3, use_graph=False)
    E.step(16, use_graph=True)
    steps = 30
    opts(sir_expose=0)
    for mid, (n, md, p, seed, gseed) in zip(ids, specs):
        got, next_step = E.read(mid)
        assert next_step == steps + 1
        g = np.ascontiguousarray(got)
        inf = int((g.view(np.uint8).reshape(n, -1)[:, 0] == vh.I).sum())
        rec = int((g.view(np.uint8).reshape(n, -1)[:, 0] == vh.R).sum())
        assert rec > 0 and inf * 32 <= n, f"member n={n}: {inf} infected, {rec} recovered"
        want = _solo(vh, n, md, p, seed, gseed, steps, max(1, n // 1000))
        assert np.array_equal(g.view(np.uint8), want), f"member n={n}"
    n, md, p, seed, gseed = specs[2]
    got, _ = E.read(ids[2])
    assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                          _solo(ora, n, md, p, seed, gseed, steps, max(1, n // 1000)))


@pytest.mark.parametrize("expose_min", ["65536", "0"])
def test_ensemble_join_and_leave(vh, gpu_ready, expose_min, opts):
    """Members added and removed between steps keep their own step counters;
    with sir_expose_min=0 the exposure pass runs across the rebuilt block
    tables (every member's first launch after a change has the pass off)."""
    opts(sir_expose_min=int(expose_min))
    print(f"[gpu] ensemble join and leave expose_min={expose_min}", flush=True)
    E = vh.ensemble.Ensemble()
    a = E.add_sir(5000, 10.0, 0.08, 21, 201, n_init=50)
    E.step(5)
    b = E.add_sir(8000, 6.0, 0.05, 22, 202, n_init=80)
    E.step(6)
    E.remove(a)
    c = E.add_sir(3000, 14.0, 0.1, 23, 203, n_init=30)
    E.step(4, use_graph=False)
    for mid, (n, md, p, seed, gseed, st) in ((b, (8000, 6.0, 0.05, 22, 202, 10)),
                                             (c, (3000, 14.0, 0.1, 23, 203, 4))):
        got, nxt = E.read(mid)
        assert nxt == st + 1
        want = _solo(vh, n, md, p, seed, gseed, st, n // 100)
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8), want)


@pytest.mark.parametrize("apt", [(4, 2, 1), (1, 4, 2), (2, 1, 4)])
def test_ensemble_tiles_per_size_class(vh, gpu_ready, apt, opts):
    """Logical blocks of 1, 2 or 4 agents per thread per size class (<1e4,
    <1e5, >=1e5 agents), with the exposure pass on: every member equals its
    solo run."""
    opts(sir_expose=1, sir_expose_min=0)
    print(f"[gpu] ensemble tiles per size class {apt}", flush=True)
    specs = [(700, 9.0, 0.08, 41, 401), (9999, 6.0, 0.05, 42, 402), (10000, 12.0, 0.03, 43, 403),
             (54321, 7.0, 0.06, 44, 404), (100000, 5.0, 0.04, 45, 405), (180000, 10.0, 0.02, 46, 406)]
    E = vh.ensemble.Ensemble()
    ids = [E.add_sir(n, md, p, seed, gseed, recovery=7, n_init=max(1, n // 200))
           for n, md, p, seed, gseed in specs]
    E.set_tiles(apt)
    E.step(9, use_graph=True)
    E.step(3, use_graph=False)
    for mid, (n, md, p, seed, gseed) in zip(ids, specs):
        got, nxt = E.read(mid)
        assert nxt == 13
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                              _solo(vh, n, md, p, seed, gseed, 12, max(1, n // 200))), f"member n={n}"


def test_ensemble_1024_replicas_config3(vh, gpu_ready):
    """Config c3 at its stated size: 1024 SIR replicas, N ~ logU[1e3, 1e6]
    (1.5e8 agents, 1.9e9 edges), mean in-degree U[4, 20], p U[0.01, 0.1],
    distinct seeds; autotuned (block, order, tiles per size class); 6 steps
    in one graph launch; 16 members spread over the sizes (the smallest and
    largest included) equal their solo runs."""
    rng = np.random.default_rng(7)
    R = 1024
    specs = [(int(10 ** rng.uniform(3, 6)), float(rng.uniform(4, 20)), float(rng.uniform(0.01, 0.1)), 1 + i)
             for i in range(R)]
    print("[gpu] ensemble 1024 replicas: building", flush=True)
    E = vh.ensemble.Ensemble()
    ids = [E.add_sir(n, md, p, seed, 10_000 + seed, 7, max(1, n // 1000)) for n, md, p, seed in specs]
    info = E.info()
    assert info["members"] == R and info["agents"] == sum(s[0] for s in specs)
    block, remap, ms = E.autotune(1)
    print(f"[gpu] ensemble 1024: {info}, autotune block {block} remap {remap} tiles {E.class_apt} "
          f"{ms:.3f} ms/step", flush=True)
    E.step(6, use_graph=True)
    E.sync()
    by_size = np.argsort([s[0] for s in specs])
    pick = sorted(set(by_size[np.linspace(0, R - 1, 16).astype(int)].tolist()))
    for k in pick:
        n, md, p, seed = specs[k]
        got, nxt = E.read(ids[k])
        assert nxt == 7
        assert np.array_equal(np.ascontiguousarray(got).view(np.uint8),
                              _solo(vh, n, md, p, seed, 10_000 + seed, 6, max(1, n // 1000))), f"member {k} n={n}"
    E.close()
