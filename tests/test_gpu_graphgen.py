"""Graphs built on the device at finish_init! (vh_add_edges_ba) and the
breadth-first agent numbering on the device (vh_finish_init_relabel): the
same CSR and the same numbering as the host generator and vh_bfs_order,
which the checker's inputs use."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr(m):
    to, fr, st = m.sim._read_edges(m.E)
    return to, fr, np.asarray(st["w"]) if st is not None else None


@pytest.mark.parametrize("n,mba,seed", [(2000, 5, 1), (50, 1, 3), (7, 3, 2), (5, 5, 4), (30000, 3, 9),
                                        (4000, 12, 5)])
def test_ba_device_equals_host(vh, ora, gpu_ready, n, mba, seed):
    """The device BA graph equals vh_graph_ba + both directions of every
    pair added on the host, edge by edge (CSR order, sources, weights), on
    the device and in the checker."""
    print(f"[gpu] device BA n={n} m={mba} seed={seed}", flush=True)
    M = vh.models
    d = M.build_opinion(vh, n, m_ba=mba, seed=seed, device_graph=True)
    h = M.build_opinion(vh, n, m_ba=mba, seed=seed)
    c = M.build_opinion(ora, n, m_ba=mba, seed=seed)
    a, b, w = _csr(d)
    for other in (h, c):
        x, y, z = _csr(other)
        assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(w, z)
    k = mba * (mba - 1) // 2 + (n - mba) * mba
    assert len(a) == 2 * k
    for m in (d, h):
        m.sim._destroy()


def _bfs_case(vh, name):
    """(n, [(hints, from, to)] per edge type) of a relabel test graph."""
    rng = np.random.default_rng(7)
    if name == "er_components":
        n = 3000
        k = 1400                      # sparse: many components, isolated agents
        return n, [((), rng.integers(1, n + 1, k), rng.integers(1, n + 1, k))]
    if name == "self_loops_two_types":
        n = 500
        f1, t1 = rng.integers(1, n + 1, 900), rng.integers(1, n + 1, 900)
        f1[::7] = t1[::7]             # self-loops
        f2, t2 = rng.integers(1, 40, 300), rng.integers(200, 260, 300)
        return n, [((), f1, t1), ((), f2, t2)]
    if name == "no_edges":
        return 64, [((), np.zeros(0, np.int64), np.zeros(0, np.int64))]
    if name == "star_and_path":
        n = 400
        f = np.concatenate([np.full(100, 150), np.arange(300, 399)])
        t = np.concatenate([np.arange(1, 101), np.arange(301, 400)])
        return n, [((), f, t)]
    raise KeyError(name)


@pytest.mark.parametrize("case", ["er_components", "self_loops_two_types", "no_edges", "star_and_path"])
def test_relabel_device_equals_host(vh, gpu_ready, case):
    """vh_finish_init_relabel numbers the agents on the device exactly as
    vh_bfs_order (highest-degree root, then the lowest unvisited agents,
    isolated ones included) over the init edges of every edge type in order."""
    print(f"[gpu] device BFS numbering {case}", flush=True)
    n, etypes = _bfs_case(vh, case)
    A = vh.struct("A", [("v", "<i8")])
    types = vh.ModelTypes() | vh.register_agenttype_(A)
    Es = []
    for i, (hints, _f, _t) in enumerate(etypes):
        E = vh.struct(f"E{i}")
        types = types | vh.register_edgetype_(E, *hints)
        Es.append(E)
    sim = vh.create_simulation(types | vh.create_model(f"bfs_{case}"))
    st = np.zeros(n, dtype=A.dtype)
    st["v"] = np.arange(1, n + 1)
    sim._add_agents_array(A, st)
    tid = np.uint64(1) << np.uint64(56)
    for E, (_h, f, t) in zip(Es, etypes):
        if len(f):
            sim._add_edges_array(E, tid | f.astype(np.uint64), tid | t.astype(np.uint64))
    got = vh.finish_init_(sim, relabel=A)
    fa = np.concatenate([f for _h, f, _t in etypes]).astype(np.uint32)
    ta = np.concatenate([t for _h, _f, t in etypes]).astype(np.uint32)
    want = vh.inputs.bfs_order(n, fa, ta)
    assert np.array_equal(np.asarray(got, dtype=np.uint64), want)
    # states travel with their agents
    s, _ = sim._read_agents(A)
    v = np.asarray(s["v"])
    assert np.array_equal(v[(want - 1).astype(np.int64)], np.arange(1, n + 1))


def test_relabel_device_ba_fullsize(vh, gpu_ready):
    """The headline graph (10^7 agents, BA m = 5, symmetrised): device BA and
    device numbering give the numbering of vh_bfs_order over the host graph
    and, edge by edge, the CSR of the host graph renumbered on the host."""
    import time
    M = vh.models
    n = 10_000_000
    print("[gpu] device BA + numbering at 1e7", flush=True)
    t0 = time.perf_counter()
    d = M.build_opinion(vh, n, seed=1, device_graph=True, relabel=True)
    t_dev = time.perf_counter() - t0
    a, b = vh.inputs.graph_ba(n, 5, 1)
    fr, to, w = vh.inputs.symmetric_weighted(a, b, 2)
    del a, b
    want = vh.inputs.bfs_order(n, fr, to)
    assert np.array_equal(d.idmap, want)
    # CSR of the host graph renumbered on the host: rows by new target,
    # each row in insertion order
    nf = want[fr.astype(np.int64) - 1]
    nt = want[to.astype(np.int64) - 1]
    del fr, to
    order = np.argsort(nt, kind="stable")
    x, y, z = _csr(d)
    assert np.array_equal(x & np.uint64(vh._abi.VH_NR_MASK), nt[order])
    assert np.array_equal(y & np.uint64(vh._abi.VH_NR_MASK), nf[order])
    assert np.array_equal(z, w[order])
    print(f"[gpu] device BA + numbering at 1e7: {t_dev:.2f} s to a ready simulation", flush=True)
    d.sim._destroy()
