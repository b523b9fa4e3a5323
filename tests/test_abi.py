"""C ABI surface and host-side helpers (CPU only: nothing here launches a kernel)."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vahana_hip.h")

# SplitMix64 reference output for seed 1234567 (Vigna, splitmix64.c)
SPLITMIX_1234567 = [6457827717110365317, 3203168211198807973, 9817491932198370423,
                    4593380528125082431, 16408922859458223821]


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vh_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(vh):
    lib = vh._abi.lib()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib.dll, n)]
    assert not missing, missing


def test_struct_layouts_match_header(vh):
    A = vh._abi
    assert C.sizeof(A.ApplyDesc) == 80
    assert C.sizeof(A.ValueBind) == 8
    assert C.sizeof(A.Options) == 32
    assert C.sizeof(A.MapDesc) == 24
    assert C.sizeof(A.TestParams) == 32
    assert C.sizeof(A.SirParams) == 24
    assert C.sizeof(A.ContactParams) == 40
    assert C.sizeof(A.RasterParams) == 72


def _mix64(z):
    M = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_splitmix_known_answers():
    G = 0x9E3779B97F4A7C15
    s, out = 1234567, []
    for _ in range(5):
        s = (s + G) & ((1 << 64) - 1)
        out.append(_mix64(s))
    assert out == SPLITMIX_1234567


def test_hash_agrees_everywhere(vh, ora):
    # first stage of vh_hash4 is the first SplitMix64 output of the seed
    G = 0x9E3779B97F4A7C15
    M = (1 << 64) - 1

    def h4(s, a, b, c):
        h = _mix64((s + G) & M)
        for w in (a, b, c):
            h = _mix64((h + w + G) & M)
        return h

    assert _mix64((1234567 + G) & M) == SPLITMIX_1234567[0]
    rng = np.random.default_rng(3)
    L, O = vh._abi.lib(), ora.lib()
    for _ in range(200):
        s, a, b, c = (int(x) for x in rng.integers(0, 2**63, 4, dtype=np.int64))
        ref = h4(s, a, b, c)
        assert L.hash4(s, a, b, c) == ref
        assert O.hash4(s, a, b, c) == ref
        assert int(vh.inputs.hash4(s, a, b, c)) == ref
    # the vectorised form
    a = np.arange(1000, dtype=np.uint64)
    got = vh.inputs.hash4(7, a, 3, 5)
    assert all(int(got[i]) == h4(7, i, 3, 5) for i in range(0, 1000, 97))


def test_uniform_conversions(vh):
    assert vh.inputs.u01(np.uint64(0)) == 0.0
    assert vh.inputs.u01(np.uint64(2**64 - 1)) == 1.0 - 2.0**-53
    w = vh.inputs.open01f(np.array([0, 2**64 - 1], dtype=np.uint64))
    assert w[0] == np.float32(2.0**-24) and w[1] == np.float32(1.0 - 2.0**-24)


def test_er_generator_properties(vh):
    n, p = 20000, 10.0 / 19999
    fr, to = vh.inputs.graph_er(n, p, 5)
    assert np.all(fr != to)
    assert np.all(np.diff(to.astype(np.int64)) >= 0)          # targets ascending
    # sources ascending within a target, no duplicates
    key = to.astype(np.int64) * (n + 1) + fr
    assert np.all(np.diff(key) > 0)
    mean = len(to) / n
    assert abs(mean - 10.0) < 0.2                              # E[d] = (n-1) p = 10
    deg = np.bincount(to, minlength=n + 1)[1:]
    assert abs(deg.var() - 10.0 * (1 - p)) < 0.6               # binomial variance
    # determinism
    fr2, to2 = vh.inputs.graph_er(n, p, 5)
    assert np.array_equal(fr, fr2) and np.array_equal(to, to2)


def test_ba_generator_properties(vh):
    n, m = 5000, 5
    a, b = vh.inputs.graph_ba(n, m, 11)
    assert len(a) == m * (m - 1) // 2 + (n - m) * m
    k0 = m * (m - 1) // 2
    assert np.all(a[k0:] > b[k0:])                             # new node -> earlier node
    assert np.all(np.bincount(a[k0:], minlength=n + 1)[m + 1:] == m)
    pairs = set(zip(a.tolist(), b.tolist()))
    assert len(pairs) == len(a)                                # no multi-edges
    deg = np.bincount(np.concatenate([a, b]), minlength=n + 1)[1:]
    assert deg.min() >= m - 1 and deg.sum() == 2 * len(a)
    assert deg.max() > 10 * m                                  # heavy tail (hubs)


def test_weight_fill_matches_hash(vh):
    a, b = vh.inputs.graph_ba(300, 3, 1)
    fr, to, w = vh.inputs.symmetric_weighted(a, b, 9)
    idx = np.arange(len(w), dtype=np.uint64)
    ref = vh.inputs.open01f(vh.inputs.hash4(9, idx, 0, vh.inputs.KEY_WEIGHT))
    assert np.array_equal(w, ref)
    assert np.all((w > 0) & (w < 1))
    assert np.array_equal(fr[0::2], to[1::2]) and np.array_equal(to[0::2], fr[1::2])


def test_gnp_generator_properties(vh):
    n, md = 30000, 12.0
    fr, to = vh.inputs.graph_gnp(n, md, 4)
    assert np.all(fr != to) and np.all((fr >= 1) & (fr <= n))
    assert np.all(np.diff(to.astype(np.int64)) >= 0)
    deg = np.bincount(to, minlength=n + 1)[1:]
    p = md / (n - 1)
    assert abs(deg.mean() - md) < 0.1                      # E = (n-1) p
    assert abs(deg.var() - md * (1 - p)) < 0.5             # Binomial variance
    key = to.astype(np.int64) * (n + 1) + fr
    assert len(np.unique(key)) == len(key)                  # distinct sources per target
    # sources are uniform: mean source id ~ (n + 1) / 2
    assert abs(fr.mean() / ((n + 1) / 2) - 1) < 0.01
    fr2, to2 = vh.inputs.graph_gnp(n, md, 4)
    assert np.array_equal(fr, fr2) and np.array_equal(to, to2)


def test_bfs_order_known_answers(vh):
    bfs = vh.inputs.bfs_order
    # path 1-2-3-4-5: root 2 (highest degree, lowest number), then 1, 3, 4, 5
    assert bfs(5, [1, 2, 3, 4], [2, 3, 4, 5]).tolist() == [2, 1, 3, 4, 5]
    # star around 3 (edges 3-1, 3-2, 3-4): 3 first, its neighbours in edge order
    assert bfs(4, [3, 3, 3], [1, 2, 4]).tolist() == [2, 3, 1, 4]
    # two components and an isolated vertex: component by component
    assert bfs(5, [1, 4], [2, 5]).tolist() == [1, 2, 3, 4, 5]
    # random graph: a permutation whose first vertex has the highest degree,
    # and every later vertex of a component has a neighbour numbered before it
    rng = np.random.default_rng(3)
    n = 300
    a = rng.integers(1, n + 1, 900).astype(np.uint32)
    b = rng.integers(1, n + 1, 900).astype(np.uint32)
    nn = bfs(n, a, b)
    assert sorted(nn.tolist()) == list(range(1, n + 1))
    deg = np.bincount(a, minlength=n + 1) + np.bincount(b, minlength=n + 1)
    root = int(np.nonzero(nn == 1)[0][0]) + 1
    assert deg[root] == deg[1:].max()
    nbr = {v: set() for v in range(1, n + 1)}
    for x, y in zip(a.tolist(), b.tolist()):
        nbr[x].add(y)
        nbr[y].add(x)
    for v in range(1, n + 1):
        if nbr[v] and nn[v - 1] != 1:
            earlier = [u for u in nbr[v] if nn[u - 1] < nn[v - 1]]
            later_root = all(nn[u - 1] > nn[v - 1] for u in nbr[v])
            assert earlier or later_root   # v is the root of a later component


def test_relabelled_checker_run_maps_to_plain_run(vh, ora):
    """The opinion model on a breadth-first renumbered graph (the checker's
    host renumbering) equals the plain run agent by agent: states travel with
    their agents and every target keeps its in-edge order."""
    M = vh.models
    a, b = vh.inputs.graph_ba(2000, 3, 5)
    edges = vh.inputs.symmetric_weighted(a, b, 6)
    p = M.build_opinion(ora, 2000, seed=5, edges=edges)
    r = M.build_opinion(ora, 2000, seed=5, edges=edges, relabel=True)
    for _ in range(3):
        M.step_opinion(ora, p, eps=0.3)
        M.step_opinion(ora, r, eps=0.3)
    xp = np.asarray(p.sim._read_agents(p.T)[0]["opinion"])
    xr = np.asarray(r.sim._read_agents(r.T)[0]["opinion"])
    assert np.array_equal(xr[(r.idmap - 1).astype(np.int64)], xp)
