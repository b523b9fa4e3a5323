"""The device graph algorithms restated in plain Python on small inputs,
against the host generators the checker's inputs use: the Barabasi-Albert
rounds of vh_graphgen.hip (each node resumes at its first draw whose target
is another node's choice not made yet) give vh_graph_ba's pairs, and a
level-synchronous numbering with smallest-enumeration-index winners gives
vh_bfs_order's numbering."""
import numpy as np
import pytest

KEY_BA = 10 << 40


def _clique(k, m):
    r, ln = 1, m - 1
    while k >= ln:
        k -= ln
        r += 1
        ln -= 1
    return r, r + 1 + k


def ba_rounds(vh, n, m, seed, budget=64):
    """The rounds of k_ba_round; a round reads only the choices made before
    it (the slowest order the device can observe). Returns (a, b, rounds)."""
    h4 = vh._abi.lib().hash4
    C = m * (m - 1) // 2
    K = C + (n - m) * m
    a = np.zeros(K, np.int64)
    b = np.zeros(K, np.int64)
    for k in range(C):
        i, j = _clique(k, m)
        a[k], b[k] = j, i
    for x in range(n - m):
        a[C + x * m: C + (x + 1) * m] = m + 1 + x
    draws = [0] * (n - m)
    got = [0] * (n - m)
    rounds = 0
    while True:
        rounds += 1
        prev = b.copy()
        left = 0
        for x in range(n - m):
            g = got[x]
            if g >= m:
                continue
            v, base = m + 1 + x, C + x * m
            L, d = 2 * base, draws[x]
            for _ in range(budget):
                if g >= m:
                    break
                if L == 0:
                    u = g + 1
                else:
                    idx = ((h4(seed, v, d, KEY_BA) >> 32) * L) >> 32
                    k = idx >> 1
                    if k < C:
                        i, j = _clique(k, m)
                        u = j if idx & 1 else i
                    elif not idx & 1:
                        u = m + 1 + (k - C) // m
                    else:
                        u = int(prev[k])
                        if u == 0:
                            break
                    d += 1
                if u not in b[base:base + g]:
                    b[base + g] = u
                    g += 1
            draws[x], got[x] = d, g
            left += g < m
        if not left:
            return a, b, rounds


@pytest.mark.parametrize("n,m,seed", [(600, 5, 1), (50, 1, 3), (7, 3, 2), (5, 5, 4), (300, 12, 5)])
def test_ba_rounds_equal_sequential(vh, n, m, seed):
    a, b, rounds = ba_rounds(vh, n, m, seed)
    ha, hb = vh.inputs.graph_ba(n, m, seed)
    assert np.array_equal(a, ha) and np.array_equal(b, hb)
    assert rounds <= 64


def bfs_levels(n, ea, eb):
    """bfs_device restated: incidence lists in edge order, the
    highest-degree root, levels by smallest enumeration index, then the
    lowest unvisited agents (those without edges in order)."""
    adj = [[] for _ in range(n + 1)]
    for x, y in zip(ea, eb):
        adj[x].append(y)
        adj[y].append(x)
    deg = [len(v) for v in adj]
    root = max(range(1, n + 1), key=lambda v: (deg[v], -v))
    pos = [None] * (n + 1)
    order = []
    cursor, first = 1, True
    while len(order) < n:
        if not first:
            r2 = next((v for v in range(cursor, n + 1) if pos[v] is None and deg[v] > 0), None)
            lim = r2 if r2 is not None else n + 1
            for v in range(cursor, lim):
                if pos[v] is None:
                    pos[v] = len(order)
                    order.append(v)
            if r2 is None:
                break
            root, cursor = r2, r2 + 1
        first = False
        pos[root] = len(order)
        order.append(root)
        lo, hi = len(order) - 1, len(order)
        while lo < hi:
            key = {}
            e = 0
            for p in range(lo, hi):
                for w in adj[order[p]]:
                    if pos[w] is None and w not in key:
                        key[w] = e
                    e += 1
            lo = hi
            for w in sorted(key, key=key.get):
                pos[w] = len(order)
                order.append(w)
            hi = len(order)
    return np.array([pos[v] + 1 for v in range(1, n + 1)], dtype=np.uint64)


@pytest.mark.parametrize("case", ["er", "ba", "selfloops"])
def test_bfs_levels_equal_sequential(vh, case):
    rng = np.random.default_rng(3)
    if case == "er":
        n = 400
        ea, eb = rng.integers(1, n + 1, 250), rng.integers(1, n + 1, 250)
    elif case == "ba":
        a, b = vh.inputs.graph_ba(300, 3, 2)
        fr, to, _w = vh.inputs.symmetric_weighted(a, b, 3)
        n, ea, eb = 300, fr.astype(np.int64), to.astype(np.int64)
    else:
        n = 120
        ea, eb = rng.integers(1, n + 1, 200), rng.integers(1, n + 1, 200)
        ea[::5] = eb[::5]
    want = vh.inputs.bfs_order(n, ea.astype(np.uint32), eb.astype(np.uint32))
    assert np.array_equal(bfs_levels(n, ea.tolist(), eb.tolist()), want)
