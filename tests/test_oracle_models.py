"""The checker's model bodies pinned by what the mathematics fixes (CPU only):
closed forms, a textbook BFS, invariants and exact replays of the canonical
summation order."""
import ctypes as C
import math

import numpy as np
import pytest


def _states(m):
    st, died = m.sim._read_agents(m.T)
    return np.asarray(st), died


# ------------------------------------------------------------------ SIR
def test_sir_p0_recovers_after_exactly_recovery_steps(vh, ora):
    M = vh.models
    n = 500
    m = M.build_sir(ora, n, mean_deg=6.0, seed=4, n_init=40)
    st0, _ = _states(m)
    inf0 = st0["health"] == vh.I
    for step in range(1, 10):
        M.step_sir(ora, m, p_inf=0.0, seed=4, recovery=7)
        st, _ = _states(m)
        assert not np.any((st["health"] != vh.S) & ~inf0)          # nobody new
        if step < 7:
            assert np.all(st["health"][inf0] == vh.I)
            assert np.all(st["days"][inf0] == step)
        else:
            assert np.all(st["health"][inf0] == vh.R)


def test_sir_p1_is_breadth_first_search(vh, ora):
    M = vh.models
    n = 800
    fr, to = vh.inputs.graph_er(n, 3.0 / (n - 1), 17)
    m = M.build_sir(ora, n, seed=2, n_init=3, edges=(fr, to))
    st, _ = _states(m)
    reached = st["health"] == vh.I
    adj = [[] for _ in range(n + 1)]          # source -> targets
    for a, b in zip(fr.tolist(), to.tolist()):
        adj[a].append(b)
    frontier = set((np.nonzero(reached)[0] + 1).tolist())
    seen = set(frontier)
    for _ in range(6):
        # an agent infected at step t can infect only while it is I
        M.step_sir(ora, m, p_inf=1.0, seed=2, recovery=100)
        nxt = {b for a in frontier for b in adj[a] if b not in seen}
        seen |= nxt
        frontier = nxt
        st, _ = _states(m)
        got = set((np.nonzero(st["health"] != vh.S)[0] + 1).tolist())
        assert got == seen


def _py_sir_step(st, off, src, seed, step, p, rec, type_id=1):
    """Brute force, pure Python: S -> I if any infected in-neighbour k
    (1-based position) draws u < p; I ages and recovers at `rec` days."""
    G, M64 = 0x9E3779B97F4A7C15, (1 << 64) - 1

    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def h4(s, a, b, c):
        h = mix((s + G) & M64)
        for w in (a, b, c):
            h = mix((h + w + G) & M64)
        return h

    out = [tuple(x) for x in st]
    for i in range(len(st)):
        h, pad, days = st[i]
        if h == 0:
            for k in range(off[i], off[i + 1]):
                j = src[k] - 1
                u = (h4(seed, (type_id << 56) | (i + 1), step, k - off[i] + 1) >> 11) * 2.0 ** -53
                if st[j][0] == 1 and u < p:
                    out[i] = (1, pad, 0)
                    break
        elif h == 1:
            d = (days + 1) & 0xFFFF
            out[i] = (2 if d >= rec else 1, pad, d)
    return out


def test_sir_brute_force_small(vh, ora):
    M = vh.models
    n = 40
    fr, to = vh.inputs.graph_er(n, 0.2, 3)
    m = M.build_sir(ora, n, seed=9, n_init=4, edges=(fr, to))
    order = np.argsort(to, kind="stable")
    src = fr[order].tolist()
    off = np.concatenate([[0], np.cumsum(np.bincount(to, minlength=n + 1)[1:])]).tolist()
    st = [tuple(int(v) for v in x) for x in _states(m)[0]]
    for step in range(1, 15):
        st = _py_sir_step(st, off, src, 9, step, 0.3, 3)
        M.step_sir(ora, m, p_inf=0.3, seed=9, recovery=3)
        got = [tuple(int(v) for v in x) for x in _states(m)[0]]
        assert got == st, step


def test_sir_invariants(vh, ora):
    M = vh.models
    n = 3000
    m = M.build_sir(ora, n, seed=1, n_init=30)
    prev_r = 0
    for _ in range(20):
        M.step_sir(ora, m, p_inf=0.1, seed=1)
        st, died = _states(m)
        assert not died.any()
        h = st["health"]
        assert ((h == 0) | (h == 1) | (h == 2)).all()
        r = int((h == 2).sum())
        assert r >= prev_r
        prev_r = r
    assert ora.num_agents(m.sim, m.T) == n


# ------------------------------------------------------------------ opinion
def test_canonical_sum_replay(ora):
    L = ora.lib()
    rng = np.random.default_rng(5)
    for d in (0, 1, 7, 64, 65, 200, 1000):
        a = rng.random(d) * 2 - 0.5
        b = rng.random(d)
        num, den = C.c_double(), C.c_double()
        L.canon_sum_f64(a.ctypes.data_as(C.POINTER(C.c_double)),
                        b.ctypes.data_as(C.POINTER(C.c_double)), d, C.byref(num), C.byref(den))
        if d <= 64:
            ra = np.cumsum(a)[-1] if d else 0.0      # sequential left fold
            rb = np.cumsum(b)[-1] if d else 0.0
        else:
            pa = np.array([np.cumsum(a[l::64])[-1] if len(a[l::64]) else 0.0 for l in range(64)])
            pb = np.array([np.cumsum(b[l::64])[-1] if len(b[l::64]) else 0.0 for l in range(64)])
            for o in (32, 16, 8, 4, 2, 1):
                pa = pa + pa[np.arange(64) ^ o]
                pb = pb + pb[np.arange(64) ^ o]
            ra, rb = pa[0], pb[0]
        assert num.value == ra and den.value == rb, d
        if d:
            bound = d * 2.0 ** -53 * np.abs(a).sum()
            assert abs(num.value - math.fsum(a)) <= bound


def _build_opinion_with(ora, xs, edges, w, f32=False):
    from __graft_entry__ import load_package
    vh = load_package()
    HK = ora.struct("HK32" if f32 else "HK", [("opinion", "<f4" if f32 else "<f8")])
    Knows = ora.struct("Knows", [("w", "<f4")])
    model = (ora.ModelTypes() | ora.register_agenttype_(HK, "Immortal")
             | ora.register_edgetype_(Knows, target=HK) | ora.create_model("opinion"))
    sim = ora.create_simulation(model)
    st = np.zeros(len(xs), dtype=HK.dtype)
    st["opinion"] = xs
    sim._add_agents_array(HK, st)
    tid = np.uint64(1) << np.uint64(56)
    fr = np.array([a for a, _ in edges], dtype=np.uint64) | tid
    to = np.array([b for _, b in edges], dtype=np.uint64) | tid
    ws = np.zeros(len(edges), dtype=Knows.dtype)
    ws["w"] = w
    sim._add_edges_array(Knows, fr, to, ws)
    ora.finish_init_(sim)
    import types
    return types.SimpleNamespace(sim=sim, T=HK, E=Knows, f32=f32)


def test_opinion_consensus_closed_form(vh, ora):
    # complete graph, uniform weights, eps >= 1: x_i' = (sum_j x_j - x_i) / (n - 1);
    # opinions are multiples of 2^-10 so every partial sum is exact
    n = 9
    xs = np.array([k / 1024.0 for k in (3, 100, 512, 77, 1000, 256, 9, 640, 1)])
    edges = [(j, i) for i in range(1, n + 1) for j in range(1, n + 1) if j != i]
    m = _build_opinion_with(ora, xs, edges, 0.5)
    vh.models.step_opinion(ora, m, eps=1.0)
    st, _ = _states(m)
    expect = (xs.sum() - xs) * 0.5 / (0.5 * (n - 1))
    assert np.array_equal(st["opinion"], expect)


def test_opinion_eps0_fixed_point_and_range(vh, ora):
    M = vh.models
    n = 2000
    m = M.build_opinion(ora, n, seed=3)
    x0, _ = _states(m)
    m0 = M.build_opinion(ora, n, seed=3)
    M.step_opinion(ora, m0, eps=0.0)     # distinct opinions: nobody is accepted
    x1, _ = _states(m0)
    assert np.array_equal(x0["opinion"], x1["opinion"])
    lo, hi = x0["opinion"].min(), x0["opinion"].max()
    for _ in range(5):
        M.step_opinion(ora, m, eps=0.2)
        x, _ = _states(m)
        assert x["opinion"].min() >= lo and x["opinion"].max() <= hi


def test_opinion_accept_boundary(vh, ora):
    # |x_j - x_i| <= eps is inclusive: with eps exactly the gap the neighbour counts
    xs = np.array([0.25, 0.5, 0.875])
    edges = [(2, 1), (3, 1), (1, 2), (3, 2), (1, 3), (2, 3)]
    m = _build_opinion_with(ora, xs, edges, [0.5, 0.25, 0.5, 0.25, 0.5, 0.25])
    vh.models.step_opinion(ora, m, eps=0.25)
    st, _ = _states(m)
    assert st["opinion"][0] == 0.5                 # only x2 (gap 0.25) accepted
    assert st["opinion"][1] == 0.25                # only x1 (gap 0.25) accepted
    assert st["opinion"][2] == 0.875               # nobody within 0.25: keeps x3


# ------------------------------------------------------------------ contact
def test_contact_invariants_and_lifo(vh, ora):
    M = vh.models
    n = 1500
    m = M.build_contact(ora, n, seed=6, n_init=15)
    prev_stack = []
    prev_died = np.zeros(n, dtype=bool)
    for step in range(8):
        M.step_contact(ora, m, seed=6, p_death=0.05)
        st, died = _states(m)
        assert (~died).sum() == n                                 # 1:1 births
        stack = m.sim.reuse_stack(m.T)
        ndeath = len(stack) - len(prev_stack) + 0
        # deaths of this step: live before (slots that existed) and died now
        old = len(prev_died)
        deaths = [i + 1 for i in range(old) if died[i] and not prev_died[i]]
        births = [i + 1 for i in range(len(died)) if not died[i] and (i >= old or prev_died[i])]
        # LIFO: births took the top of the previous stack, then fresh slots;
        # this step's deaths were pushed in ascending slot order
        k = min(len(births), len(prev_stack))
        reused = sorted(prev_stack[len(prev_stack) - k:])
        assert sorted(b for b in births if b <= old) == reused
        assert stack == prev_stack[:len(prev_stack) - k] + deaths
        assert ndeath == len(deaths) - k
        to, fr, _ = m.sim._read_edges(m.E)
        ton = (to & vh._abi.VH_NR_MASK).astype(np.int64)
        frn = (fr & vh._abi.VH_NR_MASK).astype(np.int64)
        assert not died[ton - 1].any() and not died[frn - 1].any()   # no edge touches the dead
        assert not (ton == frn).any()
        cnt = np.bincount(ton, minlength=len(died) + 1)
        assert cnt.max() <= 20
        assert all(cnt[b] == 0 for b in births)                       # newborns: no contacts yet
        prev_stack, prev_died = stack, died.copy()


# ------------------------------------------------------------------ mapreduce
def test_mapreduce_orders(vh, ora):
    M = vh.models
    n = 50000
    m = M.build_opinion(ora, n, m_ba=2, seed=8)
    x, _ = _states(m)
    xs = x["opinion"]
    m.sim.order = 0
    seq = ora.mapreduce(m.sim, "opinion", "+", m.T)
    m.sim.order = 1
    tree = ora.mapreduce(m.sim, "opinion", "+", m.T)
    acc = 0.0
    for v in xs:                        # the reference's left fold, as written
        acc = v + acc
    assert seq == acc
    assert abs(tree - math.fsum(xs)) <= n * 2.0 ** -53 * np.abs(xs).sum()
    assert abs(tree - seq) <= 2 * n * 2.0 ** -53 * np.abs(xs).sum()
    # integer ops are exact in any order
    m2 = M.build_sir(ora, 10000, seed=2, n_init=77)
    assert ora.mapreduce(m2.sim, ora.Field("health", 1), "+", m2.T) == 77
    st, _ = _states(m2)
    assert ora.mapreduce(m2.sim, "days", "+", m2.T) == int(st["days"].sum())
