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


# ------------------------------------------------- brute-force replays (pure Python)
_G, _M64 = 0x9E3779B97F4A7C15, (1 << 64) - 1
KEY_RECOVER, KEY_DEATH, KEY_BIRTH, KEY_PARTNER = 1 << 40, 2 << 40, 3 << 40, 4 << 40


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _h4(s, a, b, c):
    h = _mix((s + _G) & _M64)
    for w in (a, b, c):
        h = _mix((h + w + _G) & _M64)
    return h


def _u01(h):
    return (h >> 11) * 2.0 ** -53


def _u01f(h):
    return np.float32(h >> 40) * np.float32(2.0 ** -24)


class _PyContact:
    """The contact model written out in plain Python from its specification:
    S -> I when some in-edge at 1-based position k has an infected source and
    u01(hash(seed, id, step, k)) < beta * susc (susc widened to Float64);
    I -> R with gamma (KEY_RECOVER); death with p_death (KEY_DEATH): the body
    returns nothing and adds one newborn {S, 0.5f + 0.5f * u01f(KEY_BIRTH)};
    a survivor draws n_contacts partners, draw r of contact c from
    hash(seed, id, step, KEY_PARTNER + c * max_redraw + r) as
    j = 1 + ((h >> 32) * S >> 32) over the S read slots, rejecting itself and
    slots dead in the read state, and adds the edge j -> id. Framework rules
    (src/AgentMethods.jl:37-89,300-439, src/EdgeMethods.jl:606-631,897-920):
    newborns take the reuse stack top (LIFO) else a fresh slot, in call
    order; this step's deaths are pushed in ascending slot order after the
    remaining stack; edges whose source or target died are dropped; the edge
    type is rebuilt every step (in `write` without add_existing)."""

    def __init__(self, st, seed, beta, gamma, p_death, n_contacts, max_redraw):
        self.h = [int(v) for v in st["health"]]
        self.s = [np.float32(v) for v in st["susc"]]
        self.died = [False] * len(self.h)
        self.reuse = []
        self.inedges = {}
        self.p = (seed, beta, gamma, p_death, n_contacts, max_redraw)

    def step(self, step):
        seed, beta, gamma, p_death, nc, mr = self.p
        S = len(self.h)
        new_h = list(self.h)
        births, deaths, emitted = [], [], {}
        for idx in range(1, S + 1):
            if self.died[idx - 1]:
                continue
            aid = (1 << 56) | idx
            h, susc = self.h[idx - 1], self.s[idx - 1]
            if h == 0:
                for k, j in enumerate(self.inedges.get(idx, []), start=1):
                    if self.h[j - 1] == 1 and _u01(_h4(seed, aid, step, k)) < beta * float(susc):
                        h = 1
                        break
            elif h == 1:
                if _u01(_h4(seed, aid, step, KEY_RECOVER)) < gamma:
                    h = 2
            if _u01(_h4(seed, aid, step, KEY_DEATH)) < p_death:
                births.append(np.float32(0.5) + np.float32(0.5) * _u01f(_h4(seed, aid, step, KEY_BIRTH)))
                deaths.append(idx)
                continue
            for c in range(nc):
                for r in range(mr):
                    hh = _h4(seed, aid, step, KEY_PARTNER + c * mr + r)
                    j = 1 + (((hh >> 32) * S) >> 32)
                    if j == idx or self.died[j - 1]:
                        continue
                    emitted.setdefault(idx, []).append(j)
                    break
            new_h[idx - 1] = h
        self.h = new_h
        for susc in births:
            if self.reuse:
                slot = self.reuse.pop()
            else:
                self.h.append(0)
                self.s.append(np.float32(0))
                self.died.append(False)
                slot = len(self.h)
            self.h[slot - 1], self.s[slot - 1], self.died[slot - 1] = 0, susc, False
        for d in deaths:
            self.died[d - 1] = True
        self.reuse = self.reuse + deaths
        dead = set(deaths)
        self.inedges = {t: [j for j in js if j not in dead] for t, js in emitted.items() if t not in dead}


@pytest.mark.parametrize("params", [
    dict(beta=0.9, gamma=0.3, p_death=0.1, n_contacts=4, max_redraw=3),
    dict(beta=0.6, gamma=0.2, p_death=0.35, n_contacts=6, max_redraw=1),    # exhausted redraws
    dict(beta=1.0, gamma=0.0, p_death=0.05, n_contacts=3, max_redraw=16),
])
def test_contact_brute_force_small(vh, ora, params):
    """The checker's contact body against the plain-Python replay, bit for
    bit: states (health, susceptibility bytes), died flags, reuse stack and
    every contact list, 14 steps of 48 agents."""
    M = vh.models
    n, seed = 48, 21
    m = M.build_contact(ora, n, seed=seed, n_init=6)
    st0, _ = _states(m)
    py = _PyContact(st0, seed, **params)
    events = {"deaths": 0, "infections": 0, "short": 0}
    for _ in range(14):
        step = m.sim._num_transitions()
        prev_s = sum(1 for h, d in zip(py.h, py.died) if h == 0 and not d)
        py.step(step)
        M.step_contact(ora, m, seed=seed, **params)
        st, died = _states(m)
        assert [int(v) for v in st["health"]] == py.h, step
        assert np.array_equal(np.asarray(st["susc"], dtype=np.float32).view(np.uint32),
                              np.array(py.s, dtype=np.float32).view(np.uint32)), step
        assert list(died) == py.died, step
        assert m.sim.reuse_stack(m.T) == py.reuse, step
        to, fr, _ = m.sim._read_edges(m.E)
        ton = [int(v) & int(vh._abi.VH_NR_MASK) for v in to]
        frn = [int(v) & int(vh._abi.VH_NR_MASK) for v in fr]
        want = [(t, j) for t in sorted(py.inedges) for j in py.inedges[t]]
        assert list(zip(ton, frn)) == want, step
        events["deaths"] += sum(py.died) > 0
        events["infections"] += prev_s - sum(1 for h, d in zip(py.h, py.died) if h == 0 and not d) > 0
        events["short"] += any(len(v) < params["n_contacts"] for v in py.inedges.values())
    assert events["deaths"] and events["infections"] and events["short"]


def test_contact_death_and_newborn_statistics(vh, ora):
    """At 10^5 agents: deaths per step within 5 sigma of Binomial(N, p_death),
    every newborn's susceptibility in [0.5, 1) and spread over it."""
    M = vh.models
    n, p = 100_000, 0.02
    m = M.build_contact(ora, n, seed=4)
    for _ in range(3):
        st0, died0 = _states(m)
        M.step_contact(ora, m, seed=4, p_death=p)
        st, died = _states(m)
        old = len(died0)
        deaths = int(np.sum(died[:old] & ~died0))
        assert abs(deaths - n * p) <= 5 * math.sqrt(n * p * (1 - p))
        born = np.array([i for i in range(len(died)) if not died[i] and (i >= old or died0[i])])
        s = np.asarray(st["susc"])[born]
        assert len(born) == deaths
        assert (s >= 0.5).all() and (s < 1.0).all()
        assert s.min() < 0.55 and s.max() > 0.95 and abs(s.mean() - 0.75) < 0.02
        assert (np.asarray(st["health"])[born] == 0).all()


def test_contact_infection_scales_with_susceptibility(vh, ora):
    """P(S -> I) = 1 - (1 - beta * susc)^c for c infected contacts: low- and
    high-susceptibility halves each match their prediction within 5 sigma
    (a body that drops susc, or uses it twice, fails one of them)."""
    M = vh.models
    n, beta = 40_000, 0.08
    m = M.build_contact(ora, n, seed=12, n_init=n // 2)
    kw = dict(seed=12, beta=beta, gamma=0.0, p_death=0.0, n_contacts=8)
    M.step_contact(ora, m, **kw)          # draws the contacts the next step reads
    st, _ = _states(m)
    h, susc = np.asarray(st["health"]), np.asarray(st["susc"]).astype(np.float64)
    to, fr, _ = m.sim._read_edges(m.E)
    ton = (to & vh._abi.VH_NR_MASK).astype(np.int64) - 1
    frn = (fr & vh._abi.VH_NR_MASK).astype(np.int64) - 1
    c = np.bincount(ton, weights=(h[frn] == 1), minlength=n).astype(np.int64)
    M.step_contact(ora, m, **kw)
    st2, _ = _states(m)
    newly = (h == 0) & (np.asarray(st2["health"]) == 1)
    pred = 1.0 - (1.0 - beta * susc) ** c
    sus = (h == 0) & (c > 0)
    for half in (susc < 0.75, susc >= 0.75):
        sel = sus & half
        obs, exp = newly[sel].sum(), pred[sel].sum()
        var = (pred[sel] * (1 - pred[sel])).sum()
        assert abs(obs - exp) <= 5 * math.sqrt(var), (obs, exp)
    assert not newly[(h == 0) & (c == 0)].any()


def _py_opinion_step(x, off, src, w, eps, dt):
    """One weighted bounded-confidence step in plain Python on scalars of
    type `dt` (np.float64 / np.float32): term w x_j and weight w when
    |x_j - x_i| <= eps (else +0), sums in the canonical order (sequential up
    to degree 64; above, 64 lane-strided partials then the xor butterfly
    p[l] + p[l ^ o], o = 32..1), result num / den, or x_i when den == 0."""
    out = []
    for i in range(len(x)):
        xi = dt(x[i])
        ks = range(off[i], off[i + 1])
        nt = [dt(w[k]) * dt(x[src[k]]) if abs(dt(x[src[k]]) - xi) <= dt(eps) else dt(0) for k in ks]
        de = [dt(w[k]) if abs(dt(x[src[k]]) - xi) <= dt(eps) else dt(0) for k in ks]
        d = len(nt)
        if d <= 64:
            a = b = dt(0)
            for k in range(d):
                a = dt(a + nt[k])
                b = dt(b + de[k])
        else:
            pa, pb = [dt(0)] * 64, [dt(0)] * 64
            for l in range(64):
                for k in range(l, d, 64):
                    pa[l] = dt(pa[l] + nt[k])
                    pb[l] = dt(pb[l] + de[k])
            for o in (32, 16, 8, 4, 2, 1):
                pa = [dt(pa[l] + pa[l ^ o]) for l in range(64)]
                pb = [dt(pb[l] + pb[l ^ o]) for l in range(64)]
            a, b = pa[0], pb[0]
        out.append(dt(a / b) if b > 0 else xi)
    return out


@pytest.mark.parametrize("f32", [False, True])
def test_opinion_body_replay(vh, ora, f32):
    """The checker's opinion body (Float64 and Float32) against the plain
    Python replay, bit for bit, over 4 steps on a graph with agents of
    in-degree 150, 90 and 65 (lane-strided sums), 51 and 64 (sequential sums)
    and light agents."""
    dt = np.float32 if f32 else np.float64
    rng = np.random.default_rng(17 + f32)
    n = 160
    xs = rng.random(n).astype(dt)
    edges = ([(j, 1) for j in range(2, 152)] + [(j, 2) for j in range(20, 110)]
             + [(j, 3) for j in range(10, 61)] + [(j, 4) for j in range(5, 69)]
             + [(j, 5) for j in range(90, 155)])
    for i in range(6, n + 1):
        for j in rng.choice(np.arange(1, n + 1), size=int(rng.integers(0, 9)), replace=False):
            if j != i:
                edges.append((int(j), i))
    w = rng.random(len(edges)).astype(np.float32) + np.float32(1e-3)
    m = _build_opinion_with(ora, xs, edges, w, f32=f32)
    to = np.array([b for _, b in edges])
    order = np.argsort(to, kind="stable")
    src = [edges[k][0] - 1 for k in order]
    ws = [w[k] for k in order]
    off = np.concatenate([[0], np.cumsum(np.bincount(to, minlength=n + 1)[1:])]).tolist()
    x = [dt(v) for v in xs]
    for _ in range(4):
        x = _py_opinion_step(x, off, src, ws, 0.3, dt)
        vh.models.step_opinion(ora, m, eps=0.3)
        st, _ = _states(m)
        assert np.array_equal(np.asarray(st["opinion"], dtype=dt).view(np.uint8),
                              np.array(x, dtype=dt).view(np.uint8))


def test_opinion_f32_consensus_closed_form(vh, ora):
    """Float32: complete graph K9, weights 0.5, eps >= 1: x_i' = (sum x - x_i) / 8
    exactly (dyadic opinions, every partial sum exact in Float32)."""
    n = 9
    xs = np.array([k / 1024.0 for k in (3, 100, 512, 77, 1000, 256, 9, 640, 1)], dtype=np.float32)
    edges = [(j, i) for i in range(1, n + 1) for j in range(1, n + 1) if j != i]
    m = _build_opinion_with(ora, xs, edges, 0.5, f32=True)
    vh.models.step_opinion(ora, m, eps=1.0)
    st, _ = _states(m)
    expect = ((xs.astype(np.float64).sum() - xs.astype(np.float64)) / 8).astype(np.float32)
    assert np.array_equal(np.asarray(st["opinion"]), expect)


def test_opinion_f32_accept_boundary(vh, ora):
    """Float32: |x_j - x_i| <= eps is inclusive (exact dyadic gaps)."""
    xs = np.array([0.25, 0.5, 0.875], dtype=np.float32)
    edges = [(2, 1), (3, 1), (1, 2), (3, 2), (1, 3), (2, 3)]
    m = _build_opinion_with(ora, xs, edges, [0.5, 0.25, 0.5, 0.25, 0.5, 0.25], f32=True)
    vh.models.step_opinion(ora, m, eps=0.25)
    st, _ = _states(m)
    assert list(np.asarray(st["opinion"])) == [np.float32(0.5), np.float32(0.25), np.float32(0.875)]
