"""Full-size runs of the benchmark configurations on the GPU.

* SIR at 1e5 agents x 100 steps: bit-exact against the checker.
* Opinion at 1e7 agents (BA m=5, ~1e8 edges), Float64 renumbered (the
  headline) and Float32, 10 steps each: every sampled agent's update (light
  and hub agents) recomputed from the pre-step state in numpy, in the
  canonical order, must be bit-identical; opinions stay in range.
* Contact at 1e6 and 1e7 agents: size-independent invariants of births,
  deaths and rebuilt contacts; at 1e5 agents bit-exact against the checker.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sir_config0_bit_exact(vh, ora, gpu_ready):
    M = vh.models
    n = 100_000
    edges = vh.inputs.graph_er(n, 10.0 / (n - 1), 1)
    g = M.build_sir(vh, n, seed=1, n_init=100, edges=edges)
    c = M.build_sir(ora, n, seed=1, n_init=100, edges=edges)
    for step in range(100):
        M.step_sir(vh, g, p_inf=0.05, seed=1)
        M.step_sir(ora, c, p_inf=0.05, seed=1)
        if step % 10 == 9:
            a, _ = g.sim._read_agents(g.T)
            b, _ = c.sim._read_agents(c.T)
            assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), step
    a, _ = g.sim._read_agents(g.T)
    h = np.asarray(a)["health"]
    print(f"[gpu] SIR 1e5 x 100: S {int((h == 0).sum())} I {int((h == 1).sum())} R {int((h == 2).sum())}",
          flush=True)


def _canon(nt, dt):
    """The canonical order of the opinion sums in the arrays' dtype:
    sequential up to degree 64, else 64 lane-strided sequential partials and
    the xor butterfly p[l] + p[l ^ o], o = 32..1 (np.cumsum accumulates
    sequentially in the array's dtype)."""
    d = len(nt)
    z = nt.dtype.type(0)
    if d <= 64:
        a = np.cumsum(nt)[-1] if d else z
        b = np.cumsum(dt)[-1] if d else z
        return a, b
    pa = np.array([np.cumsum(nt[l::64])[-1] if len(nt[l::64]) else z for l in range(64)], dtype=nt.dtype)
    pb = np.array([np.cumsum(dt[l::64])[-1] if len(dt[l::64]) else z for l in range(64)], dtype=dt.dtype)
    idx = np.arange(64)
    for o in (32, 16, 8, 4, 2, 1):
        pa = pa + pa[idx ^ o]
        pb = pb + pb[idx ^ o]
    return pa[0], pb[0]


def _opinion_sampled(vh, f32, relabel, steps):
    """10^7 agents (BA m=5 symmetrised, ~10^8 edges, the headline graph):
    every sampled agent's update (3000 random agents and the 200 of highest
    degree) recomputed from the pre-step states in numpy, in the canonical
    order and the state's precision, bit-identical, for `steps` steps."""
    M = vh.models
    n = 10_000_000
    a, b = vh.inputs.graph_ba(n, 5, 1)
    fr, to, w = vh.inputs.symmetric_weighted(a, b, 2)
    del a, b
    m = M.build_opinion(vh, n, seed=1, f32=f32, edges=(fr, to, w), relabel=relabel)
    # slot of each agent (generation numbering) in the simulation
    slot = (m.idmap.astype(np.int64) - 1) if relabel else np.arange(n, dtype=np.int64)
    order = np.argsort(to, kind="stable")
    src = fr[order].astype(np.int64) - 1
    ws = w[order]
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(to, minlength=n + 1)[1:])
    del order, fr, to, w
    deg = np.diff(off)
    rng = np.random.default_rng(0)
    hubs = np.argsort(deg)[-200:]
    sample = np.concatenate([rng.choice(n, 3000, replace=False), hubs])
    dt = np.float32 if f32 else np.float64
    eps = dt(0.2)

    def states():
        return np.asarray(m.sim._read_agents(m.T)[0]["opinion"])[slot].astype(dt)

    x0 = states()
    lo, hi = x0.min(), x0.max()
    for step in range(steps):
        M.step_opinion(vh, m, eps=0.2)
        x1 = states()
        for i in sample:
            s, e = off[i], off[i + 1]
            xj = x0[src[s:e]]
            acc = np.abs(xj - x0[i]) <= eps
            wj = ws[s:e].astype(dt)
            nt = np.where(acc, wj * xj, dt(0))
            dn = np.where(acc, wj, dt(0))
            num, den = _canon(nt, dn)
            want = num / den if den > 0 else x0[i]
            assert x1[i] == want, (step, int(i), int(deg[i]))
        assert x1.min() >= lo and x1.max() <= hi
        x0 = x1
    m.sim._destroy()
    print(f"[gpu] opinion 1e7 {'f32' if f32 else 'f64'}{' renumbered' if relabel else ''}: {len(sample)} "
          f"sampled agents (max degree {int(deg.max())}) exact over {steps} steps", flush=True)


def test_opinion_config1_sampled_exact(vh, gpu_ready):
    """The headline configuration (Float64, agents renumbered breadth-first,
    two-pass gather), 10 steps."""
    _opinion_sampled(vh, f32=False, relabel=True, steps=10)


def test_opinion_config1_f32_sampled_exact(vh, gpu_ready):
    """The Float32 variant at 10^7 agents (generation order), 10 steps."""
    _opinion_sampled(vh, f32=True, relabel=False, steps=10)


@pytest.mark.parametrize("n,steps", [(1_000_000, 4), (10_000_000, 5)])
def test_contact_invariants_large(vh, gpu_ready, n, steps):
    """Config c2 size (10^7 agents, 20 contacts, 2% deaths): population
    constant, no contact touches a dead agent or itself, at most 20 per
    agent, about 20 per survivor, deaths within 6 sigma of Binomial(N, 0.02),
    newborns without contacts and with susceptibility in [0.5, 1)."""
    M = vh.models
    print(f"[gpu] contact invariants n={n} steps={steps}", flush=True)
    m = M.build_contact(vh, n, seed=9)
    _, died0 = m.sim._read_agents(m.T)
    for step in range(steps):
        M.step_contact(vh, m, seed=9)
        st, died = m.sim._read_agents(m.T)
        assert int((~died).sum()) == n
        old = len(died0)
        deaths = int((died[:old] & ~died0).sum())
        assert abs(deaths - 0.02 * n) <= 6 * np.sqrt(n * 0.02 * 0.98), deaths
        to, fr, _ = m.sim._read_edges(m.E)
        ton = (to & vh._abi.VH_NR_MASK).astype(np.int64)
        frn = (fr & vh._abi.VH_NR_MASK).astype(np.int64)
        del to, fr
        assert not died[ton - 1].any() and not died[frn - 1].any()
        assert not (ton == frn).any()
        cnt = np.bincount(ton, minlength=len(died) + 1)
        assert cnt.max() <= 20
        # about 20 contacts per survivor, minus the ~2% that died as partners
        assert 0.9 * 20 * n * 0.98 < len(ton) <= 20 * n
        born = np.nonzero(~died & np.concatenate([died0, np.ones(len(died) - old, bool)]))[0]
        assert len(born) == deaths and (cnt[born + 1] == 0).all()
        sb = np.asarray(st["susc"])[born]
        assert (sb >= 0.5).all() and (sb < 1.0).all() and (np.asarray(st["health"])[born] == 0).all()
        del ton, frn, cnt
        died0 = died


def test_contact_1e5_bit_exact(vh, ora, gpu_ready):
    """The contact model at 10^5 agents for 20 steps: states, died flags (the
    slot map) and contact lists equal the checker's."""
    M = vh.models
    n = 100_000
    g = M.build_contact(vh, n, seed=13)
    c = M.build_contact(ora, n, seed=13)
    for step in range(20):
        M.step_contact(vh, g, seed=13)
        M.step_contact(ora, c, seed=13)
        if step % 5 == 4:
            a, da = g.sim._read_agents(g.T)
            b, db = c.sim._read_agents(c.T)
            assert len(a) == len(b) and np.array_equal(da, db), step
            live = ~da
            ab = np.ascontiguousarray(a).view(np.uint8).reshape(len(a), -1)[live]
            bb = np.ascontiguousarray(b).view(np.uint8).reshape(len(b), -1)[live]
            assert np.array_equal(ab, bb), step
            tg, fg, _ = g.sim._read_edges(g.E)
            tc, fc, _ = c.sim._read_edges(c.E)
            assert np.array_equal(tg, tc) and np.array_equal(fg, fc), step
    print(f"[gpu] contact 1e5 x 20: {int((~da).sum())} live, {len(tg)} contacts, bit-exact", flush=True)
