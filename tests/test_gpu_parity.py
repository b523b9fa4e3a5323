"""Parity of the HIP path with the checker (GPU).

Every scenario of the reference's own tests runs through the C ABI on cuda:0
and through the checker on the CPU; observations must be identical and equal
the reference's known answers.  The model bodies are compared state-by-state
(bit-exact, including the Float64/Float32 opinions, whose summation order is
fixed by the model spec).
"""
import numpy as np
import pytest

import scenarios as SC

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return repr(a) == repr(b)


@pytest.mark.parametrize("name", ["core_aggregate", "core_sum_neighbors", "core_queries",
                                  "edges_matrix", "edges_transition", "remove_agents",
                                  "independent", "addexisting", "graphs_k4",
                                  "edges_iter_removal", "raster_diffusion", "raster_nodeid",
                                  "raster_move_to_dist", "raster_random", "remove_edges"])
def test_reference_scenarios(vh, ora, gpu_ready, name):
    fn = getattr(SC, name)
    print(f"[gpu] scenario {name}", flush=True)
    got = fn(vh)
    want = fn(ora)
    assert _eq(got, want), f"{name}:\n device {got}\n checker {want}"


def _agents(m):
    st, died = m.sim._read_agents(m.T)
    return np.asarray(st), died


def _bytes(st):
    """Raw bytes per record (padding included).  Index the result, never the
    structured array: numpy's fancy indexing leaves padding uninitialised."""
    st = np.ascontiguousarray(st)
    return st.view(np.uint8).reshape(len(st), st.dtype.itemsize)


@pytest.mark.parametrize("expose", ["1", "0"])
def test_sir_parity(vh, ora, gpu_ready, expose, opts):
    """SIR, with (default) and without the exposure pass (option sir_expose)."""
    opts(sir_expose=int(expose), sir_expose_min=0)   # (3000 agents)
    print(f"[gpu] SIR parity expose={expose}", flush=True)
    M = vh.models
    n = 3000
    edges = vh.inputs.graph_er(n, 8.0 / (n - 1), 21)
    g = M.build_sir(vh, n, seed=3, n_init=30, edges=edges)
    c = M.build_sir(ora, n, seed=3, n_init=30, edges=edges)
    for step in range(25):
        M.step_sir(vh, g, p_inf=0.15, seed=3)
        M.step_sir(ora, c, p_inf=0.15, seed=3)
        a, _ = _agents(g)
        b, _ = _agents(c)
        assert np.array_equal(_bytes(a), _bytes(b)), f"step {step}"
    assert (a["health"] == vh.R).sum() > 0


# neighbour-state gather (options opinion_plan_bytes, opinion_part_bytes):
# one pass; two passes through a plan with 2 KiB source windows (16 for
# Float64, 8 for Float32 states), with 4 KiB windows, with the default
# 160 KiB (one window at this size), and with 1-byte windows requested
# (widened to the most a block's run table holds: 1000 windows of 4 agents)
GATHER_MODES = {"direct": (-1, 0), "plan": (0, 2048), "plan_4k": (0, 4096), "plan_auto": (0, 0),
                "plan_1src": (0, 1)}


@pytest.mark.parametrize("mode", list(GATHER_MODES))
@pytest.mark.parametrize("f32", [False, True])
def test_opinion_parity(vh, ora, gpu_ready, f32, mode, opts):
    plan_bytes, part_bytes = GATHER_MODES[mode]
    opts(opinion_plan_bytes=plan_bytes, opinion_part_bytes=part_bytes)
    print(f"[gpu] opinion parity f32={f32} gather={mode}", flush=True)
    M = vh.models
    n = 4000
    a, b = vh.inputs.graph_ba(n, 5, 8)
    edges = vh.inputs.symmetric_weighted(a, b, 9)
    deg = np.bincount(edges[1], minlength=n + 1)
    assert deg.max() > 64       # the wave-per-agent path is exercised
    g = M.build_opinion(vh, n, seed=8, f32=f32, edges=edges)
    c = M.build_opinion(ora, n, seed=8, f32=f32, edges=edges)
    for step in range(8):
        M.step_opinion(vh, g, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
        x, _ = _agents(g)
        y, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}"


@pytest.mark.parametrize("mode", ["direct", "plan", "plan_1src"])
@pytest.mark.parametrize("f32", [False, True])
def test_opinion_parity_hub(vh, ora, gpu_ready, f32, mode, opts):
    """Every in-degree class of the gather: a Barabasi-Albert graph (degrees
    up to a few hundred: agents summed by their own thread and by a wave)
    with two hubs of 1500 and 3000 undirected neighbours added (in-degree
    above 1024: the chunk lists)."""
    plan_bytes, part_bytes = GATHER_MODES[mode]
    opts(opinion_plan_bytes=plan_bytes, opinion_part_bytes=part_bytes)
    print(f"[gpu] opinion parity with hubs f32={f32} gather={mode}", flush=True)
    M = vh.models
    n = 4000
    a, b = vh.inputs.graph_ba(n, 5, 8)
    ha = np.concatenate([np.full(3000, 17, np.uint32), np.full(1500, 900, np.uint32)])
    hb = np.concatenate([np.arange(1000, 4000, dtype=np.uint32), np.arange(2, 1502, dtype=np.uint32)])
    edges = vh.inputs.symmetric_weighted(np.concatenate([a, ha]), np.concatenate([b, hb]), 9)
    deg = np.bincount(edges[1], minlength=n + 1)
    assert deg.max() > 3000 and ((deg > 64) & (deg <= 1024)).sum() > 5
    g = M.build_opinion(vh, n, seed=8, f32=f32, edges=edges)
    c = M.build_opinion(ora, n, seed=8, f32=f32, edges=edges)
    for step in range(6):
        M.step_opinion(vh, g, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
        x, _ = _agents(g)
        y, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}"


@pytest.mark.parametrize("mode", ["plan", "plan_auto"])
@pytest.mark.parametrize("f32", [False, True])
def test_opinion_parity_isolated(vh, ora, gpu_ready, f32, mode, opts):
    """Agents without in-edges: a 3000-agent Barabasi-Albert graph inside
    40000 agents, so that the gather plan's blocks of isolated agents hold
    up to ~12000 agents (many rounds of the block scan) and no staged edges
    (no runs), next to blocks with both."""
    plan_bytes, part_bytes = GATHER_MODES[mode]
    opts(opinion_plan_bytes=plan_bytes, opinion_part_bytes=part_bytes)
    print(f"[gpu] opinion parity with isolated agents f32={f32} gather={mode}", flush=True)
    M = vh.models
    n = 40_000
    a, b = vh.inputs.graph_ba(3000, 5, 8)
    edges = vh.inputs.symmetric_weighted(a + 17_000, b + 17_000, 9)   # agents 17001..20000
    deg = np.bincount(edges[1], minlength=n + 1)
    assert (deg[1:] == 0).sum() > 30_000
    g = M.build_opinion(vh, n, seed=8, f32=f32, edges=edges)
    c = M.build_opinion(ora, n, seed=8, f32=f32, edges=edges)
    for step in range(4):
        M.step_opinion(vh, g, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
        x, _ = _agents(g)
        y, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}"


@pytest.mark.parametrize("f32", [False, True])
def test_opinion_prepare(vh, ora, gpu_ready, f32, opts):
    """vh_opinion_prepare builds the plan ahead of the first apply! (twice in
    a row: the second keeps it); the steps after it match the checker's, and
    a state type that does not match the transition's is refused."""
    opts(opinion_plan_bytes=0, opinion_part_bytes=2048)
    print(f"[gpu] opinion prepare f32={f32}", flush=True)
    M = vh.models
    n = 4000
    a, b = vh.inputs.graph_ba(n, 5, 8)
    edges = vh.inputs.symmetric_weighted(a, b, 9)
    g = M.build_opinion(vh, n, seed=8, f32=f32, edges=edges)
    c = M.build_opinion(ora, n, seed=8, f32=f32, edges=edges)
    with pytest.raises(vh.VahanaError):
        g.sim._opinion_prepare(g.T, g.E, not f32)
    g.sim._opinion_prepare(g.T, g.E, f32)
    g.sim._opinion_prepare(g.T, g.E, f32)
    for step in range(3):
        M.step_opinion(vh, g, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
        x, _ = _agents(g)
        y, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}"


@pytest.mark.parametrize("f32", [False, True])
def test_opinion_plan_after_edge_change(vh, ora, gpu_ready, f32, opts):
    """The two-pass plan (its lists and the edge weights copied into list
    order) follows a change of the edge store: after three steps every agent
    adds a self-loop of weight 0.75 (add_existing), and the next steps must
    match the checker's."""
    opts(opinion_plan_bytes=0, opinion_part_bytes=2048)
    print(f"[gpu] opinion plan after edge change f32={f32}", flush=True)
    M = vh.models
    n = 4000
    a, b = vh.inputs.graph_ba(n, 5, 8)
    edges = vh.inputs.symmetric_weighted(a, b, 9)
    g = M.build_opinion(vh, n, seed=8, f32=f32, edges=edges)
    c = M.build_opinion(ora, n, seed=8, f32=f32, edges=edges)
    ne = vh.num_edges(g.sim, g.E)
    for step in range(6):
        if step == 3:
            for V, m in ((vh, g), (ora, c)):
                st = np.zeros(1, dtype=m.E.dtype)
                st["w"] = 0.75
                V.apply_(m.sim, V.transitions.add_self_edge(m.E, st), [m.T], [], [m.E], add_existing=[m.E])
            assert vh.num_edges(g.sim, g.E) == ne + n == ora.num_edges(c.sim, c.E)
        M.step_opinion(vh, g, eps=0.2)
        M.step_opinion(ora, c, eps=0.2)
        x, _ = _agents(g)
        y, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}"


# (contacts, redraws, beta, p_death): the default model; more contacts than
# one LDS row holds (the unstaged path); no contacts; a single draw per
# contact under heavy deaths (draws exhausted on dead partners); certain
# transmission (every infected neighbour is looked up)
CONTACT_CASES = [(20, 16, 0.05, 0.05), (40, 4, 0.05, 0.05), (0, 16, 0.05, 0.05), (5, 1, 0.05, 0.3),
                 (20, 16, 1.0, 0.02)]


@pytest.mark.parametrize("case", CONTACT_CASES, ids=lambda c: "C%d_R%d_b%g_d%g" % c)
def test_contact_parity(vh, ora, gpu_ready, case):
    M = vh.models
    C, R, beta, pd = case
    n = 3000
    print(f"[gpu] contact parity {case}", flush=True)
    g = M.build_contact(vh, n, seed=5, n_init=30)
    c = M.build_contact(ora, n, seed=5, n_init=30)
    kw = dict(seed=5, beta=beta, p_death=pd, n_contacts=C, max_redraw=R)
    for step in range(10):
        M.step_contact(vh, g, **kw)
        M.step_contact(ora, c, **kw)
        a, da = _agents(g)
        b, db = _agents(c)
        assert len(a) == len(b), f"step {step}: slots {len(a)} vs {len(b)}"
        assert np.array_equal(da, db), f"step {step}: died flags"
        live = ~da
        assert np.array_equal(_bytes(a)[live], _bytes(b)[live]), f"step {step}: states"
        tg, fg, _ = g.sim._read_edges(g.E)
        tc, fc, _ = c.sim._read_edges(c.E)
        assert np.array_equal(tg, tc) and np.array_equal(fg, fc), f"step {step}: contacts"
        assert vh.num_agents(g.sim, g.T) == n


def test_mapreduce_tree_parity(vh, ora, gpu_ready):
    M = vh.models
    n = 20000
    a, b = vh.inputs.graph_ba(n, 3, 2)
    edges = vh.inputs.symmetric_weighted(a, b, 4)
    g = M.build_opinion(vh, n, m_ba=3, seed=2, edges=edges)
    c = M.build_opinion(ora, n, m_ba=3, seed=2, edges=edges)
    for op in ("+", "*", "max", "min"):
        x = vh.mapreduce(g.sim, "opinion", op, g.T)
        y = ora.mapreduce(c.sim, "opinion", op, c.T)
        assert x == y or (np.isnan(x) and np.isnan(y)), op
    xw = vh.mapreduce(g.sim, "w", "+", g.E)
    assert xw == ora.mapreduce(c.sim, "w", "+", c.E)


def test_opinion_device_gnp_graph(vh, ora, gpu_ready):
    """vh_add_edges_gnp builds on the device the CSR that its host twin
    (graph_gnp + gnp_weights, added edge by edge) gives; the model then
    steps bit-identically, through apply_n as through apply!, and equals the
    checker."""
    M = vh.models
    n = 5000
    print("[gpu] opinion on device G(n, p)", flush=True)
    d = M.build_opinion_gnp(vh, n, mean_deg=10.0, seed=3, graph_seed=4, weight_seed=5)
    h = M.build_opinion_gnp(vh, n, mean_deg=10.0, seed=3, graph_seed=4, weight_seed=5, device_graph=False)
    c = M.build_opinion_gnp(ora, n, mean_deg=10.0, seed=3, graph_seed=4, weight_seed=5, device_graph=False)
    td, fd, wd = d.sim._read_edges(d.E)
    th, fh, wh = h.sim._read_edges(h.E)
    assert len(td) > 9 * n
    assert np.array_equal(td, th) and np.array_equal(fd, fh)
    assert np.array_equal(_bytes(wd), _bytes(wh))
    for step in range(6):
        M.step_opinion_n(vh, d, 1, eps=0.25)
        M.step_opinion(vh, h, eps=0.25)
        M.step_opinion(ora, c, eps=0.25)
        x, _ = _agents(d)
        y, _ = _agents(h)
        z, _ = _agents(c)
        assert np.array_equal(_bytes(x), _bytes(y)), f"step {step}: device graph vs host twin"
        assert np.array_equal(_bytes(x), _bytes(z)), f"step {step}: vs checker"
    M.step_opinion_n(vh, d, 5, eps=0.25)
    for _ in range(5):
        M.step_opinion(vh, h, eps=0.25)
    assert np.array_equal(_bytes(_agents(d)[0]), _bytes(_agents(h)[0])), "apply_n(5)"


def test_device_gnp_argument_errors(vh, gpu_ready):
    M = vh.models
    d = M.build_opinion_gnp(vh, 100, mean_deg=4.0)
    with pytest.raises(AssertionError):
        d.sim._add_edges_gnp(d.E, 4.0, 1, 2)          # after finish_init!
    _, _, HK, _, Knows, Person, _ = M._types(vh)
    model = (vh.ModelTypes() | vh.register_agenttype_(HK, "Immortal")
             | vh.register_agenttype_(Person) | vh.register_edgetype_(Knows, target=HK)
             | vh.create_model("two types"))
    sim = vh.create_simulation(model)
    with pytest.raises(vh.VahanaError):
        sim._add_edges_gnp(Knows, 4.0, 1, 2)           # more than one agent type
    model = (vh.ModelTypes() | vh.register_agenttype_(HK, "Immortal")
             | vh.register_edgetype_(Knows, target=HK) | vh.create_model("twice"))
    sim = vh.create_simulation(model)
    sim._add_edges_gnp(Knows, 4.0, 1, 2)
    with pytest.raises(vh.VahanaError):
        sim._add_edges_gnp(Knows, 4.0, 1, 2)           # generated twice
    with pytest.raises(vh.VahanaError):
        sim._add_edges_array(Knows, np.array([1 << 56 | 1], dtype=np.uint64),
                             np.array([1 << 56 | 1], dtype=np.uint64))


def test_raster_device_stencil_known_answers(vh, ora, gpu_ready):
    """The raster known answers with the stencil edges generated on the
    device (vh_add_edges_stencil) equal the checker's host-built run."""
    got = SC.raster_diffusion(vh, uniform=True)
    want = SC.raster_diffusion(ora)
    assert got == want
    assert got["2d_10x8"] == [1, 9, 25] and got["3d_euclidean"] == [1, 19]


def test_raster_stencil_diffuse_known_answers(vh, ora, gpu_ready):
    """With :Immortal cells on a device stencil, diffuse runs the stencil
    kernel (neighbours from the stencil arithmetic, not the CSR): the known
    answers of every metric, distance and periodicity case are unchanged."""
    print("[gpu] raster stencil diffuse", flush=True)
    got = SC.raster_diffusion(vh, uniform=True, immortal=True)
    want = SC.raster_diffusion(ora)
    assert got == want
    assert got["2d_10x8"] == [1, 9, 25] and got["3d_euclidean"] == [1, 19]


STENCIL_CASES = [((10, 8), 1, "chebyshev", True), ((10, 8), 1, "chebyshev", False),
                 ((2, 3), 1, "chebyshev", True),       # wrapped duplicates
                 ((1, 5), 2, "chebyshev", True),
                 ((13,), 3, "chebyshev", True), ((6, 7, 6), 2, "manhatten", False),
                 ((6, 7, 6), 1.5, "euclidean", True), ((5, 4, 3, 3), 1, "chebyshev", True),
                 ((40, 30), 2.5, "euclidean", False)]


@pytest.mark.parametrize("case", STENCIL_CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_{c[2]}{c[1]}"
                         + ("_p" if c[3] else ""))
def test_raster_device_stencil_csr(vh, gpu_ready, case):
    """The device stencil CSR equals the host path's, edge by edge (targets,
    sources in insertion order, states), in a model with other agent types
    before the cells and a stateful edge type."""
    dims, dist, metric, periodic = case
    Other = vh.struct("Other", [("x", "<i8")])
    Cell = vh.struct("Cell", [("v", "<i4")])
    Link = vh.struct("Link", [("w", "<f4"), ("tag", "<i4")])
    model = (vh.ModelTypes() | vh.register_agenttype_(Other) | vh.register_agenttype_(Cell)
             | vh.register_edgetype_(Link) | vh.create_model("stencil"))
    res = []
    for device in (True, False):
        sim = vh.create_simulation(model)
        vh.add_agents_(sim, [Other(i) for i in range(3)])
        vh.raster.add_raster_(sim, "r", dims, lambda p: Cell(sum(p)))
        vh.raster.connect_raster_neighbors_(sim, "r", Link(0.5, 7), distance=dist, metric=metric,
                                            periodic=periodic, device=device)
        vh.finish_init_(sim)
        to, fr, st = sim._read_edges(Link)
        res.append((to, fr, np.ascontiguousarray(st).view(np.uint8)))
        vh.finish_simulation_(sim)
    (a, b, c), (x, y, z) = res
    assert len(a) > 0
    assert np.array_equal(a, x) and np.array_equal(b, y) and np.array_equal(c, z)


def test_raster_diffusion_model_parity(vh, ora, gpu_ready):
    """The benchmark's raster diffusion (device stencil, diffuse body) equals
    the checker's (host-added stencil edges) step by step."""
    M = vh.models
    g = M.build_raster_diffusion(vh, 300, seed=4, frac=0.002)
    c = M.build_raster_diffusion(ora, 300, seed=4, frac=0.002)
    tg, fg, _ = g.sim._read_edges(g.E)
    tc, fc, _ = c.sim._read_edges(c.E)
    assert np.array_equal(tg, tc) and np.array_equal(fg, fc)
    for step in range(6):
        M.step_raster_diffusion(vh, g)
        M.step_raster_diffusion(ora, c)
        a, _ = _agents(g)
        b, _ = _agents(c)
        assert np.array_equal(_bytes(a), _bytes(b)), step


def test_snapshot_resume_device(vh, gpu_ready, tmp_path):
    out = SC.snapshot_sir_resume(vh, str(tmp_path / "sir"))
    assert out["t"] == out["t_read"] == 5 and out["resumed"] == 6
    assert out["same"]
    assert out["list"] == [(5, "after five")]


def test_snapshot_cross_restore_contact(vh, ora, gpu_ready, tmp_path):
    """A device snapshot of the mortal contact model restored into the
    device and into the checker (free slots rebuilt ascending on both):
    four more steps stay bit-identical, slot map included."""
    import types
    path = str(tmp_path / "c")
    out = SC.snapshot_contact_restore(vh, path)
    assert out["died"] and out["states"] and out["edges"] and out["free"]
    g = out["b"]
    model, Person, Met = SC.contact_model(ora)
    c = types.SimpleNamespace(sim=ora.create_simulation(model), T=Person, E=Met)
    ora.snapshot.read_snapshot(c.sim, path)
    assert c.sim.reuse_stack(c.T) == out["free"]
    M = vh.models
    for step in range(4):
        M.step_contact(vh, g, seed=8, p_death=0.05)
        M.step_contact(ora, c, seed=8, p_death=0.05)
        a, da = _agents(g)
        b, db = _agents(c)
        assert np.array_equal(da, db), step
        assert np.array_equal(_bytes(a)[~da], _bytes(b)[~db]), step
        tg, fg, _ = g.sim._read_edges(g.E)
        tc, fc, _ = c.sim._read_edges(c.E)
        assert np.array_equal(tg, tc) and np.array_equal(fg, fc), step


def test_relabelled_opinion_device(vh, ora, gpu_ready):
    """finish_init_(relabel=T) on the device: the mapping equals the
    checker's host renumbering, the run equals the checker's relabelled run
    bit for bit and the device's plain run agent by agent."""
    M = vh.models
    a, b = vh.inputs.graph_ba(20000, 5, 7)
    edges = vh.inputs.symmetric_weighted(a, b, 8)
    d = M.build_opinion(vh, 20000, seed=7, edges=edges, relabel=True)
    p = M.build_opinion(vh, 20000, seed=7, edges=edges)
    c = M.build_opinion(ora, 20000, seed=7, edges=edges, relabel=True)
    assert np.array_equal(d.idmap, c.idmap)
    for _ in range(4):
        for m, V in ((d, vh), (p, vh), (c, ora)):
            M.step_opinion(V, m, eps=0.2)
    xd = np.asarray(d.sim._read_agents(d.T)[0]["opinion"])
    xp = np.asarray(p.sim._read_agents(p.T)[0]["opinion"])
    xc = np.asarray(c.sim._read_agents(c.T)[0]["opinion"])
    assert np.array_equal(xd, xc)
    assert np.array_equal(xd[(d.idmap - 1).astype(np.int64)], xp)
    assert vh.updateids(d.idmap, vh.agent_id(1, 5)) == vh.agent_id(1, int(d.idmap[4]))


@pytest.mark.parametrize("variant", ["default", "immortal", "neighbors"])
def test_snapshot_hints_device(vh, ora, gpu_ready, tmp_path, variant):
    """The three models of test/hdf5_snapshot.jl on the device: the snapshot
    files (index with the skipped records, every record array) equal the
    checker's; the device snapshot restores into the device and into the
    checker, and both advance like the simulation that wrote it."""
    import json
    import os
    print(f"[gpu] snapshot hints {variant}", flush=True)
    pg, pc = str(tmp_path / "g"), str(tmp_path / "c")
    mg, tg = SC.snapshot_hint_model(vh, variant)
    mc, tc = SC.snapshot_hint_model(ora, variant)
    g = SC.snapshot_hint_runsim(vh, mg, tg, pg)
    c = SC.snapshot_hint_runsim(ora, mc, tc, pc)
    with open(os.path.join(pg, "index.json")) as a, open(os.path.join(pc, "index.json")) as b:
        assert json.load(a) == json.load(b)
    for kind in ("agents", "edges"):
        for name in sorted(os.listdir(os.path.join(pc, kind))):
            for f in sorted(os.listdir(os.path.join(pc, kind, name))):
                with np.load(os.path.join(pg, kind, name, f)) as za, \
                        np.load(os.path.join(pc, kind, name, f)) as zb:
                    assert sorted(za.files) == sorted(zb.files), (kind, name, f)
                    for k in za.files:
                        x, y = za[k], zb[k]
                        if kind == "agents" and k == "state" and "died" in za.files and x.size:
                            # a died slot's bytes are not state
                            live = ~za["died"].astype(bool)
                            x = x.reshape(len(live), -1)[live]
                            y = y.reshape(len(live), -1)[live]
                        assert np.array_equal(x, y), (kind, name, f, k)
    rg = SC.snapshot_hint_restore(vh, mg, pg)
    rc = SC.snapshot_hint_restore(ora, mc, pg)
    want = SC.snapshot_hint_state(ora, c, tc)
    assert SC.snapshot_hint_state(vh, rg, tg) == want
    assert SC.snapshot_hint_state(ora, rc, tc) == want
    for V, s, t in ((vh, g, tg), (vh, rg, tg), (ora, c, tc), (ora, rc, tc)):
        SC.snapshot_hint_raster_step(V, s, t)
    want = SC.snapshot_hint_state(ora, c, tc)
    for V, s, t in ((vh, g, tg), (vh, rg, tg), (ora, rc, tc)):
        assert SC.snapshot_hint_state(V, s, t) == want
