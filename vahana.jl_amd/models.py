"""The three benchmark models built through the Vahana-style API.

Every builder takes the API module `V` (this package) and returns a small
namespace with the simulation and its types; the step functions issue one
apply! each.

* SIR contagion: `SIR` agents (:Immortal) {health::UInt8, days::UInt16},
  `Contact` edges (:Stateless, :SingleType) on a directed Erdos-Renyi graph.
* Weighted bounded confidence: `HK` agents {opinion::Float64 (or Float32)},
  `Knows` edges {w::Float32} (:SingleType) on a symmetrised Barabasi-Albert
  graph.
* Dynamic contacts: mortal `Person` agents {health::UInt8, susc::Float32},
  `Met` edges (:Stateless, :SingleType) rebuilt by every step.
* Raster diffusion: a side x side periodic raster of `Cell` agents
  (:Immortal) {active::Bool}, Moore neighbourhood `Adj` edges (:Stateless,
  :SingleType) from the device stencil; a step is test/raster.jl's diffuse.
"""
import types

import numpy as np

from . import inputs


def _types(V):
    SIR = V.struct("SIR", [("health", "u1"), ("pad", "u1"), ("days", "<u2")])
    Contact = V.struct("Contact")
    HK = V.struct("HK", [("opinion", "<f8")])
    HK32 = V.struct("HK32", [("opinion", "<f4")])
    Knows = V.struct("Knows", [("w", "<f4")])
    Person = V.struct("Person", [("health", "u1"), ("susc", "<f4")])
    Met = V.struct("Met")
    return SIR, Contact, HK, HK32, Knows, Person, Met


def build_sir(V, n, mean_deg=10.0, seed=1, n_init=100, graph_seed=None, edges=None,
              init="smallest"):
    SIR, Contact, *_ = _types(V)
    model = (V.ModelTypes()
             | V.register_agenttype_(SIR, "Immortal")
             | V.register_edgetype_(Contact, "Stateless", target=SIR)
             | V.create_model("sir"))
    sim = V.create_simulation(model)
    st0 = (inputs.sir_init(n, seed, n_init) if init == "smallest"
           else inputs.sir_init_frac(n, seed, n_init))
    ids = sim._add_agents_array(SIR, st0)
    if edges is None:
        p = mean_deg / max(n - 1, 1)
        edges = inputs.graph_er(n, p, seed if graph_seed is None else graph_seed)
    fr, to = edges
    tid = np.uint64(1) << np.uint64(56)
    sim._add_edges_array(Contact, tid | fr.astype(np.uint64), tid | to.astype(np.uint64))
    V.finish_init_(sim)
    return types.SimpleNamespace(sim=sim, T=SIR, E=Contact, ids=ids, nedges=len(to))


def step_sir(V, m, p_inf=0.05, seed=1, recovery=7):
    V.apply_(m.sim, V.transitions.sir(m.E, p_inf, seed, recovery), m.T, [m.T, m.E], m.T)


def build_opinion(V, n, m_ba=5, seed=1, f32=False, edges=None, relabel=False, device_graph=False):
    """`relabel`: agents renumbered breadth-first at finish_init! (m.idmap =
    new number of each agent number); a simulation without
    vh_finish_init_relabel (the checker) gets its inputs renumbered the same
    way (inputs.bfs_order). `device_graph`: the BA graph (m_ba, seed, weights
    seed + 1) is generated on the device at finish_init! (vh_add_edges_ba)
    instead of being added from `edges` / the host generator; the same graph."""
    _, _, HK, HK32, Knows, *_ = _types(V)
    A = HK32 if f32 else HK
    model = (V.ModelTypes()
             | V.register_agenttype_(A, "Immortal")
             | V.register_edgetype_(Knows, target=A)
             | V.create_model("opinion"))
    sim = V.create_simulation(model)
    x = inputs.opinion_init(n, seed, f32)
    st = np.zeros(n, dtype=A.dtype)
    st["opinion"] = x
    if device_graph and hasattr(sim.lib, "add_edges_ba"):
        ids = sim._add_agents_array(A, st)
        sim._add_edges_ba(Knows, m_ba, seed, seed + 1)
        idmap = V.finish_init_(sim, relabel=A) if relabel else V.finish_init_(sim)
        nedges = V.num_edges(sim, Knows)
        return types.SimpleNamespace(sim=sim, T=A, E=Knows, ids=ids, nedges=nedges, f32=f32, idmap=idmap)
    if edges is None:
        a, b = inputs.graph_ba(n, m_ba, seed)
        edges = inputs.symmetric_weighted(a, b, seed + 1)
    fr, to, w = edges
    idmap = None
    host_relabel = relabel and not hasattr(sim.lib, "finish_init_relabel")
    if host_relabel:
        idmap = inputs.bfs_order(n, fr, to)
        st2 = np.zeros_like(st)
        st2[(idmap - 1).astype(np.int64)] = st
        st = st2
        fr = idmap[fr.astype(np.int64) - 1]
        to = idmap[to.astype(np.int64) - 1]
    ids = sim._add_agents_array(A, st)
    tid = np.uint64(1) << np.uint64(56)
    ws = np.zeros(len(w), dtype=Knows.dtype)
    ws["w"] = w
    sim._add_edges_array(Knows, tid | fr.astype(np.uint64), tid | to.astype(np.uint64), ws)
    if relabel and not host_relabel:
        idmap = V.finish_init_(sim, relabel=A)
    else:
        V.finish_init_(sim)
    return types.SimpleNamespace(sim=sim, T=A, E=Knows, ids=ids, nedges=len(to), f32=f32,
                                 idmap=idmap)


def step_opinion(V, m, eps=0.2):
    V.apply_(m.sim, V.transitions.opinion(m.E, eps, m.f32), m.T, [m.T, m.E], m.T)


def build_opinion_gnp(V, n, mean_deg=10.0, seed=1, graph_seed=1, weight_seed=2, f32=False,
                      device_graph=True, stream=None, keep=None):
    """Opinion dynamics on a directed G(n, p) with Float32 weights. With
    `device_graph` the graph is generated on the device at finish_init!
    (vh_add_edges_gnp); otherwise its host twin is added edge by edge (the
    checker's path), giving the same CSR. `stream`: a raw hipStream_t the
    simulation is built on (and stays on). `keep(m)` receives the namespace as
    soon as the simulation exists (so that a caller can release it when a
    later build step fails)."""
    _, _, HK, HK32, Knows, *_ = _types(V)
    A = HK32 if f32 else HK
    model = (V.ModelTypes()
             | V.register_agenttype_(A, "Immortal")
             | V.register_edgetype_(Knows, target=A)
             | V.create_model("opinion"))
    sim = V.create_simulation(model)
    out = types.SimpleNamespace(sim=sim, T=A, E=Knows, ids=None, f32=f32)
    if keep is not None:
        keep(out)
    if stream is not None:
        sim._set_streams(stream, stream)
    st = np.zeros(n, dtype=A.dtype)
    st["opinion"] = inputs.opinion_init(n, seed, f32)
    ids = sim._add_agents_array(A, st)
    if device_graph:
        sim._add_edges_gnp(Knows, mean_deg, graph_seed, weight_seed)
    else:
        fr, to = inputs.graph_gnp(n, mean_deg, graph_seed)
        tid = np.uint64(1) << np.uint64(56)
        ws = np.zeros(len(to), dtype=Knows.dtype)
        ws["w"] = inputs.gnp_weights(len(to), weight_seed)
        sim._add_edges_array(Knows, tid | fr.astype(np.uint64), tid | to.astype(np.uint64), ws)
    V.finish_init_(sim)
    out.ids = ids
    return out


def step_opinion_n(V, m, steps, eps=0.2):
    """`steps` opinion apply! calls enqueued through one prepared descriptor."""
    if not hasattr(m, "prepared"):
        m.prepared = m.sim._prepare(V.transitions.opinion(m.E, eps, m.f32), m.T, [m.T, m.E], m.T)
        m.prepared_eps = eps
    assert m.prepared_eps == eps
    m.sim._apply_prepared(m.prepared, steps)


def build_contact(V, n, seed=1, n_init=None):
    *_, Person, Met = _types(V)
    model = (V.ModelTypes()
             | V.register_agenttype_(Person)
             | V.register_edgetype_(Met, "Stateless", target=Person)
             | V.create_model("contact"))
    sim = V.create_simulation(model)
    if n_init is None:
        n_init = max(1, n // 1000)
    ids = sim._add_agents_array(Person, inputs.contact_init(n, seed, n_init))
    V.finish_init_(sim)
    return types.SimpleNamespace(sim=sim, T=Person, E=Met, ids=ids)


def step_contact(V, m, beta=0.05, gamma=0.1, p_death=0.02, seed=1, n_contacts=20, max_redraw=16):
    V.apply_(m.sim, V.transitions.contact(m.E, beta, gamma, p_death, seed, n_contacts, max_redraw),
             m.T, [m.T, m.E], [m.T, m.E])


def build_raster_diffusion(V, side, seed=1, frac=0.01, device_stencil=True):
    """`side` x `side` periodic raster, a cell active when
    u01(hash(seed, id, 0, KEY_INIT)) < frac."""
    Cell = V.struct("Cell", [("active", "?")])
    Adj = V.struct("Adj")
    model = (V.ModelTypes() | V.register_agenttype_(Cell, "Immortal")
             | V.register_edgetype_(Adj, "Stateless", target=Cell) | V.create_model("raster"))
    sim = V.create_simulation(model)
    n = side * side
    u = np.empty(n, dtype=np.float64)
    inputs._abi.lib().fill_agent_u01(seed, 1, 0, inputs.KEY_INIT, n,
                                     u.ctypes.data_as(inputs.C.POINTER(inputs.C.c_double)))
    st = np.zeros(n, dtype=Cell.dtype)
    st["active"] = u < frac
    V.raster.add_raster_(sim, "grid", (side, side), st, T=Cell)
    # on the device the stencil is generated there (the checker adds the edges)
    V.raster.connect_raster_neighbors_(sim, "grid", Adj(), device=None if device_stencil else False)
    V.finish_init_(sim)
    return types.SimpleNamespace(sim=sim, T=Cell, E=Adj, n=n)


def step_raster_diffusion(V, m):
    V.apply_(m.sim, V.transitions.diffuse(m.E, m.T), m.T, [m.T, m.E], m.T)
