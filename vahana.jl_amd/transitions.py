"""Device transition bodies compiled into libvahana_hip.so.

Each factory returns a :class:`Transition` naming the body, the agent/edge
types bound to its role parameters and its parameter blob — the counterpart
of the Julia closures the reference passes to apply! (src/Simulation.jl:720).
"""
import ctypes as C

from . import _abi


class ValueRef:
    """param(sim, name) (a parameter) or get_global(sim, name) (a global)
    read by a transition body: the parameter-blob field that receives it is
    filled from the simulation's value when apply! is enqueued
    (vh_value_bind). Both are constant while a transition runs."""

    def __init__(self, kind, name):
        self.kind = kind
        self.name = str(name).lstrip(":")

    def __repr__(self):
        return f"{'param' if self.kind == _abi.VALUE_PARAM else 'get_global'}(sim, :{self.name})"


def param_ref(name):
    return ValueRef(_abi.VALUE_PARAM, name)


def global_ref(name):
    return ValueRef(_abi.VALUE_GLOBAL, name)


# vh_dtype of the ctypes field types of the parameter blobs
_CT_DT = {C.c_double: _abi.DT["f8"], C.c_float: _abi.DT["f4"], C.c_int64: _abi.DT["i8"],
          C.c_uint64: _abi.DT["u8"], C.c_uint32: _abi.DT["u4"], C.c_int32: _abi.DT["i4"],
          C.c_uint16: _abi.DT["u2"], C.c_uint8: _abi.DT["u1"]}


def _set_fields(p, **fields):
    """Set the blob fields; a ValueRef leaves the field 0 and becomes a
    binding (byte offset, field dtype, reference)."""
    binds = []
    ftypes = dict(p._fields_)
    for k, v in fields.items():
        if isinstance(v, ValueRef):
            binds.append((getattr(type(p), k).offset, _CT_DT[ftypes[k]], v))
        else:
            setattr(p, k, float(v) if ftypes[k] in (C.c_double, C.c_float) else int(v))
    return binds


class Transition:
    def __init__(self, id_, name, role_agents=(), role_edges=(), params=None, binds=()):
        self.id = id_
        self.name = name
        self.role_agents = list(role_agents)
        self.role_edges = list(role_edges)
        self.params = params
        self.binds = list(binds)   # (offset, dtype, ValueRef)

    def params_bytes(self):
        if self.params is None:
            return b""
        if isinstance(self.params, (bytes, bytearray)):
            return bytes(self.params)
        return bytes(C.string_at(C.addressof(self.params), C.sizeof(self.params)))

    def __repr__(self):
        return f"Transition({self.name})"


def _tp(value=0, cmp="always", edge_state=b""):
    p = _abi.TestParams()
    p.value = int(value)
    p.cmp = _abi.CMP[cmp]
    for i, b in enumerate(bytes(edge_state)[:16]):
        p.edge_state[i] = b
    return p


def noop():
    """Returns the state unchanged."""
    return Transition(_abi.TR_NOOP, "noop", params=_tp())


def sum_neighbors(edgetype):
    """foo = sum of neighborstates_flexible(sim, id, edgetype) .foo (test/core.jl:71-83)."""
    return Transition(_abi.TR_SUM_NEIGHBORS, "sum_neighbors", role_edges=[edgetype], params=_tp())


def die_if(cmp, value=0, edgetype=None):
    """`nothing` (death) when the predicate on the Int64 field `foo` holds
    (cmp: always, never, lt, ge, eq, odd, even, num_edges_eq); `value` may be
    a param_ref / global_ref."""
    p = _tp(0, cmp)
    binds = _set_fields(p, value=value)
    return Transition(_abi.TR_DIE_IF, f"die_if_{cmp}", role_edges=[edgetype], params=p, binds=binds)


def add_self_edge(edgetype, state=None):
    """add_edge!(sim, id, id, state)."""
    raw = b"" if state is None else bytes(state.tobytes() if hasattr(state, "tobytes") else state)
    return Transition(_abi.TR_ADD_SELF_EDGE, "add_self_edge", role_edges=[edgetype],
                      params=_tp(edge_state=raw))


def readd_edges(edgetype):
    """add_edges!(sim, id, edges(sim, id, edgetype)) (test/edges.jl:357-359)."""
    return Transition(_abi.TR_READD_EDGES, "readd_edges", role_edges=[edgetype], params=_tp())


def neighbor_foo_edges(read_edgetype, write_edgetype, die_value):
    """test/independent.jl:55-68: die when foo == die_value, else add foo-valued
    edges to and from every neighbour."""
    return Transition(_abi.TR_NEIGHBOR_FOO_EDGES, "neighbor_foo_edges",
                      role_edges=[read_edgetype, write_edgetype], params=_tp(die_value))


def spawn(agenttype, count, edgetype=None, both_ways=True):
    """test/independent.jl:165-176: `count` newborns T(foo + 10j), each linked
    with an edge carrying the newborn id (newborn -> caller, and caller ->
    newborn when both_ways)."""
    p = _tp(count)
    p.cmp = 0 if both_ways else 1
    return Transition(_abi.TR_SPAWN, "spawn", role_agents=[agenttype], role_edges=[edgetype],
                      params=p)


def remove_edges(mode, edgetype, sendid=None, state=None):
    """The remove_edges! bodies of test/mpi/test_edgetypes.jl:295-445
    (src/EdgeMethods.jl:527-599), `mode` one of own, neighbor, readd,
    from_two, to_neighbor, send_id, to_sendid, zero_from (enum vh_remove_mode);
    `state`: the edge state of the edges added by readd / send_id."""
    raw = b"" if state is None else bytes(state.tobytes() if hasattr(state, "tobytes") else state)
    edges = [edgetype] if sendid is None else [edgetype, sendid]
    return Transition(_abi.TR_REMOVE_EDGES, f"remove_edges_{mode}", role_edges=edges,
                      params=_tp(_abi.RM[mode], edge_state=raw))


def count_edges(edgetype):
    """foo = num_edges(sim, id, edgetype)."""
    return Transition(_abi.TR_COUNT_EDGES, "count_edges", role_edges=[edgetype], params=_tp())


def sir(edgetype, p_inf, seed, recovery=7):
    """SIR contagion over Stateless in-edges (any argument may be a
    param_ref / global_ref)."""
    p = _abi.SirParams()
    binds = _set_fields(p, p_inf=p_inf, seed=seed, recovery=recovery)
    return Transition(_abi.TR_SIR, "sir", role_edges=[edgetype], params=p, binds=binds)


def opinion(edgetype, eps, f32=False):
    """Weighted bounded confidence (Float64 opinions, or Float32 with f32=True);
    `eps` may be a param_ref / global_ref."""
    p = _abi.OpinionParams()
    binds = _set_fields(p, eps=eps)
    return Transition(_abi.TR_OPINION_F32 if f32 else _abi.TR_OPINION_F64,
                      "opinion_f32" if f32 else "opinion_f64", role_edges=[edgetype], params=p,
                      binds=binds)


def contact(edgetype, beta, gamma, p_death, seed, n_contacts=20, max_redraw=16):
    """Dynamic contact network with deaths, 1:1 births and rebuilt contacts
    (any argument may be a param_ref / global_ref)."""
    p = _abi.ContactParams()
    binds = _set_fields(p, beta=beta, gamma=gamma, p_death=p_death, seed=seed, n_contacts=n_contacts,
                        max_redraw=max_redraw)
    return Transition(_abi.TR_CONTACT, "contact", role_edges=[edgetype], params=p, binds=binds)


def diffuse(edgetype, T, field="active"):
    """a.field || any(n.field for n in neighborstates(sim, id, edgetype)) on a
    Bool field of T (test/raster.jl:38-41, 143-146)."""
    p = _abi.RasterParams()
    p.field_offset = T.dtype.fields[field][1]
    return Transition(_abi.TR_DIFFUSE, "diffuse", role_edges=[edgetype], params=p)


def move_to_value(sim, raster, edgetype, from_raster=True, to_raster=True, edge_state=None):
    """value_on_pos (test/raster.jl:279-285): v = the Int64 at offset 0 of the
    first in-neighbour over `edgetype`; move_to!(sim, raster, id, (v, ..., v),
    ...) with distance 0; the new state is v. The raster's cells must be
    consecutive agents of one type (as add_raster_ makes them)."""
    from .raster import _rasters
    R = _rasters(sim)[raster]
    contig = R.contiguous()
    if contig is None or len(R.dims) > 4:
        raise AssertionError("move_to_value needs a raster of consecutive cells, <= 4 dimensions")
    p = _abi.RasterParams()
    p.cell_type, p.first_nr = contig
    p.ndims = len(R.dims)
    for i, d in enumerate(R.dims):
        p.dims[i] = d
    p.flags = (1 if from_raster else 0) | (2 if to_raster else 0)
    raw = b"" if edge_state is None else bytes(edge_state)
    if len(raw) > 16:
        raise AssertionError("edge states of more than 16 bytes")
    for i, b in enumerate(raw):
        p.edge_state[i] = b
    return Transition(_abi.TR_MOVE_TO, "move_to_value", role_edges=[edgetype], params=p)


class UserBody:
    """A transition body loaded at run time (register_body): calling it with
    the edge / agent types of its roles and its parameter blob (a ctypes
    structure or bytes) gives the Transition to apply."""

    def __init__(self, id_, name, flags):
        self.id = id_
        self.name = name
        self.flags = flags

    def __call__(self, role_edges=(), role_agents=(), params=None):
        return Transition(self.id, self.name, role_agents=list(role_agents), role_edges=list(role_edges),
                          params=params)

    def __repr__(self):
        return f"UserBody({self.name}, id={self.id})"
