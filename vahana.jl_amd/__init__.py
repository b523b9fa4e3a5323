"""Host-side mirror of Vahana.jl's model/transition API over the MI355X backend.

Julia's mutating `f!` functions are spelled `f_` here (``add_agent_`` is
``add_agent!``, ``apply_`` is ``apply!``).  Agent and edge types are numpy
structured dtypes with the layout of the Julia bitstype; transition functions
are device bodies compiled into libvahana_hip.so and selected through
:mod:`.transitions` (an unregistered Python callable is rejected, there is no
CPU fallback).

Reference: src/ModelTypes.jl (registration), src/Simulation.jl (create_model,
create_simulation, finish_init!, apply!, apply, copy_simulation),
src/Agent.jl / src/AgentMethods.jl (agents, mapreduce), src/Edge.jl /
src/EdgeMethods.jl (edges and accessors).
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import S, I, R  # noqa: F401  (model state enums)
from . import transitions  # noqa: F401

__all__ = [
    "struct", "Edge", "ModelTypes", "register_agenttype_", "register_edgetype_",
    "create_model", "create_simulation", "add_agent_", "add_agents_", "add_edge_",
    "add_edges_", "finish_init_", "apply_", "apply", "mapreduce", "num_agents",
    "num_edges", "all_agents", "all_agentids", "agentstate", "edges", "neighborids",
    "edgestates", "has_edge", "copy_simulation", "finish_simulation_", "agent_id",
    "agent_nr", "type_nr", "register_transition", "enable_asserts", "VahanaError",
    "register_param_", "register_global_", "param", "set_param_", "get_global", "set_global_",
    "modify_global_", "push_global_", "param_ref", "global_ref", "get_options", "set_options",
    "options", "Batch", "register_body",
]
param_ref = transitions.param_ref
global_ref = transitions.global_ref


class VahanaError(RuntimeError):
    """A non-assertion error of the backend (device, memory, capacity...)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _raise(code, lib):
    msg = lib.last_error().decode(errors="replace")
    if code == _abi.VH_ERR_ASSERT:
        raise AssertionError(msg)
    raise VahanaError(code, msg)


def _check(code, lib):
    if code != _abi.VH_OK:
        _raise(code, lib)


# ------------------------------------------------------------------ types
class Value:
    """One value of a StructType (the type travels with it, like a Julia value)."""

    __slots__ = ("T", "v")

    def __init__(self, T, v):
        self.T = T
        self.v = v

    def __getitem__(self, k):
        return self.v[k]

    def __eq__(self, other):
        if isinstance(other, Value):
            return self.T is other.T and _state_eq(self.v, other.v)
        return _state_eq(self.v, other)

    def __repr__(self):
        return f"{self.T.name}{tuple(self.v) if self.v is not None else '()'}"


class StructType:
    """A Julia bitstype: `T(args...)` builds one value."""

    def __init__(self, name, fields=()):
        self.__name__ = name
        self.name = name
        self.fields = list(fields)
        self.dtype = np.dtype(self.fields, align=True) if self.fields else np.dtype([])
        self.nbytes = self.dtype.itemsize

    def __call__(self, *args, **kw):
        if kw:
            args = tuple(kw[f[0]] for f in self.fields)
        if not self.fields:
            return Value(self, None)
        return Value(self, np.array(tuple(args), dtype=self.dtype)[()])

    def pack(self, values):
        """Contiguous array of this type with zeroed padding bytes (numpy's
        structured conversions leave padding uninitialised)."""
        if isinstance(values, np.ndarray) and values.dtype.names:
            src = values
        else:
            src = np.array([v.v if isinstance(v, Value) else v for v in values], dtype=self.dtype)
        out = np.zeros(len(src), dtype=self.dtype)
        for name in self.dtype.names:
            out[name] = src[name]
        return out

    def to_bytes(self, values):
        if not self.fields:
            return b""
        return self.pack(values).tobytes()

    def from_bytes(self, buf, n):
        """Array over a private copy of the raw bytes (padding preserved:
        numpy's structured copies drop padding bytes)."""
        if not self.fields:
            return [None] * n
        return np.frombuffer(bytearray(buf[:n * self.nbytes]), dtype=self.dtype, count=n)

    def __repr__(self):
        return f"StructType({self.name})"


def struct(name, fields=()):
    """Declare an agent/edge type: ``struct("HK", [("opinion", "f8")])``."""
    return StructType(name, fields)


class Edge:
    """Edge{T}(from, state) src/Edge.jl:31-34."""

    __slots__ = ("from_", "state")

    def __init__(self, from_, state=None):
        self.from_ = int(from_)
        self.state = state

    def __eq__(self, other):
        return (isinstance(other, Edge) and self.from_ == other.from_ and
                _state_eq(self.state, other.state))

    def __repr__(self):
        return f"Edge({self.from_:#x}, {self.state})"


def _state_eq(a, b):
    if isinstance(a, Value):
        a = a.v
    if isinstance(b, Value):
        b = b.v
    if a is None or b is None:
        return a is None and b is None
    return bytes(np.asarray(a).tobytes()) == bytes(np.asarray(b).tobytes())


# ------------------------------------------------------------------ AgentID codec
def agent_id(type_nr_, nr):
    """src/Agent.jl:63-74 (rank 0)."""
    return (int(type_nr_) << _abi.VH_SHIFT_TYPE) | int(nr)


def agent_nr(id_):
    return int(id_) & _abi.VH_NR_MASK


def type_nr(id_):
    return int(id_) >> _abi.VH_SHIFT_TYPE


# ------------------------------------------------------------------ model types
class ModelTypes:
    """ModelTypes() src/ModelTypes.jl:1-74; compose with ``|``."""

    def __init__(self):
        self.agents = []       # (StructType, hints)
        self.edges = []        # (StructType, hints, target, size)
        self.params = []       # (name, default) of register_param!
        self.globals_ = []     # (name, initial value) of register_global!

    def __or__(self, step):
        return step(self)


def _hint_bits(hints, table, kind):
    bits = 0
    for h in hints:
        h = h.lstrip(":")
        if h not in table:
            raise AssertionError(f"The {kind} type hint {h} is unknown")
        bits |= table[h]
    return bits


def register_agenttype_(T, *hints):
    """register_agenttype! src/ModelTypes.jl:81-108 (pipeable)."""
    def step(types):
        assert all(t is not T for t, _ in types.agents), f"Type {T.name} is already registered"
        types.agents.append((T, _hint_bits(hints, _abi.AGENT_HINTS, "agent")))
        return types
    return step


def register_edgetype_(T, *hints, target=None, size=0):
    """register_edgetype! src/ModelTypes.jl:158-230 (pipeable)."""
    def step(types):
        assert all(t is not T for t, *_ in types.edges), f"Type {T.name} is already registered"
        types.edges.append((T, _hint_bits(hints, _abi.EDGE_HINTS, "edge"), target, size))
        return types
    return step


def register_param_(name, default):
    """register_param! src/ModelTypes.jl:260-266 (pipeable)."""
    def step(types):
        types.params.append((str(name).lstrip(":"), default))
        return types
    return step


def register_global_(name, init):
    """register_global! src/ModelTypes.jl:282-288 (pipeable)."""
    def step(types):
        types.globals_.append((str(name).lstrip(":"), init))
        return types
    return step


class Model:
    """create_model src/Simulation.jl:115-207: registered type ids and hints."""

    def __init__(self, types, name, lib):
        self.name = name
        self.lib = lib
        self.params = list(getattr(types, "params", ()))
        self.globals_ = list(getattr(types, "globals_", ()))
        self.handle = C.c_void_p()
        _check(lib.model_create(C.byref(self.handle)), lib)
        self.agent_types = []
        self.edge_types = []
        self.aid = {}
        self.eid = {}
        for T, hints in types.agents:
            tid = C.c_uint8()
            _check(lib.register_agenttype(self.handle, hints, T.nbytes, C.byref(tid)), lib)
            self.aid[T] = tid.value
            self.agent_types.append((T, hints))
        self.etarget = {}      # edge type -> target agent type (:SingleType), or None
        for T, hints, target, size in types.edges:
            self.etarget[T] = target
            eid = C.c_uint8()
            tgt = self.aid[target] if target is not None else -1
            _check(lib.register_edgetype(self.handle, hints, T.nbytes, tgt, int(size),
                                         C.byref(eid)), lib)
            self.eid[T] = eid.value
            self.edge_types.append(T)
        self.ehints = {}
        for T, e in self.eid.items():
            h = C.c_uint32()
            _check(lib.edgetype_hints(self.handle, e, C.byref(h)), lib)
            self.ehints[T] = h.value
        self.immortal = {T: bool(h & _abi.HINT_IMMORTAL) for T, h in self.agent_types}

    def has_hint(self, T, hint):
        bit = _abi.EDGE_HINTS[hint] if T in self.eid else _abi.AGENT_HINTS[hint]
        h = self.ehints[T] if T in self.eid else dict(self.agent_types)[T]
        return bool(h & bit)

    def __del__(self):
        try:
            if self.handle:
                self.lib.model_destroy(self.handle)
        except Exception:
            pass


def create_model(name):
    """create_model(types, name) (pipeable)."""
    def step(types):
        return Model(types, name, _abi.lib())
    return step


# ------------------------------------------------------------------ transitions
_REGISTRY = {}


def register_transition(func, transition):
    """Bind a Python function object to a device body (the Julia side keeps a
    Dict{Function,UInt32} of its model functions)."""
    _REGISTRY[func] = transition
    return func


def _resolve(func):
    if isinstance(func, transitions.Transition):
        return func
    if func in _REGISTRY:
        return _REGISTRY[func]
    raise VahanaError(_abi.VH_ERR_UNREGISTERED,
                      f"{getattr(func, '__name__', func)!r} is not a registered device "
                      "transition (no CPU fallback exists)")


def _as_list(x):
    if x is None:
        return []
    if isinstance(x, (list, tuple, set)):
        return list(x)
    return [x]


# ------------------------------------------------------------------ simulation
class Record:
    """A mutable struct of named fields: the Params_/Globals_ structs that
    create_model generates from register_param!/register_global!
    (src/Simulation.jl:180-200)."""

    def __init__(self, fields):
        self.__dict__.update(fields)

    def __repr__(self):
        return "Record(" + ", ".join(f"{k}={v!r}" for k, v in vars(self).items()) + ")"


def _fields_of(obj):
    if obj is None:
        return {}
    if isinstance(obj, dict):
        return dict(obj)
    if hasattr(obj, "_asdict"):
        return dict(obj._asdict())
    return dict(vars(obj))


def _getf(obj, name):
    return obj[name] if isinstance(obj, dict) else getattr(obj, name)


def _hasf(obj, name):
    return name in obj if isinstance(obj, dict) else hasattr(obj, name)


def _setf(obj, name, value):
    if isinstance(obj, dict):
        obj[name] = value
    else:
        setattr(obj, name, value)


def _scalar_dtype(v):
    """(vh_dtype, numpy dtype) of a scalar value, None otherwise."""
    if isinstance(v, (bool, np.bool_)):
        return _abi.DT["b1"], np.uint8
    if isinstance(v, np.unsignedinteger):
        return _abi.DT["u8"], np.uint64
    if isinstance(v, (int, np.integer)):
        return _abi.DT["i8"], np.int64
    if isinstance(v, np.float32):
        return _abi.DT["f4"], np.float32
    if isinstance(v, (float, np.floating)):
        return _abi.DT["f8"], np.float64
    return None


class Simulation:
    """create_simulation src/Simulation.jl:261-339 on the device backend."""

    def __init__(self, model, name=None, handle=None, params=None, globals_=None):
        self.model = model
        self.lib = model.lib
        self.name = name or model.name
        self.handle = C.c_void_p()
        if handle is not None:
            self.handle = handle
        else:
            _check(self.lib.sim_create(model.handle, C.byref(self.handle)), self.lib)
        self.initialized = False
        self._init_values(params, globals_)

    # -- params / globals (src/Simulation.jl:270-292, src/Global.jl)
    def _init_values(self, params, globals_):
        import copy
        m = self.model
        if getattr(m, "params", None):
            assert params is None, ("You can not use `register_param!` and the `param` argument of "
                                    "`create_simulation` at the same time.")
            params = Record(copy.deepcopy(dict(m.params)))
        if getattr(m, "globals_", None):
            assert globals_ is None, ("You can not use `register_globals!` and the `global` argument "
                                      "of `create_simulation` at the same time.")
            globals_ = Record(copy.deepcopy(dict(m.globals_)))
        self.params, self.globals = params, globals_
        self.globals_last_change = 0
        # scalar fields get numbered value slots in the library (the values a
        # transition body can read); others stay on the host
        self._slots = ({}, {})
        for kind, obj in ((_abi.VALUE_PARAM, params), (_abi.VALUE_GLOBAL, globals_)):
            for k, v in _fields_of(obj).items():
                if _scalar_dtype(v) is not None and len(self._slots[kind]) < _abi.MAX_VALUES:
                    self._slots[kind][k] = len(self._slots[kind])
                    self._push_value(kind, k, v)

    def _push_value(self, kind, name, v):
        slot = self._slots[kind].get(name)
        if slot is None:
            return
        dt = _scalar_dtype(v)
        if dt is None:
            raise TypeError(f"{name} holds a scalar; {type(v).__name__} can not be stored in it")
        buf = np.array([v], dtype=dt[1])
        _check(self.lib.sim_set_value(self.handle, kind, slot, dt[0], buf.ctypes.data_as(C.c_void_p)),
               self.lib)

    def _value(self, kind, name):
        """The library's copy of a scalar param/global (what bodies read)."""
        slot = self._slots[kind][name]
        dt = C.c_int()
        raw = (C.c_uint8 * 8)()
        _check(self.lib.sim_get_value(self.handle, kind, slot, C.byref(dt), raw), self.lib)
        code = {v: k for k, v in _abi.DT.items()}[dt.value]
        npt = {"b1": np.uint8, "u8": np.uint64, "i8": np.int64, "f4": np.float32, "f8": np.float64,
               "u4": np.uint32, "i4": np.int32, "u2": np.uint16, "u1": np.uint8}[code]
        v = np.frombuffer(bytes(raw), dtype=npt, count=1)[0]
        return bool(v) if code == "b1" else v.item()

    # -- type ids
    def _aid(self, T):
        if T not in self.model.aid:
            raise VahanaError(_abi.VH_ERR_UNREGISTERED, f"{T} is not a registered agent type")
        return self.model.aid[T]

    def _eid(self, T):
        if T not in self.model.eid:
            raise VahanaError(_abi.VH_ERR_UNREGISTERED, f"{T} is not a registered edge type")
        return self.model.eid[T]

    # -- init
    def _add_agents(self, T, values):
        vals = list(values) if not isinstance(values, np.ndarray) else values
        for v in vals if not isinstance(vals, np.ndarray) else ():
            if isinstance(v, Value) and v.T is not T:
                raise AssertionError(f"{v.T.name} value added as {T.name}")
        n = len(vals)
        buf = T.to_bytes(vals) if T.fields else b""
        ids = (C.c_uint64 * max(n, 1))()
        cbuf = C.create_string_buffer(buf, max(len(buf), 1))
        _check(self.lib.add_agents(self.handle, self._aid(T), cbuf, n, ids), self.lib)
        return [ids[i] for i in range(n)]

    def _add_agents_array(self, T, arr):
        arr = T.pack(arr) if T.fields else np.zeros(len(arr), dtype=np.uint8)
        n = len(arr)
        ids = np.empty(n, dtype=np.uint64)
        _check(self.lib.add_agents(self.handle, self._aid(T), arr.ctypes.data_as(C.c_void_p), n,
                                   ids.ctypes.data_as(C.POINTER(C.c_uint64))), self.lib)
        return ids

    def _add_edges_array(self, T, from_, to, states=None):
        from_ = np.ascontiguousarray(from_, dtype=np.uint64)
        to = np.ascontiguousarray(to, dtype=np.uint64)
        n = len(to)
        sp = None
        if T.fields and states is not None:
            states = T.pack(states)
            sp = states.ctypes.data_as(C.c_void_p)
        _check(self.lib.add_edges(self.handle, self._eid(T),
                                  from_.ctypes.data_as(C.POINTER(C.c_uint64)),
                                  to.ctypes.data_as(C.POINTER(C.c_uint64)), sp, n), self.lib)

    def _finish_init(self):
        _check(self.lib.finish_init(self.handle), self.lib)
        self.initialized = True

    def _finish_init_relabel(self, T):
        """finish_init! with the agents of type T renumbered breadth-first
        (vh_finish_init_relabel); returns new_nr[old_nr - 1]."""
        n = C.c_uint64()
        self.lib.read_agents(self.handle, self._aid(T), None, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(self.lib.finish_init_relabel(self.handle, self._aid(T),
                                            out.ctypes.data_as(C.POINTER(C.c_uint64))), self.lib)
        self.initialized = True
        return out[:n.value]

    # -- apply!
    def _apply(self, func, call, read, write, add_existing=(), with_edge=None):
        d, keep = self._desc(func, call, read, write, add_existing, with_edge)
        self._last_desc = (d, keep)
        _check(self._call_apply(d), self.lib)
        return self

    def _prepare(self, func, call, read, write, add_existing=(), with_edge=None):
        """The apply! descriptor of one transition, built once for repeated
        steps with `_apply_prepared` (same meaning as `_apply`)."""
        return self._desc(func, call, read, write, add_existing, with_edge)

    def _apply_prepared(self, prepared, steps=1):
        """`steps` apply! calls of a prepared transition, enqueued in one ABI
        call (vh_apply_n); asynchronous like apply!."""
        d, _keep = prepared
        _check(self.lib.apply_n(self.handle, C.byref(d), steps), self.lib)
        return self

    def _opinion_prepare(self, T, E, f32=False):
        """Build what the first opinion apply! of agent type T over edge type
        E needs (degree bins, the two-pass gather plan) now, on the
        simulation's stream (vh_opinion_prepare)."""
        _check(self.lib.opinion_prepare(self.handle, self._aid(T), self._eid(E), 1 if f32 else 0), self.lib)

    def _set_streams(self, stream, io_stream=None):
        """Run apply! on `stream` (a raw hipStream_t, None: own stream) and
        reads on `io_stream` (vh_set_io_stream)."""
        _check(self.lib.set_stream(self.handle, stream), self.lib)
        _check(self.lib.set_io_stream(self.handle, io_stream), self.lib)

    def _add_edges_gnp(self, T, mean_deg, graph_seed, weight_seed=0):
        """Directed G(n, p) in-edges of edge type T generated on the device at
        finish_init! (vh_add_edges_gnp); the host twin is inputs.graph_gnp with
        weights inputs.gnp_weights."""
        _check(self.lib.add_edges_gnp(self.handle, self._eid(T), float(mean_deg), int(graph_seed),
                                      int(weight_seed)), self.lib)

    def _add_edges_ba(self, T, m, graph_seed, weight_seed=0):
        """The symmetrised Barabasi-Albert graph of edge type T generated on
        the device at finish_init! (vh_add_edges_ba); the host twin is
        inputs.symmetric_weighted(*inputs.graph_ba(n, m, graph_seed),
        weight_seed)."""
        _check(self.lib.add_edges_ba(self.handle, self._eid(T), int(m), int(graph_seed), int(weight_seed)),
               self.lib)

    def _desc(self, func, call, read, write, add_existing=(), with_edge=None):
        tr = _resolve(func)
        call, read, write = _as_list(call), _as_list(read), _as_list(write)
        add_existing = _as_list(add_existing)
        m = self.model
        d = _abi.ApplyDesc()
        d.transition = tr.id
        d.n_call = len(call)
        for i, T in enumerate(call):
            d.call[i] = self._aid(T)
        for lst, an, en in ((read, "read_agents", "read_edges"),
                            (write, "write_agents", "write_edges"),
                            (add_existing, "add_existing_agents", "add_existing_edges")):
            am = em = 0
            for T in lst:
                if T in m.aid:
                    am |= 1 << m.aid[T]
                elif T in m.eid:
                    em |= 1 << m.eid[T]
                else:
                    raise VahanaError(_abi.VH_ERR_UNREGISTERED, f"{T} is not registered")
            setattr(d, an, am)
            setattr(d, en, em)
        d.with_edge = self._eid(with_edge) if with_edge is not None else -1
        for i in range(4):
            d.role_agent[i] = -1
            d.role_edge[i] = -1
        for i, T in enumerate(tr.role_agents):
            d.role_agent[i] = self._aid(T) if T is not None else -1
        for i, T in enumerate(tr.role_edges):
            d.role_edge[i] = self._eid(T) if T is not None else -1
        pbuf = tr.params_bytes()
        keep = C.create_string_buffer(pbuf, max(len(pbuf), 1))
        d.params = C.cast(keep, C.c_void_p)
        d.params_bytes = len(pbuf)
        binds = getattr(tr, "binds", ())
        barr = None
        if binds:
            barr = (_abi.ValueBind * len(binds))()
            for i, (off, dt, ref) in enumerate(binds):
                slot = self._slots[ref.kind].get(ref.name)
                if slot is None:
                    raise AssertionError(f"{ref!r}: the simulation has no scalar "
                                         f"{'parameter' if ref.kind == _abi.VALUE_PARAM else 'global'} "
                                         f"named {ref.name}")
                barr[i].offset, barr[i].kind, barr[i].dtype, barr[i].slot = off, ref.kind, dt, slot
            d.n_bind = len(binds)
            d.bind = C.cast(barr, C.POINTER(_abi.ValueBind))
        return d, (keep, barr)

    def _call_apply(self, d):
        return self.lib.apply(self.handle, C.byref(d))

    # -- queries
    def _num_agents(self, T):
        n = C.c_uint64()
        _check(self.lib.num_agents(self.handle, self._aid(T), C.byref(n)), self.lib)
        return n.value

    def _num_edges(self, T):
        n = C.c_uint64()
        _check(self.lib.num_edges(self.handle, self._eid(T), C.byref(n)), self.lib)
        return n.value

    def _read_agents(self, T):
        """(states array, died bool array) of all slots (raw view)."""
        n = C.c_uint64()
        cap = 0
        rc = self.lib.read_agents(self.handle, self._aid(T), None, None, 0, C.byref(n))
        if rc not in (_abi.VH_OK, _abi.VH_ERR_CAPACITY):
            _raise(rc, self.lib)
        cap = n.value
        buf = C.create_string_buffer(max(cap * T.nbytes, 1))
        died = (C.c_uint8 * max(cap, 1))()
        _check(self.lib.read_agents(self.handle, self._aid(T), buf, died, cap, C.byref(n)), self.lib)
        states = T.from_bytes(buf.raw[:cap * T.nbytes], cap) if T.fields else [T()] * cap
        return states, np.frombuffer(bytes(died), dtype=np.uint8, count=cap).astype(bool)

    def _read_edges(self, T):
        """(to, from, states) of all edges in container order."""
        n = C.c_uint64()
        rc = self.lib.read_edges(self.handle, self._eid(T), None, None, None, 0, C.byref(n))
        if rc not in (_abi.VH_OK, _abi.VH_ERR_CAPACITY):
            _raise(rc, self.lib)
        cap = n.value
        to = np.zeros(max(cap, 1), dtype=np.uint64)
        fr = np.zeros(max(cap, 1), dtype=np.uint64)
        st = C.create_string_buffer(max(cap * T.nbytes, 1))
        _check(self.lib.read_edges(self.handle, self._eid(T),
                                   to.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   fr.ctypes.data_as(C.POINTER(C.c_uint64)), st, cap,
                                   C.byref(n)), self.lib)
        states = T.from_bytes(st.raw[:cap * T.nbytes], cap) if T.fields else None
        return to[:cap], fr[:cap], states

    def _edges_of(self, id_, T):
        n = C.c_uint64()
        rc = self.lib.edges_of(self.handle, self._eid(T), int(id_), None, None, 0, C.byref(n))
        if rc not in (_abi.VH_OK, _abi.VH_ERR_CAPACITY):
            _raise(rc, self.lib)
        cap = n.value
        if cap == 0:
            return [], []
        fr = (C.c_uint64 * cap)()
        st = C.create_string_buffer(max(cap * T.nbytes, 1))
        _check(self.lib.edges_of(self.handle, self._eid(T), int(id_), fr, st, cap, C.byref(n)),
               self.lib)
        states = list(T.from_bytes(st.raw[:cap * T.nbytes], cap)) if T.fields else [None] * cap
        return [fr[i] for i in range(cap)], states

    def _agentstate(self, id_, T):
        buf = C.create_string_buffer(max(T.nbytes, 1))
        _check(self.lib.agentstate(self.handle, int(id_), self._aid(T), buf), self.lib)
        return T.from_bytes(buf.raw[:T.nbytes], 1)[0] if T.fields else T()

    def _mapreduce(self, kind, tid, md, op, dt, initp, outp):
        return self.lib.mapreduce(self.handle, kind, tid, C.byref(md), op, dt, initp, outp)

    def _mapreduce_expr(self, kind, tid, prog, nops, op, dt, initp, outp):
        return self.lib.mapreduce_expr(self.handle, kind, tid, prog, nops, op, dt, initp, outp)

    def _num_transitions(self):
        n = C.c_uint64()
        _check(self.lib.num_transitions(self.handle, C.byref(n)), self.lib)
        return n.value

    def _copy(self):
        h = C.c_void_p()
        _check(self.lib.sim_copy(self.handle, C.byref(h)), self.lib)
        c = type(self).__new__(type(self))
        c.model, c.lib, c.name, c.handle = self.model, self.lib, self.name, h
        c.initialized = self.initialized
        import copy
        c.params, c.globals = copy.deepcopy(self.params), copy.deepcopy(self.globals)
        c.globals_last_change = self.globals_last_change
        c._slots = (dict(self._slots[0]), dict(self._slots[1]))
        return c

    def _destroy(self):
        if self.handle:
            self.lib.sim_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass


def create_simulation(model, params=None, globals_=None, name=None):
    """create_simulation src/Simulation.jl:261-339. `params` / `globals_`: an
    object with fields (or a dict), or None when the model registered them
    (register_param_ / register_global_); scalar fields are what transition
    bodies can read (transitions.param_ref / global_ref)."""
    return Simulation(model, name, params=params, globals_=globals_)


def param(sim, name):
    """param(sim, name) src/Simulation.jl:587."""
    return _getf(sim.params, str(name).lstrip(":"))


def set_param_(*args):
    """set_param!(sim, name, value) src/Simulation.jl:602-625, or its pipeable
    form set_param_(name, value)(sim). Only before finish_init!; an unknown
    name is a warning."""
    if len(args) == 2:
        name, value = args
        return lambda sim: (set_param_(sim, name, value), sim)[1]
    sim, name, value = args
    name = str(name).lstrip(":")
    assert not sim.initialized, ("You can call `set_param!` only before `finish_init!`. For values "
                                 "that change while the simulation is running, register them as "
                                 "globals and use `set_global!` instead.")
    if sim.params is None or not _hasf(sim.params, name):
        import warnings
        warnings.warn(f"Parameter {name} is unknown")
        return
    _setf(sim.params, name, value)
    sim._push_value(_abi.VALUE_PARAM, name, value)


def get_global(sim, name):
    """get_global(sim, name) src/Global.jl:10."""
    return _getf(sim.globals, str(name).lstrip(":"))


def set_global_(sim, name, value):
    """set_global!(sim, name, value) src/Global.jl:26-29 (not inside a
    transition: here globals change only between apply! calls)."""
    name = str(name).lstrip(":")
    sim.globals_last_change = sim._num_transitions() - 1
    _setf(sim.globals, name, value)
    sim._push_value(_abi.VALUE_GLOBAL, name, value)


def modify_global_(sim, name, f):
    """modify_global!(sim, name, f) src/Global.jl:53-55."""
    set_global_(sim, name, f(get_global(sim, name)))


def push_global_(sim, name, value):
    """push_global!(sim, name, value) src/Global.jl:86-89: appends to a
    vector-valued global (a time series)."""
    name = str(name).lstrip(":")
    sim.globals_last_change = sim._num_transitions() - 1
    _getf(sim.globals, name).append(value)


def add_agent_(sim, agent):
    """add_agent! src/AgentMethods.jl:65-89 (initialisation phase)."""
    T = _type_of_value(sim, agent)
    return sim._add_agents(T, [agent])[0]


def add_agents_(sim, *agents):
    """add_agents! src/Agent.jl:174-180."""
    if len(agents) == 1 and isinstance(agents[0], (list, tuple)):
        agents = agents[0]
    if not agents:
        return []
    T = _type_of_value(sim, agents[0])
    return sim._add_agents(T, agents)


def _type_of_value(sim, v):
    if isinstance(v, Value):
        return v.T
    if isinstance(v, tuple) and len(v) == 2 and isinstance(v[0], StructType):
        return v[0]
    raise AssertionError("agent values are built with their type: T(fields...)")


def add_edge_(sim, *args):
    """add_edge!(sim, from, to, state) / add_edge!(sim, to, Edge(from, state))
    src/EdgeMethods.jl:388-523, src/Edge.jl:36-66 (initialisation phase).
    `state` is a value T(...) of the edge type (or a (T, value) pair)."""
    if len(args) == 2:
        to, e = args
        from_, state = e.from_, e.state
    else:
        from_, to, state = args
    if isinstance(state, Value):
        T, st = state.T, state.v
    elif isinstance(state, tuple) and len(state) == 2 and isinstance(state[0], StructType):
        T, st = state
        if isinstance(st, Value):
            st = st.v
    else:
        raise AssertionError("edge states are built with their type: T(fields...)")
    sim._add_edges_array(T, [from_], [to], [st] if T.fields else None)


def add_edges_(sim, to, edges_):
    """add_edges!(sim, to, edges) src/Edge.jl:81-97."""
    for e in edges_:
        add_edge_(sim, to, e)


def finish_init_(sim, relabel=None):
    """finish_init! src/Simulation.jl:403-476. With `relabel` = an agent type,
    its agents are renumbered breadth-first over their edges (as the
    reference renumbers agents when it distributes them, returning the
    mapping with `return_idmapping`); returns new_nr[old_nr - 1]."""
    if relabel is None:
        sim._finish_init()
        return None
    return sim._finish_init_relabel(relabel)


def updateids(idmapping, ids):
    """updateids for a mapping from finish_init_ (the reference's idmapping,
    src/Simulation.jl:455-475): the new
    AgentIDs of `ids` (scalar or array) of the renumbered type."""
    a = np.asarray(ids, dtype=np.uint64)
    t = a & ~np.uint64(_abi.VH_NR_MASK)
    nr = (a & np.uint64(_abi.VH_NR_MASK)).astype(np.int64)
    out = t | np.asarray(idmapping, dtype=np.uint64)[nr - 1]
    return int(out) if out.ndim == 0 else out


def apply_(sim, func, call, read, write, add_existing=(), with_edge=None):
    """apply! src/Simulation.jl:720-821."""
    return sim._apply(func, call, read, write, add_existing, with_edge)


def apply(sim, func, call, read, write, add_existing=(), with_edge=None):
    """apply src/Simulation.jl:847-856: apply! on a copy, returns the copy."""
    c = copy_simulation(sim)
    c._apply(func, call, read, write, add_existing, with_edge)
    return c


def copy_simulation(sim):
    """copy_simulation src/Simulation.jl:500-510."""
    return sim._copy()


def finish_simulation_(sim):
    """finish_simulation! src/Simulation.jl:551-585."""
    sim._destroy()


# ------------------------------------------------------------------ mapreduce
class Field:
    """f = x -> x.field (optionally x.field == value)."""

    def __init__(self, name, eq=None):
        self.name = name
        self.eq = eq


def _map_desc(T, f):
    md = _abi.MapDesc()
    if f is None or f == 1:
        md.kind = _abi.MAP_ONE
        return md, None
    if isinstance(f, str):
        f = Field(f)
    fdt = T.dtype.fields[f.name][0]
    md.offset = T.dtype.fields[f.name][1]
    code = {"i8": "i8", "f8": "f8", "b1": "b1", "f4": "f4", "u1": "u1", "u2": "u2", "u4": "u4",
            "i4": "i4", "u8": "u8", "?": "b1"}[fdt.kind + str(fdt.itemsize) if fdt.kind != "b" else "b1"]
    md.field_dtype = _abi.DT[code]
    if f.eq is not None:
        md.kind = _abi.MAP_FIELD_EQ
        md.value = int(f.eq)
    else:
        md.kind = _abi.MAP_FIELD
    return md, code


def mapreduce(sim, f, op, T, init=None, datatype=None):
    """mapreduce(sim, f, op, T; init, datatype) src/AgentMethods.jl:533-565 and
    src/EdgeMethods.jl:972-1001.  f: a field name, Field(name[, eq]) or None for 1;
    op: '+', '*', 'max', 'min', '&', '|'; datatype: 'i8' (Int), 'f8', 'b1' (Bool)."""
    kind = 0 if T in sim.model.aid else 1
    tid = sim._aid(T) if kind == 0 else sim._eid(T)
    if isinstance(f, expr.Expr):
        return _mapreduce_expr(sim, f, op, T, kind, tid, init, datatype)
    md, fcode = _map_desc(T, f)
    if datatype is None:
        if init is not None:
            datatype = "f8" if isinstance(init, float) else ("b1" if isinstance(init, bool) else "i8")
        elif op in ("&", "|"):
            datatype = "b1"
        elif fcode in ("f8", "f4") and (f is not None and getattr(f, "eq", None) is None):
            datatype = "f8"
        else:
            datatype = "i8"
    dt = _abi.DT[datatype]
    out = C.c_uint64()
    initp = None
    if init is not None:
        if datatype == "f8":
            iv = C.c_double(float(init))
        else:
            iv = C.c_int64(int(init))
        initp = C.cast(C.pointer(iv), C.c_void_p)
    rc = sim._mapreduce(kind, tid, md, _abi.OP[op], dt, initp, C.cast(C.pointer(out), C.c_void_p))
    _check(rc, sim.lib)
    raw = out.value
    if datatype == "f8":
        return float(np.array([raw], dtype=np.uint64).view(np.float64)[0])
    if datatype == "b1":
        return bool(raw)
    return int(np.array([raw], dtype=np.uint64).view(np.int64)[0])


def _init_arg(init, datatype):
    if init is None:
        return None, None
    iv = C.c_double(float(init)) if datatype == "f8" else C.c_int64(int(init))
    return iv, C.cast(C.pointer(iv), C.c_void_p)


def _result(raw, datatype):
    if datatype == "f8":
        return float(np.array([raw], dtype=np.uint64).view(np.float64)[0])
    if datatype == "b1":
        return bool(raw)
    return int(np.array([raw], dtype=np.uint64).view(np.int64)[0])


def _mapreduce_expr(sim, f, op, T, kind, tid, init, datatype):
    """mapreduce with a computed f (vahana_amd.expr): vh_mapreduce_expr."""
    prog, nops, vkind = expr.compile_expr(f, T, sim)
    if datatype is None:
        if init is not None:
            datatype = "f8" if isinstance(init, float) else ("b1" if isinstance(init, bool) else "i8")
        elif vkind == "f":
            datatype = "f8"
        elif vkind == "b":
            datatype = "i8" if op in ("+", "*") else "b1"   # Bool + Bool is an Int
        else:
            datatype = "i8"
    out = C.c_uint64()
    _iv, initp = _init_arg(init, datatype)
    rc = sim._mapreduce_expr(kind, tid, prog, nops, _abi.OP[op], _abi.DT[datatype], initp,
                             C.cast(C.pointer(out), C.c_void_p))
    _check(rc, sim.lib)
    return _result(out.value, datatype)


def register_body(code_object, kernel, adds_edges=False, adds_agents=False):
    """A transition body compiled against include/vahana_device.h into a code
    object (hipcc --genco --offload-arch=gfx950), loaded at run time
    (vh_register_body): apply! then takes any such body, as the reference's
    takes any function (src/Simulation.jl:720-730). `adds_edges`: the body
    calls add_edge! / remove_edges! (edge role 0); `adds_agents`: add_agent!
    (agent role 0). Returns a transitions.UserBody."""
    lib = _abi.lib()
    flags = (_abi.BODY_ADDS_EDGES if adds_edges else 0) | (_abi.BODY_ADDS_AGENTS if adds_agents else 0)
    tid = C.c_uint32()
    _check(lib.register_body(os.fsencode(os.path.abspath(code_object)), kernel.encode(), flags,
                             C.byref(tid)), lib)
    return transitions.UserBody(tid.value, kernel, flags)


def example_bodies():
    """Path of the code object built from examples/bodies.hip."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "examples", "bodies.co")


# ------------------------------------------------------------------ queries
def num_agents(sim, T):
    """num_agents src/Agent.jl:324-343."""
    return sim._num_agents(T)


def num_edges(sim, a, b=None):
    """num_edges(sim, T) src/Edge.jl:373 or num_edges(sim, id, T) src/EdgeMethods.jl:850-870."""
    if b is None:
        return sim._num_edges(a)
    T = b
    assert not (sim.model.has_hint(T, "SingleEdge")), \
        "num_edges is not defined for the hint combination"
    fr, _ = sim._edges_of(a, T)
    return len(fr)


def all_agents(sim, T):
    """all_agents src/Agent.jl:234-260: live states in slot order."""
    states, died = sim._read_agents(T)
    return [states[i] for i in range(len(died)) if not died[i]]


def all_agentids(sim, T):
    """all_agentids src/Agent.jl:287-313."""
    _, died = sim._read_agents(T)
    t = sim._aid(T)
    return [agent_id(t, i + 1) for i in range(len(died)) if not died[i]]


def agentstate(sim, id_, T):
    """agentstate outside a transition src/AgentMethods.jl:91-123."""
    return sim._agentstate(id_, T)


def _need(sim, T, ok, what):
    assert ok, f"{what} is not defined for the hint combination of {T.name}"


def edges(sim, id_, T):
    """edges src/EdgeMethods.jl:704-716: list of Edge or None (SingleEdge: one Edge)."""
    m = sim.model
    _need(sim, T, not m.has_hint(T, "Stateless") and not m.has_hint(T, "IgnoreFrom"), "edges")
    fr, st = sim._edges_of(id_, T)
    if not fr:
        return None
    es = [Edge(f, s) for f, s in zip(fr, st)]
    return es[-1] if m.has_hint(T, "SingleEdge") else es


def neighborids(sim, id_, T):
    """neighborids src/EdgeMethods.jl:718-743."""
    m = sim.model
    _need(sim, T, not m.has_hint(T, "IgnoreFrom"), "neighborids")
    fr, _ = sim._edges_of(id_, T)
    if not fr:
        return None
    return fr[-1] if m.has_hint(T, "SingleEdge") else fr


def edgestates(sim, id_, T):
    """edgestates src/EdgeMethods.jl:800-826."""
    m = sim.model
    _need(sim, T, not m.has_hint(T, "Stateless"), "edgestates")
    fr, st = sim._edges_of(id_, T)
    if not fr:
        return None
    return st[-1] if m.has_hint(T, "SingleEdge") else st


def has_edge(sim, id_, T):
    """has_edge src/EdgeMethods.jl:873-892."""
    m = sim.model
    _need(sim, T, not (m.has_hint(T, "SingleEdge") and m.has_hint(T, "SingleType") and
                       not (m.has_hint(T, "Stateless") and m.has_hint(T, "IgnoreFrom"))),
          "has_edge")
    fr, _ = sim._edges_of(id_, T)
    return len(fr) > 0


class Batch:
    """apply! of one transition to many simulations at once (vh_batch_apply):
    the one-pass opinion bodies of all of them run as one launch per step.
    The simulations must run on one stream and come from models with the
    same type ids (e.g. built by the same builder)."""

    def __init__(self):
        self.lib = _abi.lib()
        self.handle = C.c_void_p()
        _check(self.lib.batch_create(C.byref(self.handle)), self.lib)

    def apply(self, sims, func, call, read, write, steps=1, add_existing=(), with_edge=None):
        """`steps`: one count for all, or one per simulation (simulation i
        takes part in the first steps[i] steps)."""
        n = len(sims)
        if not n:
            return
        steps = [int(steps)] * n if np.isscalar(steps) else [int(k) for k in steps]
        d, keep = sims[0]._desc(func, call, read, write, add_existing, with_edge)
        hs = (C.c_void_p * n)(*[s.handle.value for s in sims])
        st = (C.c_uint32 * n)(*steps)
        _check(self.lib.batch_apply(self.handle, hs, n, C.byref(d), st), self.lib)
        del keep
        for s in sims:
            s.initialized = True

    def close(self):
        if self.handle:
            self.lib.batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_options():
    """The kernel options (vh_get_options) as a dict."""
    o = _abi.Options()
    L = _abi.lib()
    _check(L.get_options(C.byref(o)), L)
    return {k: getattr(o, k) for k, _ in _abi.Options._fields_ if k != "pad"}


def set_options(**kw):
    """Set kernel options (vh_set_options): opinion_plan_bytes,
    opinion_part_bytes, sir_expose, sir_expose_min. Returns the previous
    values. Results are bit-identical under every setting."""
    prev = get_options()
    cur = dict(prev)
    for k, v in kw.items():
        if k not in cur:
            raise KeyError(f"unknown option {k}")
        cur[k] = int(v)
    o = _abi.Options(**cur)
    L = _abi.lib()
    _check(L.set_options(C.byref(o)), L)
    return prev


class options:
    """Context manager: ``with options(sir_expose=0): ...`` restores the
    previous options on exit."""

    def __init__(self, **kw):
        self.kw = kw
        self.prev = None

    def __enter__(self):
        self.prev = set_options(**self.kw)
        return self

    def __exit__(self, *exc):
        set_options(**self.prev)
        return False


def enable_asserts(enable=True):
    """enable_asserts src/Vahana.jl:42-48."""
    _abi.lib().enable_asserts(1 if enable else 0)


from . import inputs, models, ensemble, serve, raster, snapshot, expr  # noqa: E402,F401
