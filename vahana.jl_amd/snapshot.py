"""Checkpoint / resume of a simulation (src/HDF5.jl).

A snapshot directory holds, per agent type, per edge type and for the
globals, one record for each snapshot at which that part had changed since
its previous record (the last_change test of write_agents / write_edges /
write_globals, src/HDF5.jl:483-541, 577-673, 361-411). Records are filed
under k = num_transitions - 1 (transition_str, src/HDF5.jl:119), so after
apply! number k they carry the k that apply! stamped into last_change:

    index.json               model name, snapshots [k, comment, ignored],
                             and per part the records [k, last_change]
    params.npz               the parameters (written by the first snapshot)
    globals/t_<k>.npz        the globals
    agents/<T>/t_<k>.npz     slot states, died flags (not for :Immortal
                             types), slot count
    edges/<E>/t_<k>.npz      one record kind chosen by the hints
                             (src/HDF5.jl:543-575):
        complete       to, from, state   the default (CompleteEdge)
        stateless      to, from          :Stateless or no fields (StatelessEdge)
        ignore_from    to, state         :IgnoreFrom with a state (IgnoreFromEdge)
        edge_count     to, count         neighbours only: (:Stateless or no
                                         fields) and :IgnoreFrom (EdgeCount)
        count_vector   count per target  the same with :SingleType: the count
                       slot              of target slot i at i - 1, Bool for
                                         :SingleEdge (the raw field vector)
    rasters.npz              the rasters' cell ids (immutable, written once)

Reading snapshot k takes, per part, its latest record <= k
(find_transition_nr, src/HDF5.jl:781-791), sets that part's last_change to
the record's number (src/HDF5.jl:1265-1267, 1434-1436, 1615-1618), restores
the slots with their died flags and rebuilds the free slots in ascending
order (_read_agents_restore!, src/HDF5.jl:801-850), re-adds the edges in
record order (every container keeps its order) and sets num_transitions to
k + 1, the value it had when the snapshot was written.

The files are NumPy archives and JSON, not HDF5 (no HDF5 library is
available to this build); they restore into the device library and into
the checker alike.
"""
import json
import os
import shutil

import numpy as np

from . import _abi, _check

_INDEX = "index.json"


# ------------------------------------------------------------------ helpers
def _tdir(path, kind, name):
    return os.path.join(path, kind, name)


def _rec(path, kind, name, k):
    return os.path.join(path, kind, name, f"t_{int(k)}.npz") if name else \
        os.path.join(path, kind, f"t_{int(k)}.npz")


def _load_index(path):
    f = os.path.join(path, _INDEX)
    if not os.path.isfile(f):
        return None
    with open(f) as fh:
        return json.load(fh)


def _save_index(path, idx):
    tmp = os.path.join(path, _INDEX + ".tmp")
    with open(tmp, "w") as fh:
        json.dump(idx, fh)
    os.replace(tmp, os.path.join(path, _INDEX))


def _dtype_meta(dt):
    return {"names": list(dt.names or ()), "formats": [dt.fields[n][0].str for n in dt.names or ()],
            "offsets": [int(dt.fields[n][1]) for n in dt.names or ()], "itemsize": int(dt.itemsize)}


def _dtype_of(meta):
    if not meta["names"]:
        return None
    return np.dtype({"names": meta["names"], "formats": meta["formats"], "offsets": meta["offsets"],
                     "itemsize": meta["itemsize"]})


def _raw(st, T, n):
    return (np.ascontiguousarray(st).view(np.uint8).reshape(-1) if T.fields and n
            else np.zeros(0, np.uint8))


def _last_change(sim, kind, tid):
    v = _abi.C.c_uint64()
    _check(sim.lib.last_change(sim.handle, kind, tid, _abi.C.byref(v)), sim.lib)
    return int(v.value)


def _latest(records, k):
    """find_transition_nr: the largest record number <= k, or None."""
    ks = [r[0] for r in records if r[0] <= k]
    return max(ks) if ks else None


# params / globals values: arrays, numpy scalars and lists in the archive,
# other values (Python scalars, strings, tuples, dicts) as JSON
def _pack_values(obj):
    from . import _fields_of
    arrays, kinds = {}, {}
    for name, v in _fields_of(obj).items():
        if isinstance(v, np.ndarray):
            arrays[name], kinds[name] = v, "array"
        elif isinstance(v, np.generic):
            arrays[name], kinds[name] = np.asarray(v), "npscalar"
        elif isinstance(v, list) and all(isinstance(x, (bool, int, float, np.number)) for x in v):
            arrays[name], kinds[name] = np.asarray(v, dtype=None if v else np.float64), "list"
        else:
            try:
                kinds[name] = ["json", json.loads(json.dumps(v)), isinstance(v, tuple)]
            except TypeError:
                raise TypeError(f"{name}: a {type(v).__name__} can not be stored in a snapshot")
    return arrays, kinds


def _unpack_values(z, kinds):
    out = {}
    for name, kd in kinds.items():
        if kd == "array":
            out[name] = np.array(z[name])
        elif kd == "npscalar":
            out[name] = np.array(z[name])[()]
        elif kd == "list":
            out[name] = np.array(z[name]).tolist()
        else:
            out[name] = tuple(kd[1]) if kd[2] else kd[1]
    return out


def _write_values(fname, obj):
    arrays, kinds = _pack_values(obj)
    np.savez(fname, _kinds=np.frombuffer(json.dumps(kinds).encode(), dtype=np.uint8), **arrays)


def _read_values(fname):
    with np.load(fname) as z:
        kinds = json.loads(bytes(z["_kinds"]).decode())
        return _unpack_values(z, kinds)


def _edge_kind(model, E):
    """The record kind of an edge type (src/HDF5.jl:543-575, 610-632)."""
    h = model.ehints[E]
    stateless = bool(h & _abi.HINT_STATELESS) or not E.fields
    ignorefrom = bool(h & _abi.HINT_IGNOREFROM)
    if stateless and ignorefrom:
        return "count_vector" if h & _abi.HINT_SINGLETYPE else "edge_count"
    if stateless:
        return "stateless"
    return "ignore_from" if ignorefrom else "complete"


# ------------------------------------------------------------------ writing
def _start_file(sim, path):
    """create_h5file! (src/HDF5.jl:171-...): a simulation's first snapshot
    under `path` starts it afresh (overwrite_file = true); the parameters
    and rasters are written once."""
    if os.path.isfile(os.path.join(path, _INDEX)):
        for sub in ("agents", "edges", "globals"):
            d = os.path.join(path, sub)
            if os.path.isdir(d):
                shutil.rmtree(d)
        for f in ("params.npz", "rasters.npz", _INDEX):
            if os.path.isfile(os.path.join(path, f)):
                os.remove(os.path.join(path, f))
    os.makedirs(path, exist_ok=True)
    idx = {"model": sim.model.name, "snapshots": [], "last_snapshot": None,
           "agents": {}, "edges": {}, "globals": []}
    if sim.params is not None:
        _write_values(os.path.join(path, "params.npz"), sim.params)
    rasters = getattr(sim, "rasters", {})
    if rasters:
        np.savez(os.path.join(path, "rasters.npz"), **{n: R.ids for n, R in rasters.items()})
    _save_index(path, idx)
    sim._snapshot_path = path
    return idx


def _write_agents(sim, path, idx, T, k):
    tid = sim._aid(T)
    lc = _last_change(sim, 0, tid)
    recs = idx["agents"].setdefault(T.name, [])
    if recs and recs[-1][1] == lc:
        return                      # unchanged since its last record
    st, died = sim._read_agents(T)
    n = len(died)
    arrays = {"state": _raw(st, T, n)}
    if not sim.model.immortal[T]:
        arrays["died"] = died.astype(np.uint8)
    meta = {"slots": n, "last_change": lc, "dtype": _dtype_meta(T.dtype)}
    os.makedirs(_tdir(path, "agents", T.name), exist_ok=True)
    np.savez(_rec(path, "agents", T.name, k), meta=np.frombuffer(json.dumps(meta).encode(), np.uint8),
             **arrays)
    recs.append([k, lc])


def _write_edges(sim, path, idx, E, k):
    eid = sim._eid(E)
    lc = _last_change(sim, 1, eid)
    recs = idx["edges"].setdefault(E.name, [])
    if recs and recs[-1][1] == lc:
        return
    kind = _edge_kind(sim.model, E)
    to, fr, st = sim._read_edges(E)
    if kind == "complete":
        arrays = {"to": to, "from": fr, "state": _raw(st, E, len(to))}
    elif kind == "stateless":
        arrays = {"to": to, "from": fr}
    elif kind == "ignore_from":
        arrays = {"to": to, "state": _raw(st, E, len(to))}
    elif kind == "edge_count":
        # containers in order of first appearance (the read order)
        u, first, cnt = np.unique(to, return_index=True, return_counts=True)
        o = np.argsort(first, kind="stable")
        arrays = {"to": u[o], "count": cnt[o].astype(np.int64)}
    else:
        T = sim.model.etarget[E]
        nr = (to & _abi.VH_NR_MASK).astype(np.int64)
        cnt = np.bincount(nr, minlength=1)[1:] if len(nr) else np.zeros(0, np.int64)
        if sim.model.ehints[E] & _abi.HINT_SINGLEEDGE:
            cnt = cnt.astype(bool)
        arrays = {"count": cnt, "target": np.array([sim._aid(T)], np.uint8)}
    meta = {"kind": kind, "count": int(len(to)), "last_change": lc, "dtype": _dtype_meta(E.dtype)}
    os.makedirs(_tdir(path, "edges", E.name), exist_ok=True)
    np.savez(_rec(path, "edges", E.name, k), meta=np.frombuffer(json.dumps(meta).encode(), np.uint8),
             **arrays)
    recs.append([k, lc])


def write_snapshot(sim, path, comment="", ignore=()):
    """write_snapshot(sim, comment; ignore) src/HDF5.jl:688-708: the
    current state of an initialised `sim` under directory `path`; the types
    in `ignore` are not written. Returns the number the snapshot is filed
    under (num_transitions - 1)."""
    if not sim.initialized:
        raise AssertionError("write_snapshot needs an initialized simulation")
    ignore = list(ignore) if isinstance(ignore, (list, tuple, set)) else [ignore]
    if getattr(sim, "_snapshot_path", None) == path:
        idx = _load_index(path) or _start_file(sim, path)
    else:
        idx = _start_file(sim, path)
    k = sim._num_transitions() - 1
    idx["snapshots"] = [s for s in idx["snapshots"] if s[0] != k]
    idx["snapshots"].append([k, comment, [getattr(t, "name", str(t)) for t in ignore]])
    idx["last_snapshot"] = k
    if sim.globals is not None:
        recs = idx["globals"]
        if not recs or recs[-1][1] != sim.globals_last_change:
            os.makedirs(os.path.join(path, "globals"), exist_ok=True)
            _write_values(_rec(path, "globals", None, k), sim.globals)
            recs.append([k, sim.globals_last_change])
    for T, _h in sim.model.agent_types:
        if T not in ignore:
            _write_agents(sim, path, idx, T, k)
    for E in sim.model.edge_types:
        if E not in ignore:
            _write_edges(sim, path, idx, E, k)
    _save_index(path, idx)
    return k


# ------------------------------------------------------------------ reading
def list_snapshots(path):
    """list_snapshots src/HDF5.jl:1676-1705: (k, comment) of every snapshot
    under `path`, ascending."""
    idx = _load_index(path)
    if idx is None:
        return []
    return sorted((int(s[0]), s[1]) for s in idx["snapshots"])


def _resolve(idx, path, transition, comment):
    """The snapshot to read: the one with `comment` (_transition_from_comment
    src/HDF5.jl:999-1011), else the latest one <= transition."""
    if idx is None:
        raise FileNotFoundError(f"no snapshot under {path}")
    snaps = sorted((int(s[0]), s[1]) for s in idx["snapshots"])
    if comment:
        hit = [s[0] for s in snaps if s[1] == comment]
        if not hit:
            raise FileNotFoundError(f"no snapshot with comment {comment!r} under {path}")
        return hit[0]
    if transition is None:
        return int(idx["last_snapshot"])
    ks = [s[0] for s in snaps if s[0] <= transition]
    if not ks:
        raise FileNotFoundError(f"no snapshot at or before transition {transition} under {path}")
    return max(ks)


def _open_record(path, idx, kind, name, k):
    """(record number, loaded npz, meta) of the latest record <= k, or None."""
    recs = idx[kind].get(name, [])
    r = _latest(recs, k)
    if r is None:
        return None
    z = np.load(_rec(path, kind, name, r))
    return r, z, json.loads(bytes(z["meta"]).decode())


def read_agents(path, type_, transition=None, comment=""):
    """read_agents(filename, type) src/HDF5.jl:1113-1160: the states of the
    living agents of `type_` (an agent type or its name) as a structured
    array (for a type without fields: the number of living agents)."""
    idx = _load_index(path)
    k = _resolve(idx, path, transition, comment)
    name = getattr(type_, "name", str(type_).lstrip(":"))
    got = _open_record(path, idx, "agents", name, k)
    if got is None:
        raise KeyError(f"no record of agent type {name} under {path}")
    _r, z, meta = got
    with z:
        n = meta["slots"]
        live = ~z["died"].astype(bool) if "died" in z.files else np.ones(n, bool)
        dt = _dtype_of(meta["dtype"])
        if dt is None:
            return int(live.sum())
        st = np.frombuffer(bytearray(z["state"].tobytes()), dtype=dt, count=n)
        return st[live]


def read_edges(path, type_, transition=None, comment=""):
    """read_edges(filename, type) src/HDF5.jl:1292-1330: the edge records of
    `type_` (a structured array with the fields of its record kind: to,
    from, and the state's fields). Types whose records hold no state
    (stateless, edge_count, count_vector) raise, as there."""
    idx = _load_index(path)
    k = _resolve(idx, path, transition, comment)
    name = getattr(type_, "name", str(type_).lstrip(":"))
    got = _open_record(path, idx, "edges", name, k)
    if got is None:
        raise KeyError(f"no record of edge type {name} under {path}")
    _r, z, meta = got
    with z:
        if meta["kind"] not in ("complete", "ignore_from"):
            raise AssertionError("Can not read the state of stateless edges")
        dt = _dtype_of(meta["dtype"])
        n = meta["count"]
        st = np.frombuffer(bytearray(z["state"].tobytes()), dtype=dt, count=n)
        fields = [("to", "<u8")] + ([("from", "<u8")] if meta["kind"] == "complete" else [])
        out = np.zeros(n, dtype=fields + [(f, st.dtype.fields[f][0]) for f in st.dtype.names])
        out["to"] = z["to"]
        if meta["kind"] == "complete":
            out["from"] = z["from"]
        for f in st.dtype.names:
            out[f] = st[f]
        return out


def read_params(path):
    """read_params(filename) src/HDF5.jl:938-995: a dict of the parameters."""
    f = os.path.join(path, "params.npz")
    return _read_values(f) if os.path.isfile(f) else None


def read_globals(path, transition=None, comment=""):
    """read_globals(filename) src/HDF5.jl:1013-1090: a dict of the globals of
    the snapshot (their latest record <= it)."""
    idx = _load_index(path)
    k = _resolve(idx, path, transition, comment)
    r = _latest(idx["globals"], k)
    return None if r is None else _read_values(_rec(path, "globals", None, r))


def read_snapshot(sim, path, transition=None, comment="", writeable=False, ignore_params=False):
    """read_snapshot!(sim, name; transition, comment, writeable,
    ignore_params) src/HDF5.jl:1594-1646 into `sim`, a new simulation of the
    same model (no agents or edges added yet); `sim` is initialised
    afterwards. `writeable`: later write_snapshot calls of `sim` under `path`
    add to these files. Returns the number of the snapshot read."""
    from . import _hasf, _setf
    from .raster import Raster, _rasters
    if sim.initialized:
        raise AssertionError("read_snapshot restores into a new simulation of the model")
    idx = _load_index(path)
    k = _resolve(idx, path, transition, comment)
    model = sim.model
    lcs = []
    died_of = {}
    for T, _h in model.agent_types:
        got = _open_record(path, idx, "agents", T.name, k)
        if got is None:
            continue
        r, z, meta = got
        with z:
            if meta["dtype"]["itemsize"] != T.nbytes:
                raise AssertionError(f"snapshot agent type {T.name} does not match the model")
            n = meta["slots"]
            if n:
                if T.fields:
                    st = np.frombuffer(bytearray(z["state"].tobytes()), dtype=T.dtype, count=n)
                    sim._add_agents_array(T, st)
                else:
                    sim._add_agents(T, [T()] * n)
            died_of[T] = z["died"].astype(np.uint8) if "died" in z.files else None
        lcs.append((0, sim._aid(T), r))
    for E in model.edge_types:
        got = _open_record(path, idx, "edges", E.name, k)
        if got is None:
            continue
        r, z, meta = got
        with z:
            if meta["dtype"]["itemsize"] != E.nbytes or meta["kind"] != _edge_kind(model, E):
                raise AssertionError(f"snapshot edge type {E.name} does not match the model")
            kind = meta["kind"]
            n = meta["count"]
            if n:
                st = None
                if kind in ("complete", "ignore_from") and E.fields:
                    st = np.frombuffer(bytearray(z["state"].tobytes()), dtype=E.dtype, count=n)
                if kind in ("complete", "stateless"):
                    to, fr = z["to"], z["from"]
                elif kind == "ignore_from":
                    to = z["to"]
                    fr = np.zeros(n, np.uint64)       # a dummy source (src/HDF5.jl:1500-1505)
                elif kind == "edge_count":
                    to = np.repeat(z["to"], z["count"])
                    fr = np.zeros(len(to), np.uint64)
                else:
                    cnt = z["count"].astype(np.int64)
                    tt = int(z["target"][0])
                    nr = np.repeat(np.arange(1, len(cnt) + 1, dtype=np.uint64), cnt)
                    to = (np.uint64(tt) << np.uint64(_abi.VH_SHIFT_TYPE)) | nr
                    fr = np.zeros(len(to), np.uint64)
                sim._add_edges_array(E, fr, to, st)
        lcs.append((1, sim._eid(E), r))
    rp = os.path.join(path, "rasters.npz")
    if os.path.isfile(rp):
        with np.load(rp) as z:
            for name in z.files:
                _rasters(sim)[name] = Raster(np.array(z[name], dtype=np.uint64))
    # parameters, unless ignore_params (src/HDF5.jl:1620-1622); constant
    # from finish_init! on
    if not ignore_params and sim.params is not None:
        for name, v in (read_params(path) or {}).items():
            if _hasf(sim.params, name):
                _setf(sim.params, name, v)
                sim._push_value(_abi.VALUE_PARAM, name, v)
    sim._finish_init()
    lib = sim.lib
    for T, died in died_of.items():
        if died is None:
            continue
        d = np.ascontiguousarray(died)
        _check(lib.restore_died(sim.handle, sim._aid(T),
                                d.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint8)), len(d)), lib)
    _check(lib.set_num_transitions(sim.handle, k + 1), lib)
    for kind, tid, r in lcs:
        _check(lib.set_last_change(sim.handle, kind, tid, r), lib)
    # globals (src/HDF5.jl:1615-1627)
    rg = _latest(idx["globals"], k)
    if rg is not None:
        sim.globals_last_change = rg
        if sim.globals is not None:
            for name, v in _read_values(_rec(path, "globals", None, rg)).items():
                if _hasf(sim.globals, name):
                    _setf(sim.globals, name, v)
                    sim._push_value(_abi.VALUE_GLOBAL, name, v)
    if writeable:
        sim._snapshot_path = path
    return k


def last_change(sim, type_):
    """The transition that last changed an agent or edge type (its
    last_change field): apply! stamps it on the types it writes, and on edge
    types that lost edges to deaths (src/AgentMethods.jl:347-353, 433,
    src/EdgeMethods.jl:682)."""
    if type_ in sim.model.eid:
        return _last_change(sim, 1, sim._eid(type_))
    return _last_change(sim, 0, sim._aid(type_))
