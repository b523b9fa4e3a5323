"""Checkpoint / resume of a simulation (src/HDF5.jl).

`write_snapshot` (src/HDF5.jl:688-708) records, per agent type, the state
and died flag of every slot (write_agents :483-541), per edge type the
edges as (to, from, state) records in container order (write_edges
:577-673), the rasters' cell ids (:1533-1592), the transition number and a
comment. `read_snapshot` (src/HDF5.jl:1594-1646) rebuilds a simulation of
the same model from it: slots with their states, the edges re-added in
record order (so every container keeps its order), died flags with the
free slots rebuilt in ascending order (_read_agents_restore! :801-850), the
transition counter. `list_snapshots` (src/HDF5.jl:1676-1705) lists them.

The file layout follows the reference's groups (`agents/<T>/state`,
`agents/<T>/died`, `edges/<E>/to|from|state`, `rasters/<name>`) but is a
NumPy archive per snapshot, `<path>/t_<transition>.npz`: no HDF5 library is
available to this build. The same files restore into the device library
and into the checker.
"""
import json
import os

import numpy as np

from . import _abi, _check


def _file(path, transition):
    return os.path.join(path, f"t_{int(transition)}.npz")


def write_snapshot(sim, path, comment=""):
    """Write the current state of an initialised `sim` under directory
    `path`; returns the transition number it is filed under."""
    if not sim.initialized:
        raise AssertionError("write_snapshot needs an initialized simulation")
    t = sim._num_transitions()
    arrays = {}
    meta = {"transition": int(t), "comment": comment, "model": sim.model.name,
            "agents": [], "edges": [], "rasters": []}
    for T, _hints in sim.model.agent_types:
        st, died = sim._read_agents(T)
        n = len(died)
        raw = (np.ascontiguousarray(st).view(np.uint8).reshape(-1) if T.fields and n
               else np.zeros(0, np.uint8))
        arrays[f"agents/{T.name}/state"] = raw
        arrays[f"agents/{T.name}/died"] = died.astype(np.uint8)
        meta["agents"].append({"name": T.name, "bytes": T.nbytes, "slots": n})
    for E in sim.model.edge_types:
        to, fr, st = sim._read_edges(E)
        arrays[f"edges/{E.name}/to"] = to
        arrays[f"edges/{E.name}/from"] = fr
        arrays[f"edges/{E.name}/state"] = (np.ascontiguousarray(st).view(np.uint8).reshape(-1)
                                           if E.fields and len(to) else np.zeros(0, np.uint8))
        meta["edges"].append({"name": E.name, "bytes": E.nbytes, "count": int(len(to))})
    for name, R in getattr(sim, "rasters", {}).items():
        arrays[f"rasters/{name}"] = R.ids
        meta["rasters"].append(name)
    os.makedirs(path, exist_ok=True)
    np.savez(_file(path, t), meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8),
             **arrays)
    return t


def list_snapshots(path):
    """(transition, comment) of every snapshot under `path`, ascending."""
    out = []
    if not os.path.isdir(path):
        return out
    for f in os.listdir(path):
        if f.startswith("t_") and f.endswith(".npz"):
            with np.load(os.path.join(path, f)) as z:
                meta = json.loads(bytes(z["meta"]).decode())
            out.append((meta["transition"], meta["comment"]))
    return sorted(out)


def read_snapshot(sim, path, transition=None):
    """Restore the snapshot filed under `transition` (default: the latest)
    into `sim`, a new simulation of the same model (no agents or edges added
    yet); `sim` is initialised afterwards. Returns the transition number."""
    from .raster import Raster, _rasters
    if sim.initialized:
        raise AssertionError("read_snapshot restores into a new simulation of the model")
    if transition is None:
        snaps = list_snapshots(path)
        if not snaps:
            raise FileNotFoundError(f"no snapshot under {path}")
        transition = snaps[-1][0]
    with np.load(_file(path, transition)) as z:
        meta = json.loads(bytes(z["meta"]).decode())
        types = {T.name: T for T, _h in sim.model.agent_types}
        etypes = {E.name: E for E in sim.model.edge_types}
        died_of = {}
        for a in meta["agents"]:
            T = types.get(a["name"])
            if T is None or T.nbytes != a["bytes"]:
                raise AssertionError(f"snapshot agent type {a['name']} does not match the model")
            n = a["slots"]
            if n:
                if T.fields:
                    st = np.frombuffer(bytearray(z[f"agents/{T.name}/state"].tobytes()),
                                       dtype=T.dtype, count=n)
                    sim._add_agents_array(T, st)
                else:
                    sim._add_agents(T, [T()] * n)
            died_of[T] = z[f"agents/{T.name}/died"].astype(np.uint8)
        for e in meta["edges"]:
            E = etypes.get(e["name"])
            if E is None or E.nbytes != e["bytes"]:
                raise AssertionError(f"snapshot edge type {e['name']} does not match the model")
            if e["count"]:
                st = (np.frombuffer(bytearray(z[f"edges/{E.name}/state"].tobytes()), dtype=E.dtype,
                                    count=e["count"]) if E.fields else None)
                sim._add_edges_array(E, z[f"edges/{E.name}/from"], z[f"edges/{E.name}/to"], st)
        for name in meta["rasters"]:
            _rasters(sim)[name] = Raster(np.array(z[f"rasters/{name}"], dtype=np.uint64))
    sim._finish_init()
    lib = sim.lib
    for T, died in died_of.items():
        d = np.ascontiguousarray(died)
        _check(lib.restore_died(sim.handle, sim._aid(T),
                                d.ctypes.data_as(_abi.C.POINTER(_abi.C.c_uint8)), len(d)), lib)
    _check(lib.set_num_transitions(sim.handle, int(meta["transition"])), lib)
    return int(meta["transition"])
