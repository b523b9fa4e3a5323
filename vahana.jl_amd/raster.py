"""Raster subsystem (src/Raster.jl): n-dimensional grids of cell agents.

Host-side construction and readout, as in the reference (it runs these on
the root process): `add_raster_` adds one agent per cell in column-major
order (CartesianIndices, first dimension fastest), `connect_raster_neighbors_`
adds the stencil edges between cells, `move_to_` connects an agent to a cell
and its surroundings, `calc_raster` / `calc_rasterstate` / `rastervalues`
read per-cell values back as an array of the raster's shape. The transitions
that run on rasters are device bodies (transitions.diffuse,
transitions.move_to_value); large stencils can be generated on the device
(`connect_raster_neighbors_` with a uniform edge state, vh_add_edges_stencil).
Positions are 1-based tuples, like Julia's CartesianIndex.
"""
import ctypes as C
import itertools
import math

import numpy as np

from . import _abi

_METRICS = {"chebyshev": 0, "euclidean": 1, "manhatten": 2, "manhattan": 2}


class Raster:
    """The cell ids of one raster: `ids[pos - 1]` is the agent at `pos`."""

    def __init__(self, ids):
        self.ids = ids

    @property
    def dims(self):
        return self.ids.shape

    def contiguous(self):
        """(cell type, agent number of cell (1,...,1)) when the cells are one
        agent type with consecutive numbers in column-major order, else None."""
        flat = self.ids.reshape(-1, order="F")
        t = flat >> np.uint64(56)
        nr = flat & np.uint64(_abi.VH_NR_MASK)
        if not (t == t[0]).all():
            return None
        if not (nr == nr[0] + np.arange(len(nr), dtype=np.uint64)).all():
            return None
        return int(t[0]), int(nr[0])


def _cell_type(sim, R):
    tid = int(R.ids.flat[0] >> np.uint64(56))
    for T, i in sim.model.aid.items():
        if i == tid:
            return T
    raise AssertionError("raster cells of an unregistered type")


def _rasters(sim):
    if not hasattr(sim, "rasters"):
        sim.rasters = {}
    return sim.rasters


def positions(dims):
    """CartesianIndices(dims) in iteration order (first dimension fastest)."""
    for t in itertools.product(*[range(1, d + 1) for d in reversed(dims)]):
        yield tuple(reversed(t))


def add_raster_(sim, name, dims, agent_constructor, T=None):
    """add_raster! src/Raster.jl:32-56: one agent per cell, built by
    `agent_constructor(pos)`, added in column-major order; returns the ids.
    Bulk form: `agent_constructor` a structured array of T with the cells'
    states in column-major order."""
    from . import add_agent_, add_agents_, _type_of_value
    if sim.initialized:
        raise AssertionError("add_raster! can be only called before finish_init!")
    dims = tuple(int(d) for d in dims)
    if isinstance(agent_constructor, np.ndarray):
        if T is None or len(agent_constructor) != int(np.prod(dims)):
            raise AssertionError("bulk add_raster_: the type and one state per cell")
        ids = sim._add_agents_array(T, agent_constructor)
        _rasters(sim)[name] = Raster(np.asarray(ids, dtype=np.uint64).reshape(dims, order="F"))
        return ids
    values = [agent_constructor(p) for p in positions(dims)]
    types = {id(_type_of_value(sim, v)) for v in values}
    if len(types) == 1:
        ids = add_agents_(sim, values)
    else:
        ids = [add_agent_(sim, v) for v in values]
    grid = np.array(ids, dtype=np.uint64).reshape(dims, order="F")
    _rasters(sim)[name] = Raster(grid)
    return list(ids)


def stencil(metric, n, distance):
    """The offsets of `_stencil_core` src/Raster.jl:80-96: the cube
    [-d, d]^n (d = floor(distance)), first component fastest, without the
    origin, filtered by the metric (Euclidean norm or Manhattan sum
    <= distance; Chebyshev keeps the cube)."""
    if metric not in _METRICS:
        raise AssertionError(f"unknown metric {metric}")
    d = int(math.floor(distance))
    out = []
    for t in itertools.product(range(-d, d + 1), repeat=n):
        s = tuple(reversed(t))
        if not any(s):
            continue
        if _METRICS[metric] == 1 and math.sqrt(sum(x * x for x in s)) > distance:
            continue
        if _METRICS[metric] == 2 and sum(abs(x) for x in s) > distance:
            continue
        out.append(s)
    return out


def _checkpos(pos, dims, periodic):
    """src/Raster.jl:477-495: the position, wrapped (periodic) or None when
    outside the raster."""
    if all(1 <= p <= d for p, d in zip(pos, dims)):
        return pos
    if not periodic:
        return None
    return tuple((p - 1) % d + 1 for p, d in zip(pos, dims))


def connect_raster_neighbors_(sim, name, edge_constructor, distance=1, metric="chebyshev",
                              periodic=True, device=None):
    """connect_raster_neighbors! src/Raster.jl:134-164: for every cell `org`
    (column-major) and every stencil offset `s` (in stencil order) an edge
    org -> org + s, state `edge_constructor(org, shifted)`.

    `edge_constructor` may also be a value (the same state for every edge);
    then, on a device simulation whose cells are consecutive agents of one
    type, the edges are generated on the device (vh_add_edges_stencil)
    unless device=False. The CSR is the same either way."""
    from . import Value, StructType
    R = _rasters(sim)[name]
    dims = R.dims
    st = stencil(metric, len(dims), distance)
    uniform = isinstance(edge_constructor, Value) or (
        isinstance(edge_constructor, tuple) and len(edge_constructor) == 2
        and isinstance(edge_constructor[0], StructType))
    if uniform:
        val = edge_constructor
        T = val.T if isinstance(val, Value) else val[0]
        raw = T.to_bytes([val.v if isinstance(val, Value) else val[1]]) if T.fields else b""
        contig = R.contiguous()
        can = (contig is not None and hasattr(sim.lib, "add_edges_stencil") and len(dims) <= 4
               and len(raw) <= 16)
        if device is None:
            device = can
        if device:
            if not can:
                raise AssertionError("the device stencil needs consecutive cells of one agent "
                                     "type, at most 4 dimensions and an edge state of <= 16 bytes")
            return _device_stencil(sim, T, R, contig, distance, metric, periodic, raw)
    org, tgt = _stencil_pairs(dims, st, periodic)
    if len(org) == 0:
        return
    froms = R.ids.reshape(-1, order="F")[org]
    tos = R.ids.reshape(-1, order="F")[tgt]
    if uniform:
        v = edge_constructor
        T = v.T if isinstance(v, Value) else v[0]
        one = v.v if isinstance(v, Value) else v[1]
        states = [one] * len(org) if T.fields else None
    else:
        vals = [edge_constructor(_pos(o, dims), _pos(t, dims)) for o, t in zip(org, tgt)]
        T = vals[0].T if isinstance(vals[0], Value) else vals[0][0]
        states = [v.v if isinstance(v, Value) else v[1] for v in vals] if T.fields else None
    sim._add_edges_array(T, froms, tos, states)


def _pos(lin, dims):
    return tuple(int(x) + 1 for x in np.unravel_index(int(lin), dims, order="F"))


def _stencil_pairs(dims, st, periodic):
    """(org, target) cell indices (column-major) of every stencil edge, in
    insertion order: org ascending, then stencil order (src/Raster.jl:151-161)."""
    n = int(np.prod(dims))
    if not st:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    D = np.array(dims, dtype=np.int64)
    S = np.array(st, dtype=np.int64)                       # (nst, nd)
    coords = np.stack(np.unravel_index(np.arange(n), dims, order="F"), axis=1)   # (n, nd)
    sh = coords[:, None, :] + S[None, :, :]                # (n, nst, nd)
    if periodic:
        sh %= D
        valid = np.ones(sh.shape[:2], dtype=bool)
    else:
        valid = ((sh >= 0) & (sh < D)).all(axis=2)
        sh = np.where(valid[:, :, None], sh, 0)
    strides = np.cumprod(np.concatenate([[1], D[:-1]]))
    tgt = (sh * strides).sum(axis=2)
    org = np.broadcast_to(np.arange(n)[:, None], tgt.shape)
    return org[valid], tgt[valid]


def _device_stencil(sim, T, R, contig, distance, metric, periodic, raw):
    dims = (C.c_int64 * 4)(*(list(R.dims) + [1] * (4 - len(R.dims))))
    st = (C.c_uint8 * 16)(*raw.ljust(16, b"\0"))
    cell_type, first_nr = contig
    rc = sim.lib.add_edges_stencil(sim.handle, sim._eid(T), cell_type, first_nr, len(R.dims), dims,
                                   float(distance), _METRICS[metric], 1 if periodic else 0, st)
    if rc != _abi.VH_OK:
        from . import _raise
        _raise(rc, sim.lib)


def cellid(sim, name, pos):
    """cellid src/Raster.jl:404-406."""
    return int(_rasters(sim)[name].ids[tuple(int(p) - 1 for p in pos)])


def move_to_(sim, name, id_, pos, edge_from_raster, edge_to_raster, distance=0,
             metric="chebyshev", periodic=True, only_surrounding=False):
    """move_to! src/Raster.jl:437-475 (initialisation phase): edges between
    the agent and the cell at `pos` (unless only_surrounding) and the cells
    of its stencil surroundings; `edge_from_raster` cell -> agent,
    `edge_to_raster` agent -> cell (None: not added)."""
    from . import add_edge_
    R = _rasters(sim)[name]
    dims = R.dims
    pos = tuple(int(p) for p in pos)

    def link(cell_pos):
        cid = int(R.ids[tuple(p - 1 for p in cell_pos)])
        if edge_from_raster is not None:
            add_edge_(sim, cid, id_, edge_from_raster)
        if edge_to_raster is not None:
            add_edge_(sim, id_, cid, edge_to_raster)

    if not only_surrounding:
        link(pos)
    if distance >= 1:
        for s in stencil(metric, len(dims), distance):
            shifted = _checkpos(tuple(p + x for p, x in zip(pos, s)), dims, periodic)
            if shifted is not None:
                link(shifted)


def _states_by_cell(sim, name, T):
    R = _rasters(sim)[name]
    st, _ = sim._read_agents(T)
    nr = (R.ids & np.uint64(_abi.VH_NR_MASK)).astype(np.int64) - 1
    return st, nr, R


def calc_rasterstate(sim, name, f, f_returns=None, T=None):
    """calc_rasterstate src/Raster.jl:298-338: f(state) per cell, as an array
    of the raster's shape (T: the cells' type; None: the type of the first
    cell, all cells of that type)."""
    R = _rasters(sim)[name]
    if T is None:
        T = _cell_type(sim, R)
    st, nr, R = _states_by_cell(sim, name, T)
    vals = [f(st[i]) for i in nr.reshape(-1, order="F")]
    out = np.array(vals, dtype=f_returns if f_returns is not None else None)
    return out.reshape(R.dims, order="F")


def rastervalues(sim, name, field):
    """rastervalues src/Raster.jl:363-390: the field of every cell."""
    T = _cell_type(sim, _rasters(sim)[name])
    st, nr, R = _states_by_cell(sim, name, T)
    return np.asarray(st[field])[nr]


def calc_raster(sim, name, f, f_returns=None, accessible=None):
    """calc_raster src/Raster.jl:200-230: f(cell id) per cell."""
    R = _rasters(sim)[name]
    vals = [f(int(i)) for i in R.ids.reshape(-1, order="F")]
    return np.array(vals, dtype=f_returns).reshape(R.dims, order="F")


def random_pos(sim, name, weights=None, rng=None):
    """random_pos src/Raster.jl:506-533: a position, uniform or weighted by
    `weights` (an array of the raster's shape)."""
    R = _rasters(sim)[name]
    rng = rng if rng is not None else np.random.default_rng()
    n = R.ids.size
    if weights is None:
        k = int(rng.integers(n))
    else:
        w = np.asarray(weights, dtype=np.float64)
        if w.shape != R.dims:
            raise AssertionError(f"`weights` must have the same dimension as the raster :{name}")
        w = w.reshape(-1, order="F")
        k = int(rng.choice(n, p=w / w.sum()))
    return tuple(int(x) + 1 for x in np.unravel_index(k, R.dims, order="F"))


def random_cell(sim, name, weights=None, rng=None):
    """random_cell src/Raster.jl:541-568: the id of a random cell."""
    return cellid(sim, name, random_pos(sim, name, weights, rng))
