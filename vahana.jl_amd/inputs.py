"""Seeded synthetic inputs of the benchmark models (counter-hash driven).

The graph generators are the library's host functions vh_graph_er /
vh_graph_ba; the per-agent initial values are numpy restatements of the same
SplitMix64-chained hash that the device bodies use.
"""
import ctypes as C

import numpy as np

from . import _abi

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
KEY_INIT = 5 << 40
KEY_WEIGHT = 6 << 40
KEY_OPINION = 7 << 40
KEY_SUSC = 8 << 40


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def hash4(seed, a, b, c):
    """Vectorised vh_hash4(seed, a, b, c) on uint64 arrays/scalars."""
    with np.errstate(over="ignore"):
        s = np.asarray(seed, dtype=np.uint64)
        h = _mix64(s + GOLDEN)
        h = _mix64(h + np.asarray(a, dtype=np.uint64) + GOLDEN)
        h = _mix64(h + np.asarray(b, dtype=np.uint64) + GOLDEN)
        return _mix64(h + np.asarray(c, dtype=np.uint64) + GOLDEN)


def u01(h):
    return (np.asarray(h, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def u01f(h):
    return ((np.asarray(h, dtype=np.uint64) >> np.uint64(40)).astype(np.float32) *
            np.float32(2.0 ** -24))


def open01f(h):
    """Float32 in (0, 1): an odd multiple of 2^-24 (exact in Float32)."""
    k = (np.asarray(h, dtype=np.uint64) >> np.uint64(41)).astype(np.float64) * 2.0 + 1.0
    return (k * 2.0 ** -24).astype(np.float32)


def graph_er(n, p, seed):
    """Directed G(n, p): (from, to) 1-based agent numbers, targets ascending."""
    L = _abi.lib()
    m = C.c_uint64()
    L.graph_er(n, p, seed, None, None, 0, C.byref(m))
    cap = m.value
    fr = np.empty(max(cap, 1), dtype=np.uint32)
    to = np.empty(max(cap, 1), dtype=np.uint32)
    rc = L.graph_er(n, p, seed, fr.ctypes.data_as(C.POINTER(C.c_uint32)),
                    to.ctypes.data_as(C.POINTER(C.c_uint32)), cap, C.byref(m))
    if rc != 0:
        raise RuntimeError(f"vh_graph_er failed ({rc})")
    return fr[:cap], to[:cap]


def graph_gnp(n, mean_deg, seed):
    """Directed G(n, p) with exact Binomial(n-1, p) in-degrees, the host twin of
    the ensemble's device generator: (from, to), targets ascending."""
    L = _abi.lib()
    m = C.c_uint64()
    L.graph_gnp(n, mean_deg, seed, None, None, 0, C.byref(m))
    cap = m.value
    fr = np.empty(max(cap, 1), dtype=np.uint32)
    to = np.empty(max(cap, 1), dtype=np.uint32)
    rc = L.graph_gnp(n, mean_deg, seed, fr.ctypes.data_as(C.POINTER(C.c_uint32)),
                     to.ctypes.data_as(C.POINTER(C.c_uint32)), cap, C.byref(m))
    if rc != 0:
        raise RuntimeError(f"vh_graph_gnp failed ({rc})")
    return fr[:cap], to[:cap]


def graph_ba(n, m, seed):
    """Barabasi-Albert undirected pairs (a, b) in creation order."""
    L = _abi.lib()
    cnt = C.c_uint64()
    L.graph_ba(n, m, seed, None, None, 0, C.byref(cnt))
    k = cnt.value
    a = np.empty(max(k, 1), dtype=np.uint32)
    b = np.empty(max(k, 1), dtype=np.uint32)
    rc = L.graph_ba(n, m, seed, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                    b.ctypes.data_as(C.POINTER(C.c_uint32)), k, C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"vh_graph_ba failed ({rc})")
    return a[:k], b[:k]


def bfs_order(n, a, b):
    """new_nr[v - 1] of the breadth-first numbering of vertices 1..n over the
    undirected edges (a[k], b[k]) (vh_bfs_order, the order
    finish_init_(sim, relabel=T) applies)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    rc = _abi.lib().bfs_order(n, len(a), a.ctypes.data_as(C.POINTER(C.c_uint32)),
                              b.ctypes.data_as(C.POINTER(C.c_uint32)),
                              out.ctypes.data_as(C.POINTER(C.c_uint64)))
    if rc != 0:
        raise RuntimeError(f"vh_bfs_order failed ({rc})")
    return out[:n]


def gnp_weights(n_edges, wseed):
    """Float32 weights in (0, 1) by edge position (target order), the values
    the device writes for vh_add_edges_gnp."""
    w = np.empty(max(n_edges, 1), dtype=np.float32)
    _abi.lib().fill_open01f(wseed, 0, KEY_WEIGHT, n_edges, w.ctypes.data_as(C.POINTER(C.c_float)))
    return w[:n_edges]


def symmetric_weighted(a, b, wseed):
    """Both directions of every undirected pair (u->v then v->u, in pair order)
    with a Float32 weight in (0, 1) per directed edge."""
    k = len(a)
    fr = np.empty(2 * k, dtype=np.uint32)
    to = np.empty(2 * k, dtype=np.uint32)
    fr[0::2], to[0::2] = a, b
    fr[1::2], to[1::2] = b, a
    w = np.empty(2 * k, dtype=np.float32)
    _abi.lib().fill_open01f(wseed, 0, KEY_WEIGHT, 2 * k, w.ctypes.data_as(C.POINTER(C.c_float)))
    return fr, to, w


SIR_DTYPE = np.dtype([("health", "u1"), ("pad", "u1"), ("days", "<u2")], align=True)
CONTACT_DTYPE = np.dtype([("health", "u1"), ("susc", "<f4")], align=True)


def sir_init(n, seed, n_init, type_id=1):
    """All S except the n_init agents with the smallest hash(seed, id, 0, INIT)."""
    st = np.zeros(n, dtype=SIR_DTYPE)
    ids = (np.uint64(type_id) << np.uint64(56)) | np.arange(1, n + 1, dtype=np.uint64)
    h = hash4(seed, ids, 0, KEY_INIT)
    k = min(n_init, n)
    if k:
        sel = np.argsort(h, kind="stable")[:k]
        st["health"][sel] = _abi.I
    return st


def sir_init_frac(n, seed, n_init, type_id=1):
    """All S except the agents with u01(hash(seed, id, 0, INIT)) < n_init / n
    (the ensemble's initial state)."""
    st = np.zeros(n, dtype=SIR_DTYPE)
    ids = (np.uint64(type_id) << np.uint64(56)) | np.arange(1, n + 1, dtype=np.uint64)
    st["health"][u01(hash4(seed, ids, 0, KEY_INIT)) < n_init / n] = _abi.I
    return st


def opinion_init(n, seed, f32=False, type_id=1):
    """u01(hash4(seed, id, 0, KEY_OPINION)) per agent (filled in C)."""
    x = np.empty(max(n, 1), dtype=np.float64)
    _abi.lib().fill_agent_u01(seed, type_id, 0, KEY_OPINION, n, x.ctypes.data_as(C.POINTER(C.c_double)))
    x = x[:n]
    return x.astype(np.float32) if f32 else x


def contact_init(n, seed, n_init, type_id=1):
    st = np.zeros(n, dtype=CONTACT_DTYPE)
    ids = (np.uint64(type_id) << np.uint64(56)) | np.arange(1, n + 1, dtype=np.uint64)
    st["susc"] = np.float32(0.5) + np.float32(0.5) * u01f(hash4(seed, ids, 0, KEY_SUSC))
    h = hash4(seed, ids, 0, KEY_INIT)
    k = min(n_init, n)
    if k:
        st["health"][np.argsort(h, kind="stable")[:k]] = _abi.I
    return st
