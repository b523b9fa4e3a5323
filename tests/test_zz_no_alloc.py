"""vh_mapreduce_noalloc (no device allocation) against vh_mapreduce, through
the C ABI: two calls in a row on one GPU, the second on a smaller input,
bit-identical results for agent and edge reductions of every operator."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(vh, sim, kind, T, f, op, dtype):
    """(vh_mapreduce, vh_mapreduce_noalloc) raw 8-byte results."""
    L = vh._abi.lib()
    md, _ = vh._map_desc(T, f)
    tid = sim._aid(T) if kind == 0 else sim._eid(T)
    a, b = C.c_uint64(), C.c_uint64()
    assert L.mapreduce(sim.handle, kind, tid, C.byref(md), vh._abi.OP[op], vh._abi.DT[dtype], None,
                       C.byref(a)) == 0, L.last_error()
    assert L.mapreduce_noalloc(sim.handle, kind, tid, C.byref(md), vh._abi.OP[op], vh._abi.DT[dtype], None,
                               C.byref(b)) == 0, L.last_error()
    return a.value, b.value


CASES_SIR = [("days", "+", "i8"), ("days", "max", "i8"), ("days", "min", "i8"),
             (("health", 1), "+", "i8"), (("health", 2), "|", "b1"), (None, "+", "i8")]


def _check_sir(vh, m):
    for f, op, dt in CASES_SIR:
        if isinstance(f, tuple):
            f = vh.Field(f[0], f[1])
        a, b = _both(vh, m.sim, 0, m.T, f, op, dt)
        assert a == b, (f, op, dt)


def _check_opinion(vh, m):
    for f, op in (("opinion", "+"), ("opinion", "max"), ("opinion", "min"), ("opinion", "*")):
        a, b = _both(vh, m.sim, 0, m.T, f, op, "f8")
        assert a == b, (f, op)
    for op in ("+", "max"):
        a, b = _both(vh, m.sim, 1, m.E, "w", op, "f8")
        assert a == b, ("w", op)


def test_noalloc_mapreduce_matches(vh, gpu_ready):
    M = vh.models
    print("[gpu] mapreduce_noalloc: SIR 3e5 agents (~147 level-1 blocks), opinion 2e5 agents / 2e6 edges",
          flush=True)
    big = M.build_sir(vh, 300_000, seed=5, n_init=3000)
    for _ in range(6):
        M.step_sir(vh, big, p_inf=0.2, seed=5)
    _check_sir(vh, big)
    op = M.build_opinion(vh, 200_000, seed=4)
    M.step_opinion(vh, op)
    _check_opinion(vh, op)
    print("[gpu] mapreduce_noalloc: second call on a smaller input (SIR 2000, opinion 3000)", flush=True)
    small = M.build_sir(vh, 2000, seed=6, n_init=40)
    for _ in range(3):
        M.step_sir(vh, small, p_inf=0.2, seed=6)
    _check_sir(vh, small)
    op2 = M.build_opinion(vh, 3000, seed=7)
    M.step_opinion(vh, op2)
    _check_opinion(vh, op2)
    # and the larger simulation again after the smaller one
    _check_sir(vh, big)
    # the result equals the host sum of the states
    st, _ = small.sim._read_agents(small.T)
    a, b = _both(vh, small.sim, 0, small.T, "days", "+", "i8")
    assert b == int(np.asarray(st)["days"].astype(np.int64).sum())


def test_noalloc_mapreduce_mortal_and_empty(vh, gpu_ready):
    """Removed agents are the identity (contact deaths), and an empty edge
    set gives the operator's identity."""
    M = vh.models
    print("[gpu] mapreduce_noalloc: mortal contact agents", flush=True)
    m = M.build_contact(vh, 20_000, seed=2)
    for _ in range(3):
        M.step_contact(vh, m, seed=2, p_death=0.05)
    a, b = _both(vh, m.sim, 0, m.T, None, "+", "i8")
    assert a == b == 20_000
    a, b = _both(vh, m.sim, 0, m.T, "susc", "max", "f8")
    assert a == b
