"""mapreduce with computed maps (vh_mapreduce_expr) on the checker: every
expression's per-record values from an independent numpy evaluation, folded
in the reference's sequential order (src/AgentMethods.jl:545-553), give the
checker's result exactly."""
import math

import numpy as np
import pytest

import expr_cases as EC


def _init(op, datatype):
    return {"+": 0.0 if datatype == "f8" else 0, "*": 1.0 if datatype == "f8" else 1,
            "max": -math.inf if datatype == "f8" else -(2 ** 63 - 1),
            "min": math.inf if datatype == "f8" else 2 ** 63 - 1,
            "&": True if datatype == "b1" else 2 ** 63 - 1, "|": False if datatype == "b1" else 0}[op]


def _same(a, b):
    if isinstance(a, float) and isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) == math.copysign(1, b))
    return a == b


@pytest.mark.parametrize("name", list(EC.exprs(__import__("oracle_api").expr)))
@pytest.mark.parametrize("op", ["+", "max", "min"])
def test_expr_values_sequential(ora, name, op):
    X = ora.expr
    sim, A, E = EC.build(ora)
    e = EC.exprs(X)[name]
    st = np.asarray(sim._read_agents(A)[0])
    vals = EC.np_eval(e, st)
    kind = EC._kind(vals)
    datatype = "f8" if kind == "f" else ("i8" if op == "+" or kind == "i" else "b1")
    if datatype == "b1" and op in ("max", "min"):
        pytest.skip("max/min over Bool: covered by & and |")
    want = EC.seq_fold(EC.to_dtype(vals, datatype), op, datatype, _init(op, datatype))
    sim.order = 0
    got = ora.mapreduce(sim, e, op, A, datatype=datatype)
    assert _same(got, want), (name, op, got, want)


@pytest.mark.parametrize("name,op,datatype", [("gt0", "&", "b1"), ("flag_and", "|", "b1"),
                                              ("select_int", "&", "i8"), ("lin", "+", "i8"),
                                              ("divzero", "+", "i8"), ("lin", "*", "f8")])
def test_expr_conversions(ora, name, op, datatype):
    """The program's value in another reduction type: floats truncated to
    Int64 (NaN 0, infinities saturated), any value to Bool as != 0."""
    X = ora.expr
    sim, A, E = EC.build(ora, n=300)
    e = EC.exprs(X)[name]
    st = np.asarray(sim._read_agents(A)[0])
    want = EC.seq_fold(EC.to_dtype(EC.np_eval(e, st), datatype), op, datatype, _init(op, datatype))
    sim.order = 0
    assert _same(ora.mapreduce(sim, e, op, A, datatype=datatype), want)


def test_expr_edges_and_deaths(ora):
    """Edge maps over the edge states; a mortal type's dead agents add
    nothing."""
    X = ora.expr
    sim, A, E = EC.build(ora, n=500, mortal=True)
    sim.order = 0
    _to, _fr, es = sim._read_edges(E)
    es = np.asarray(es)
    for name, e in EC.EDGE_EXPRS(X).items():
        vals = EC.np_eval(e, es)
        dt = "f8" if EC._kind(vals) == "f" else "i8"
        assert _same(ora.mapreduce(sim, e, "+", E), EC.seq_fold(EC.to_dtype(vals, dt), "+", dt, _init("+", dt)))
    ora.apply_(sim, ora.transitions.die_if("odd"), [A], [A], [A])
    st, died = sim._read_agents(A)
    st = np.asarray(st)[~died]
    e = X.field("x") + X.field("i")
    want = EC.seq_fold(EC.np_eval(e, st), "+", "f8", 0.0)
    assert died.sum() > 100 and _same(ora.mapreduce(sim, e, "+", A), want)


def test_expr_params_and_errors(ora):
    """A param_ref operand takes the simulation's value; programs that do not
    fit (unknown field, too deep) are rejected."""
    X = ora.expr
    m, A, E = EC.model(ora)
    sim = ora.create_simulation(m, params={"lim": 10})
    st = EC.states(A, 200)
    sim._add_agents_array(A, st)
    ora.finish_init_(sim)
    got = ora.mapreduce(sim, X.field("i") > ora.param_ref("lim"), "+", A)
    assert got == int((st["i"] > 10).sum())
    with pytest.raises(AssertionError):
        ora.mapreduce(sim, X.field("nope") > 0, "+", A)
    big = X.field("i")
    for _ in range(64):
        big = big + 1
    with pytest.raises(AssertionError):
        ora.mapreduce(sim, big, "+", A)
    # the library checks raw programs itself
    assert_rejects_bad_programs(ora, sim, A)


def assert_rejects_bad_programs(V, sim, A):
    """Programs with a missing operand, a field outside the state, an
    unknown operation or two values left are VH_ERR_ARG."""
    import ctypes as C
    abi = V._abi
    EX = abi.EX

    def run(ops):
        arr = (abi.ExprOp * len(ops))(*[abi.ExprOp(c, d, a) for c, d, a in ops])
        out = C.c_uint64()
        return sim._mapreduce_expr(0, sim._aid(A), arr, len(ops), abi.OP["+"], abi.DT["i8"], None,
                                   C.cast(C.pointer(out), C.c_void_p))
    f8 = abi.DT["f8"]
    assert run([(EX["field"], f8, 8), (EX["+"], 0, 0)]) == abi.VH_ERR_ARG
    assert run([(EX["field"], f8, A.nbytes - 4)]) == abi.VH_ERR_ARG
    assert run([(EX["const_i"], 0, 1), (99, 0, 0)]) == abi.VH_ERR_ARG
    assert run([(EX["const_i"], 0, 1), (EX["const_i"], 0, 2)]) == abi.VH_ERR_ARG
    assert run([(EX["const_i"], 0, 1), (EX["const_i"], 0, 2), (EX["+"], 0, 0)]) == 0
