"""vh_mapreduce_expr on the device: bit-identical to the checker's tree
fold (the device's reduction order) for every expression, reduction and
record kind; invalid programs rejected alike."""
import math

import numpy as np
import pytest

import expr_cases as EC
from test_mapreduce_expr import assert_rejects_bad_programs

pytestmark = pytest.mark.gpu


def _same(a, b):
    if isinstance(a, float) and isinstance(b, float):
        return (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) == math.copysign(1, b))
    return a == b


def test_expr_device_equals_checker(vh, ora, gpu_ready):
    print("[gpu] mapreduce expressions, device vs checker", flush=True)
    g, A, E = EC.build(vh, n=70_000)           # several reduction blocks and levels
    c, A2, E2 = EC.build(ora, n=70_000)
    c.order = 1
    for name in EC.exprs(vh.expr):
        for op in ("+", "max", "min", "&", "|"):
            ge, ce = EC.exprs(vh.expr)[name], EC.exprs(ora.expr)[name]
            kinds = EC._kind(EC.np_eval(ge, np.asarray(g._read_agents(A)[0])[:4]))
            if op in ("&", "|") and kinds == "f":
                continue
            dts = ["f8"] if kinds == "f" else (["b1", "i8"] if kinds == "b" else ["i8"])
            for dt in dts:
                a = vh.mapreduce(g, ge, op, A, datatype=dt)
                b = ora.mapreduce(c, ce, op, A2, datatype=dt)
                assert _same(a, b), (name, op, dt, a, b)
    for name in EC.EDGE_EXPRS(vh.expr):
        a = vh.mapreduce(g, EC.EDGE_EXPRS(vh.expr)[name], "+", E)
        b = ora.mapreduce(c, EC.EDGE_EXPRS(ora.expr)[name], "+", E2)
        assert _same(a, b), name
    assert_rejects_bad_programs(vh, g, A)


def test_expr_device_mortal(vh, ora, gpu_ready):
    print("[gpu] mapreduce expressions over a mortal type with deaths", flush=True)
    g, A, E = EC.build(vh, n=20_000, mortal=True, seed=8)
    c, A2, E2 = EC.build(ora, n=20_000, mortal=True, seed=8)
    c.order = 1
    vh.apply_(g, vh.transitions.die_if("odd"), [A], [A], [A])
    ora.apply_(c, ora.transitions.die_if("odd"), [A2], [A2], [A2])
    for name in ("lin", "select", "flag_and", "int_ops"):
        a = vh.mapreduce(g, EC.exprs(vh.expr)[name], "+", A)
        b = ora.mapreduce(c, EC.exprs(ora.expr)[name], "+", A2)
        assert _same(a, b), name
