"""Shared cases for mapreduce with computed maps (vahana_amd.expr): a model
with fields of every mapped type, the expressions, and an independent numpy
evaluation of the documented semantics."""
import numpy as np

FIELDS = [("i", "<i8"), ("x", "<f8"), ("y", "<f4"), ("flag", "?"), ("u", "<u2"), ("k", "<i4")]


def model(V, mortal=False):
    A = V.struct("ExprAgent", FIELDS)
    E = V.struct("ExprEdge", [("w", "<f4"), ("c", "<i8")])
    m = (V.ModelTypes() | V.register_agenttype_(A, *(() if mortal else ("Immortal",)))
         | V.register_edgetype_(E) | V.create_model("expr_model"))
    return m, A, E


def states(A, n, seed=3):
    rng = np.random.default_rng(seed)
    st = np.zeros(n, dtype=A.dtype)
    st["i"] = rng.integers(-50, 50, n)
    st["x"] = rng.normal(0, 2, n)
    st["x"][::97] = -0.0
    st["x"][5::97] = 0.0
    st["y"] = rng.uniform(-1, 1, n).astype(np.float32)
    st["flag"] = rng.random(n) < 0.3
    st["u"] = rng.integers(0, 60000, n)
    st["k"] = rng.integers(-(2 ** 31), 2 ** 31 - 1, n)
    return st


def build(V, n=2000, mortal=False, seed=3):
    m, A, E = model(V, mortal)
    sim = V.create_simulation(m)
    st = states(A, n, seed)
    ids = sim._add_agents_array(A, st)
    rng = np.random.default_rng(seed + 1)
    ne = 3 * n
    fr = ids[rng.integers(0, n, ne)]
    to = ids[rng.integers(0, n, ne)]
    es = np.zeros(ne, dtype=E.dtype)
    es["w"] = rng.uniform(0, 1, ne).astype(np.float32)
    es["c"] = rng.integers(-5, 5, ne)
    sim._add_edges_array(E, fr, to, es)
    V.finish_init_(sim)
    return sim, A, E


def exprs(X):
    f, where = X.field, X.where
    return {
        "gt0": f("i") > 0,
        "lin": f("x") * 2 - f("i") / 3,
        "absdiff": abs(f("x") - 0.5),
        "f32": f("y") * f("y") + 1,
        "flag_and": f("flag") & (f("x") < 0.25),
        "not_or": ~f("flag") | (f("u") >= 30000),
        "select": where(f("flag"), f("x"), f("i")),
        "select_int": where(f("i") > 0, f("i"), -f("i")),
        "minmax": X.max_(X.min_(f("x"), f("y")), -1.5),
        "intwrap": f("k") * f("k") * f("k"),
        "negzero": -f("x") * 0.0,
        "eq_const": f("u") == 17,
        "int_ops": X.min_(f("i"), f("k")) - X.max_(f("u"), 3),
        "divzero": f("x") / (f("i") - f("i")),
        "ne_nan": (f("x") / (f("i") - f("i"))) != (f("x") / (f("i") - f("i"))),
    }


EDGE_EXPRS = lambda X: {"w_gt": X.field("w") > 0.5, "wc": X.field("w") * X.field("c")}  # noqa: E731


# ---- numpy evaluation (independent of both libraries)
def _kind(a):
    return "f" if a.dtype.kind == "f" else ("b" if a.dtype.kind == "b" else "i")


def _f(a):
    return a.astype(np.float64)


def _i(a):
    return a.astype(np.int64)


def _jl_min(a, b):
    r = np.where(a < b, a, b)
    eq = a == b
    r = np.where(eq, np.where(np.signbit(a), a, b), r)
    r = np.where(np.isnan(b), b, r)
    return np.where(np.isnan(a), a, r)


def _jl_max(a, b):
    r = np.where(a > b, a, b)
    eq = a == b
    r = np.where(eq, np.where(np.signbit(a), b, a), r)
    r = np.where(np.isnan(b), b, r)
    return np.where(np.isnan(a), a, r)


def _truth(a):
    return a != 0


def np_eval(e, st):
    with np.errstate(all="ignore"):
        return _ev(e, st)


def _ev(e, st):
    c = e.code
    n = len(st)
    if c == "field":
        a = st[e.arg_name]
        k = _kind(a)
        return _f(a) if k == "f" else (a.astype(bool) if k == "b" else _i(a))
    if c == "const_i":
        return np.full(n, e.value, dtype=np.int64)
    if c == "const_f":
        return np.full(n, float(e.value))
    if c == "const_b":
        return np.full(n, bool(e.value))
    args = [_ev(a, st) for a in e.args]
    if c in ("neg", "abs"):
        a = args[0]
        if _kind(a) == "f":
            return -a if c == "neg" else np.abs(a)
        a = _i(a)
        return -a if c == "neg" else np.where(a < 0, -a, a)
    if c == "not":
        return ~_truth(args[0])
    if c == "select":
        cc, a, b = args
        t = _truth(cc)
        if "f" in (_kind(a), _kind(b)):
            return np.where(t, _f(a), _f(b))
        if _kind(a) == _kind(b) == "b":
            return np.where(t, a, b)
        return np.where(t, _i(a), _i(b))
    a, b = args
    fl = "f" in (_kind(a), _kind(b))
    if c in ("==", "!=", "<", "<=", ">", ">="):
        x, y = (_f(a), _f(b)) if fl else (_i(a), _i(b))
        return {"==": x == y, "!=": x != y, "<": x < y, "<=": x <= y, ">": x > y, ">=": x >= y}[c]
    if c in ("and", "or"):
        return (_truth(a) & _truth(b)) if c == "and" else (_truth(a) | _truth(b))
    if fl or c == "/":
        x, y = _f(a), _f(b)
        return {"+": x + y, "-": x - y, "*": x * y, "/": x / y, "min": _jl_min(x, y), "max": _jl_max(x, y)}[c]
    x, y = _i(a), _i(b)
    return {"+": x + y, "-": x - y, "*": x * y, "min": np.minimum(x, y), "max": np.maximum(x, y)}[c]


def to_dtype(v, datatype):
    if datatype == "f8":
        return _f(v)
    if datatype == "b1":
        return _truth(v)
    if _kind(v) != "f":
        return _i(v)
    with np.errstate(all="ignore"):
        r = np.where(np.isnan(v), 0.0, v)
        big = r >= 9223372036854775808.0
        small = r < -9223372036854775808.0
        r = np.where(big | small, 0.0, r)
        out = np.trunc(r).astype(np.int64)
    out[big] = np.iinfo(np.int64).max
    out[small] = np.iinfo(np.int64).min
    return out


def seq_fold(vals, op, datatype, init):
    """The reference's fold: reduced = op(f(x), reduced) from val4empty."""
    acc = init
    for v in vals.tolist():
        if op == "+":
            acc = v + acc
        elif op == "*":
            acc = v * acc
        elif op == "max":
            acc = float(_jl_max(np.float64(v), np.float64(acc))) if datatype == "f8" else max(v, acc)
        elif op == "min":
            acc = float(_jl_min(np.float64(v), np.float64(acc))) if datatype == "f8" else min(v, acc)
        elif op == "&":
            acc = bool(v) and bool(acc) if datatype == "b1" else (v & acc)
        elif op == "|":
            acc = bool(v) or bool(acc) if datatype == "b1" else (v | acc)
        if datatype == "i8":
            acc = (acc + 2 ** 63) % 2 ** 64 - 2 ** 63
    return acc
