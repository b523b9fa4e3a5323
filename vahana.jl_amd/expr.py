"""Computed maps for mapreduce: the function f of mapreduce(sim, f, op, T)
(src/AgentMethods.jl:533-565, src/EdgeMethods.jl:972-1001) written as an
expression over the state's fields, compiled to the postfix program that
vh_mapreduce_expr evaluates per record on the device.

    from vahana_amd.expr import field, where
    mapreduce(sim, field("energy") > 0, "+", Prey)           # a -> a.energy > 0
    mapreduce(sim, abs(field("x") - 0.5), "max", HK)         # a -> abs(a.x - 0.5)
    mapreduce(sim, where(field("alive"), field("w"), 0.0), "+", T)

Values are Int64, Float64 or Bool with Julia's promotions (an operation
with a float operand is Float64, `/` is always Float64, comparisons and
`&` / `|` / `~` give Bool). A param_ref / global_ref operand is replaced by
the simulation's value when mapreduce is called.
"""
import struct as _struct

from . import _abi


class Expr:
    """A node of a map expression; build with field(), constants and
    operators."""

    def __init__(self, code, args=(), dtype=0, arg=0, kind=None):
        self.code = code
        self.args = tuple(args)
        self.dtype = dtype
        self.arg = arg
        self.kind = kind        # 'i', 'f' or 'b': the value's type

    # arithmetic
    def __add__(self, o): return _bin("+", self, o)
    def __radd__(self, o): return _bin("+", o, self)
    def __sub__(self, o): return _bin("-", self, o)
    def __rsub__(self, o): return _bin("-", o, self)
    def __mul__(self, o): return _bin("*", self, o)
    def __rmul__(self, o): return _bin("*", o, self)
    def __truediv__(self, o): return _bin("/", self, o)
    def __rtruediv__(self, o): return _bin("/", o, self)
    def __neg__(self): return Expr("neg", (self,), kind="f" if self.kind == "f" else "i")
    def __abs__(self): return Expr("abs", (self,), kind="f" if self.kind == "f" else "i")

    # comparisons (Bool)
    def __eq__(self, o): return _bin("==", self, o)
    def __ne__(self, o): return _bin("!=", self, o)
    def __lt__(self, o): return _bin("<", self, o)
    def __le__(self, o): return _bin("<=", self, o)
    def __gt__(self, o): return _bin(">", self, o)
    def __ge__(self, o): return _bin(">=", self, o)

    # logical (Bool)
    def __and__(self, o): return _bin("and", self, o)
    def __rand__(self, o): return _bin("and", o, self)
    def __or__(self, o): return _bin("or", self, o)
    def __ror__(self, o): return _bin("or", o, self)
    def __invert__(self): return Expr("not", (self,), kind="b")

    __hash__ = object.__hash__

    def __bool__(self):
        raise TypeError("a map expression has no truth value on the host; use & | ~ and where()")

    def __repr__(self):
        if self.code == "field":
            return f"a.{self.arg_name}"
        if self.code.startswith("const"):
            return repr(self.value)
        if self.code == "ref":
            return repr(self.ref)
        return f"{self.code}({', '.join(map(repr, self.args))})"


def _const(v):
    from .transitions import ValueRef
    if isinstance(v, Expr):
        return v
    if isinstance(v, ValueRef):
        e = Expr("ref", kind=None)
        e.ref = v
        return e
    if isinstance(v, bool):
        e = Expr("const_b", kind="b", arg=int(v))
    elif isinstance(v, int):
        e = Expr("const_i", kind="i", arg=v & 0xFFFFFFFFFFFFFFFF)
    else:
        e = Expr("const_f", kind="f", arg=_struct.unpack("<Q", _struct.pack("<d", float(v)))[0])
    e.value = v
    return e


def _bin(op, a, b):
    a, b = _const(a), _const(b)
    if op in ("==", "!=", "<", "<=", ">", ">=", "and", "or"):
        kind = "b"
    elif op == "/" or "f" in (a.kind, b.kind):
        kind = "f"
    elif a.kind is None or b.kind is None:
        kind = None             # decided by a value reference at call time
    else:
        kind = "i"
    return Expr(op, (a, b), kind=kind)


def field(name):
    """a.name: a field of the state (floats as Float64, Bool as Bool, other
    integers as Int64)."""
    e = Expr("field")
    e.arg_name = str(name).lstrip(":")
    return e


def const(v):
    """A constant (Int64 for int, Float64 for float, Bool for bool)."""
    return _const(v)


def where(c, a, b):
    """ifelse(c, a, b)."""
    c, a, b = _const(c), _const(a), _const(b)
    kind = "f" if "f" in (a.kind, b.kind) else ("b" if a.kind == b.kind == "b" else "i")
    return Expr("select", (c, a, b), kind=kind)


def min_(a, b):
    """min(a, b) (Julia's on floats: NaN wins, -0.0 < +0.0)."""
    a, b = _const(a), _const(b)
    return Expr("min", (a, b), kind="f" if "f" in (a.kind, b.kind) else "i")


def max_(a, b):
    """max(a, b)."""
    a, b = _const(a), _const(b)
    return Expr("max", (a, b), kind="f" if "f" in (a.kind, b.kind) else "i")


_FIELD_DT = {"i8": "i8", "f8": "f8", "b1": "b1", "f4": "f4", "u1": "u1", "u2": "u2", "u4": "u4",
             "i4": "i4", "u8": "u8"}


def compile_expr(e, T, sim=None):
    """(ctypes array of vh_expr_op, count, result kind 'i' / 'f' / 'b') of
    expression `e` over the states of struct type T; value references are
    read from `sim`."""
    ops = []
    depth = [0, 0]

    def kind_of(x):
        if x.code == "field":
            fdt = T.dtype.fields[x.arg_name][0]
            return "f" if fdt.kind == "f" else ("b" if fdt.kind == "b" else "i")
        if x.code == "ref":
            return _ref_const(x).kind
        if x.kind is not None:
            return x.kind
        ks = [kind_of(a) for a in x.args]
        if x.code in ("neg", "abs"):
            return "f" if ks[0] == "f" else "i"
        if x.code == "select":
            return "f" if "f" in ks[1:] else ("b" if ks[1] == ks[2] == "b" else "i")
        return "f" if "f" in ks else "i"

    def _ref_const(x):
        from . import _getf
        if sim is None:
            raise AssertionError(f"{x.ref!r} needs the simulation")
        obj = sim.params if x.ref.kind == _abi.VALUE_PARAM else sim.globals
        return _const(_getf(obj, x.ref.name))

    def emit(x):
        if x.code == "ref":
            return emit(_ref_const(x))
        for a in x.args:
            emit(a)
        op = _abi.ExprOp()
        op.code = _abi.EX[x.code]
        if x.code == "field":
            if T.dtype.names is None or x.arg_name not in T.dtype.names:
                raise AssertionError(f"{T.name} has no field {x.arg_name}")
            fdt, off = T.dtype.fields[x.arg_name][:2]
            code = "b1" if fdt.kind == "b" else fdt.kind + str(fdt.itemsize)
            if code not in _FIELD_DT:
                raise AssertionError(f"field {x.arg_name}: type {fdt} can not be mapped")
            op.dtype = _abi.DT[_FIELD_DT[code]]
            op.arg = off
        elif x.code.startswith("const"):
            op.arg = x.arg
        ops.append(op)
        pops = {"field": 0, "const_i": 0, "const_f": 0, "const_b": 0, "neg": 1, "abs": 1, "not": 1,
                "select": 3}.get(x.code, 2)
        depth[0] += 1 - pops
        depth[1] = max(depth[1], depth[0])

    emit(e)
    if len(ops) > _abi.EXPR_MAX_OPS or depth[1] > _abi.EXPR_MAX_DEPTH:
        raise AssertionError(f"map expression too large ({len(ops)} operations, depth {depth[1]})")
    arr = (_abi.ExprOp * len(ops))(*ops)
    return arr, len(ops), kind_of(e)
