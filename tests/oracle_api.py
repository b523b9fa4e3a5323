"""Test-side binding of the CPU checker (oracle/build/liboracle.so).

Exposes the same Python API as the product module (`vahana_amd`): the
module-level functions are the product's (they dispatch on the simulation
object) while models and simulations created here are bound to the checker
library.  Only tests, smoke() and bench.py's CPU-baseline leg import this.
"""
import ctypes as C
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
from __graft_entry__ import load_package  # noqa: E402

vh = load_package()
_abi = vh._abi

ORACLE_SO = os.path.join(ROOT, "oracle", "build", "liboracle.so")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           stdout=subprocess.DEVNULL)
        L = _abi.Lib(ORACLE_SO, "or_", [
            ("reuse_stack", C.c_int, [C.c_void_p, C.c_uint8, C.POINTER(C.c_uint64), C.c_uint64,
                                      C.POINTER(C.c_uint64)]),
            ("hash4", C.c_uint64, [C.c_uint64] * 4),
            ("canon_sum_f64", None, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            ("register_body", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]),
        ])
        L.mapreduce.argtypes = list(L.mapreduce.argtypes) + [C.c_int]
        L.mapreduce_expr.argtypes = list(L.mapreduce_expr.argtypes) + [C.c_int]
        _LIB = L
    return _LIB


class OracleSim(vh.Simulation):
    """Simulation bound to the checker; `order` selects the mapreduce fold
    (0: the reference's sequential fold, 1: the device's fixed tree)."""

    order = 1

    def _mapreduce(self, kind, tid, md, op, dt, initp, outp):
        return self.lib.mapreduce(self.handle, kind, tid, C.byref(md), op, dt, initp, outp,
                                  self.order)

    def _mapreduce_expr(self, kind, tid, prog, nops, op, dt, initp, outp):
        return self.lib.mapreduce_expr(self.handle, kind, tid, prog, nops, op, dt, initp, outp, self.order)

    def reuse_stack(self, T):
        n = C.c_uint64()
        buf = (C.c_uint64 * 1)()
        self.lib.reuse_stack(self.handle, self._aid(T), buf, 0, C.byref(n))
        buf = (C.c_uint64 * max(n.value, 1))()
        self.lib.reuse_stack(self.handle, self._aid(T), buf, n.value, C.byref(n))
        return [buf[i] for i in range(n.value)]


def register_body(code_object, kernel, adds_edges=False, adds_agents=False):
    """The checker's counterpart of vh.register_body: the C restatement of
    device kernel `vh_<name>` is the checker function `or_<name>`
    (oracle/example_bodies.c), registered with or_register_body."""
    L = lib()
    name = "or_" + (kernel[3:] if kernel.startswith("vh_") else kernel)
    fn = C.cast(getattr(L.dll, name), C.c_void_p)
    flags = (vh._abi.BODY_ADDS_EDGES if adds_edges else 0) | (vh._abi.BODY_ADDS_AGENTS if adds_agents else 0)
    tid = C.c_uint32()
    rc = L.register_body(fn, flags, C.byref(tid))
    if rc != 0:
        raise vh.VahanaError(rc, L.last_error().decode())
    return vh.transitions.UserBody(tid.value, kernel, flags)


example_bodies = vh.example_bodies


def create_model(name):
    def step(types):
        return vh.Model(types, name, lib())
    return step


def create_simulation(model, params=None, globals_=None, name=None):
    return OracleSim(model, name, params=params, globals_=globals_)


# the product's API functions dispatch on the simulation object
struct = vh.struct
Edge = vh.Edge
ModelTypes = vh.ModelTypes
register_agenttype_ = vh.register_agenttype_
register_edgetype_ = vh.register_edgetype_
transitions = vh.transitions
raster = vh.raster
models = vh.models
inputs = vh.inputs
snapshot = vh.snapshot
expr = vh.expr
for _name in ["add_agent_", "add_agents_", "add_edge_", "add_edges_", "finish_init_", "apply_",
              "apply", "mapreduce", "num_agents", "num_edges", "all_agents", "all_agentids",
              "agentstate", "edges", "neighborids", "edgestates", "has_edge", "copy_simulation",
              "finish_simulation_", "agent_id", "agent_nr", "type_nr", "Field", "register_param_",
              "register_global_", "param", "set_param_", "get_global", "set_global_", "modify_global_",
              "push_global_", "Record", "param_ref", "global_ref"]:
    globals()[_name] = getattr(vh, _name)
