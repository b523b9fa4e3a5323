"""ctypes declarations of include/vahana_hip.h (the C ABI of libvahana_hip.so)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvahana_hip.so")

VH_OK = 0
VH_ERR_ASSERT = -1
VH_ERR_UNREGISTERED = -2
VH_ERR_OOM = -3
VH_ERR_BUDGET = -4
VH_ERR_DEVICE = -5
VH_ERR_CAPACITY = -6
VH_ERR_ARG = -7

VH_SHIFT_TYPE = 56
VH_NR_MASK = (1 << 36) - 1
MAX_AGENT_TYPES = 8
MAX_EDGE_TYPES = 24
MAX_CALL = 8

HINT_IMMORTAL = 1 << 0
HINT_INDEPENDENT = 1 << 1
HINT_STATELESS = 1 << 0
HINT_IGNOREFROM = 1 << 1
HINT_SINGLEEDGE = 1 << 2
HINT_SINGLETYPE = 1 << 3
HINT_IGNORESOURCESTATE = 1 << 4
HINT_NUMEDGESONLY = 1 << 5
HINT_HASEDGEONLY = 1 << 6

AGENT_HINTS = {"Immortal": HINT_IMMORTAL, "Independent": HINT_INDEPENDENT}
EDGE_HINTS = {"Stateless": HINT_STATELESS, "IgnoreFrom": HINT_IGNOREFROM,
              "SingleEdge": HINT_SINGLEEDGE, "SingleType": HINT_SINGLETYPE,
              "IgnoreSourceState": HINT_IGNORESOURCESTATE,
              "NumEdgesOnly": HINT_NUMEDGESONLY, "HasEdgeOnly": HINT_HASEDGEONLY}

TR_NOOP, TR_SUM_NEIGHBORS, TR_DIE_IF, TR_ADD_SELF_EDGE = 0, 1, 2, 3
TR_READD_EDGES, TR_NEIGHBOR_FOO_EDGES, TR_SPAWN, TR_COUNT_EDGES = 4, 5, 6, 7
TR_DIFFUSE, TR_MOVE_TO = 8, 9
TR_REMOVE_EDGES = 10
# enum vh_remove_mode
RM = {"own": 0, "neighbor": 1, "readd": 2, "from_two": 3, "to_neighbor": 4, "send_id": 5,
      "to_sendid": 6, "zero_from": 7}
TR_SIR, TR_OPINION_F64, TR_OPINION_F32, TR_CONTACT = 16, 17, 18, 19

CMP = {"always": 0, "never": 1, "lt": 2, "ge": 3, "eq": 4, "odd": 5, "even": 6,
       "num_edges_eq": 7}

OP = {"+": 0, "*": 1, "max": 2, "min": 3, "&": 4, "|": 5}
DT = {"i8": 0, "f8": 1, "b1": 2, "f4": 3, "u1": 4, "u2": 5, "u4": 6, "i4": 7, "u8": 8}
MAP_FIELD, MAP_FIELD_EQ, MAP_ONE = 0, 1, 2

S, I, R = 0, 1, 2


class ValueBind(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("kind", C.c_uint8), ("dtype", C.c_uint8),
                ("slot", C.c_uint16)]


VALUE_PARAM, VALUE_GLOBAL = 0, 1
MAX_VALUES = 32


class ApplyDesc(C.Structure):
    _fields_ = [("transition", C.c_uint32), ("n_call", C.c_uint32),
                ("call", C.c_uint8 * MAX_CALL),
                ("read_agents", C.c_uint32), ("read_edges", C.c_uint32),
                ("write_agents", C.c_uint32), ("write_edges", C.c_uint32),
                ("add_existing_agents", C.c_uint32), ("add_existing_edges", C.c_uint32),
                ("with_edge", C.c_int32),
                ("role_agent", C.c_int8 * 4), ("role_edge", C.c_int8 * 4),
                ("params", C.c_void_p), ("params_bytes", C.c_uint32), ("n_bind", C.c_uint32),
                ("bind", C.POINTER(ValueBind))]


class Options(C.Structure):
    _fields_ = [("opinion_plan_bytes", C.c_int64), ("opinion_part_bytes", C.c_int64),
                ("sir_expose", C.c_int32), ("pad", C.c_uint32), ("sir_expose_min", C.c_int64)]


class MapDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("offset", C.c_uint32), ("field_dtype", C.c_uint32),
                ("pad", C.c_uint32), ("value", C.c_int64)]


class ExprOp(C.Structure):
    _fields_ = [("code", C.c_uint32), ("dtype", C.c_uint32), ("arg", C.c_uint64)]


EX = {"field": 0, "const_i": 1, "const_f": 2, "const_b": 3, "+": 10, "-": 11, "*": 12, "/": 13,
      "min": 14, "max": 15, "neg": 20, "abs": 21, "not": 22, "==": 30, "!=": 31, "<": 32, "<=": 33,
      ">": 34, ">=": 35, "and": 40, "or": 41, "select": 50}
EXPR_MAX_OPS = 64
TR_USER = 64
BODY_ADDS_EDGES = 1
BODY_ADDS_AGENTS = 2
EXPR_MAX_DEPTH = 16


class TestParams(C.Structure):
    _fields_ = [("value", C.c_int64), ("cmp", C.c_int32), ("pad", C.c_uint32),
                ("edge_state", C.c_uint8 * 16)]


class RasterParams(C.Structure):
    _fields_ = [("field_offset", C.c_uint32), ("cell_type", C.c_uint32), ("first_nr", C.c_uint64),
                ("dims", C.c_int64 * 4), ("ndims", C.c_uint32), ("flags", C.c_uint32),
                ("edge_state", C.c_uint8 * 16)]


class SirParams(C.Structure):
    _fields_ = [("p_inf", C.c_double), ("seed", C.c_uint64), ("recovery", C.c_uint32),
                ("pad", C.c_uint32)]


class OpinionParams(C.Structure):
    _fields_ = [("eps", C.c_double)]


class ContactParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("gamma", C.c_double), ("p_death", C.c_double),
                ("seed", C.c_uint64), ("n_contacts", C.c_uint32), ("max_redraw", C.c_uint32)]


class SirMember(C.Structure):
    _fields_ = [("n", C.c_uint64), ("mean_deg", C.c_double), ("p_inf", C.c_double),
                ("seed", C.c_uint64), ("graph_seed", C.c_uint64), ("recovery", C.c_uint32),
                ("n_init", C.c_uint32)]


_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p

# (name, restype, argtypes) with the `vh_` prefix stripped
SIGNATURES = [
    ("model_create", C.c_int, [C.POINTER(_vp)]),
    ("model_destroy", C.c_int, [_vp]),
    ("register_agenttype", C.c_int, [_vp, C.c_uint32, C.c_uint32, _u8p]),
    ("register_edgetype", C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int32, C.c_int64, _u8p]),
    ("edgetype_hints", C.c_int, [_vp, C.c_uint8, _u32p]),
    ("sim_create", C.c_int, [_vp, C.POINTER(_vp)]),
    ("sim_destroy", C.c_int, [_vp]),
    ("sim_copy", C.c_int, [_vp, C.POINTER(_vp)]),
    ("restore_died", C.c_int, [_vp, C.c_uint8, _u8p, C.c_uint64]),
    ("set_num_transitions", C.c_int, [_vp, C.c_uint64]),
    ("last_change", C.c_int, [_vp, C.c_int, C.c_uint8, _u64p]),
    ("set_last_change", C.c_int, [_vp, C.c_int, C.c_uint8, C.c_uint64]),
    ("add_agents", C.c_int, [_vp, C.c_uint8, _vp, C.c_uint64, _u64p]),
    ("add_edges", C.c_int, [_vp, C.c_uint8, _u64p, _u64p, _vp, C.c_uint64]),
    ("finish_init", C.c_int, [_vp]),
    ("apply", C.c_int, [_vp, C.POINTER(ApplyDesc)]),
    ("num_transitions", C.c_int, [_vp, _u64p]),
    ("sim_set_value", C.c_int, [_vp, C.c_int, C.c_uint32, C.c_int, _vp]),
    ("sim_get_value", C.c_int, [_vp, C.c_int, C.c_uint32, C.POINTER(C.c_int), _vp]),
    ("num_agents", C.c_int, [_vp, C.c_uint8, _u64p]),
    ("num_edges", C.c_int, [_vp, C.c_uint8, _u64p]),
    ("read_agents", C.c_int, [_vp, C.c_uint8, _vp, _u8p, C.c_uint64, _u64p]),
    ("read_edges", C.c_int, [_vp, C.c_uint8, _u64p, _u64p, _vp, C.c_uint64, _u64p]),
    ("edges_of", C.c_int, [_vp, C.c_uint8, C.c_uint64, _u64p, _vp, C.c_uint64, _u64p]),
    ("agentstate", C.c_int, [_vp, C.c_uint64, C.c_uint8, _vp]),
]
DEVICE_ONLY = [
    ("apply_n", C.c_int, [_vp, C.POINTER(ApplyDesc), C.c_uint32]),
    ("add_edges_gnp", C.c_int, [_vp, C.c_uint8, C.c_double, C.c_uint64, C.c_uint64]),
    ("add_edges_ba", C.c_int, [_vp, C.c_uint8, C.c_uint32, C.c_uint64, C.c_uint64]),
    ("register_body", C.c_int, [C.c_char_p, C.c_char_p, C.c_uint32, _u32p]),
    ("finish_init_relabel", C.c_int, [_vp, C.c_uint8, _u64p]),
    ("add_edges_stencil", C.c_int, [_vp, C.c_uint8, C.c_uint32, C.c_uint64, C.c_uint32,
                                    C.POINTER(C.c_int64), C.c_double, C.c_int, C.c_int,
                                    C.POINTER(C.c_uint8)]),
    ("sync", C.c_int, [_vp]),
    ("opinion_prepare", C.c_int, [_vp, C.c_uint8, C.c_uint8, C.c_int]),
    ("set_stream", C.c_int, [_vp, _vp]),
    ("set_io_stream", C.c_int, [_vp, _vp]),
    ("stream_create", C.c_int, [C.POINTER(_vp)]),
    ("stream_destroy", C.c_int, [_vp]),
    ("device_bytes", C.c_int, [_vp, _u64p]),
    ("enable_asserts", C.c_int, [C.c_int]),
    ("get_options", C.c_int, [C.POINTER(Options)]),
    ("set_options", C.c_int, [C.POINTER(Options)]),
    ("debug_fail_allocs", C.c_int, [C.c_uint32]),
    ("debug_set_device_error", C.c_int, [_vp, C.c_uint32]),
    ("graph_er", C.c_int, [C.c_uint64, C.c_double, C.c_uint64, _u32p, _u32p, C.c_uint64, _u64p]),
    ("graph_ba", C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_uint64, _u64p]),
    ("hash4", C.c_uint64, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]),
    ("fill_open01f", C.c_int, [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                               C.POINTER(C.c_float)]),
    ("bfs_order", C.c_int, [C.c_uint64, C.c_uint64, _u32p, _u32p, _u64p]),
    ("fill_agent_u01", C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64,
                                 C.POINTER(C.c_double)]),
    ("graph_gnp", C.c_int, [C.c_uint64, C.c_double, C.c_uint64, _u32p, _u32p, C.c_uint64, _u64p]),
    ("ensemble_create", C.c_int, [C.POINTER(_vp)]),
    ("ensemble_destroy", C.c_int, [_vp]),
    ("ensemble_add_sir", C.c_int, [_vp, C.POINTER(SirMember), _u32p]),
    ("ensemble_remove", C.c_int, [_vp, C.c_uint32]),
    ("ensemble_set_horizon", C.c_int, [_vp, C.c_uint32, C.c_uint64]),
    ("ensemble_set_config", C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    ("ensemble_step", C.c_int, [_vp, C.c_uint32, C.c_int]),
    ("ensemble_sync", C.c_int, [_vp]),
    ("ensemble_set_streams", C.c_int, [_vp, _vp, _vp]),
    ("ensemble_read", C.c_int, [_vp, C.c_uint32, _vp, C.c_uint64, _u64p, _u64p]),
    ("ensemble_info", C.c_int, [_vp, _u64p, _u64p, _u64p, _u64p]),
    ("ensemble_autotune", C.c_int, [_vp, C.c_uint32, _u32p, _u32p, _u32p, C.POINTER(C.c_float)]),
    ("ensemble_set_tiles", C.c_int, [_vp, _u32p]),
    ("batch_create", C.c_int, [C.POINTER(_vp)]),
    ("batch_destroy", C.c_int, [_vp]),
    ("batch_apply", C.c_int, [_vp, C.POINTER(_vp), C.c_uint32, C.POINTER(ApplyDesc), _u32p]),
    ("mapreduce_noalloc", C.c_int, [_vp, C.c_int, C.c_uint8, C.POINTER(MapDesc), C.c_int, C.c_int, _vp,
                                    _vp]),
]
MAPREDUCE = ("mapreduce", C.c_int, [_vp, C.c_int, C.c_uint8, C.POINTER(MapDesc), C.c_int,
                                     C.c_int, _vp, _vp])
MAPREDUCE_EXPR = ("mapreduce_expr", C.c_int, [_vp, C.c_int, C.c_uint8, C.POINTER(ExprOp), C.c_uint32,
                                              C.c_int, C.c_int, _vp, _vp])


class Lib:
    """Bound entry points of one shared library with a name prefix."""

    def __init__(self, path, prefix, extra=()):
        if not os.path.exists(path):
            raise ImportError(
                f"{os.path.basename(path)} is not built ({path}); run "
                "`python vahana.jl_amd/build.py` (or __graft_entry__.build())")
        self.path = path
        self.dll = C.CDLL(path)
        self.prefix = prefix
        for name, res, args in list(SIGNATURES) + [MAPREDUCE, MAPREDUCE_EXPR] + list(extra):
            fn = getattr(self.dll, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        le = getattr(self.dll, prefix + "last_error")
        le.restype = C.c_char_p
        le.argtypes = []
        self.last_error = le


_LIB = None


def lib():
    """The product library (libvahana_hip.so); raises if it is not built."""
    global _LIB
    if _LIB is None:
        _LIB = Lib(LIB_PATH, "vh_", DEVICE_ONLY)
    return _LIB
