// vh_query.hip — mapreduce, counts, read-back accessors and copy_simulation.
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "vh_engine.h"

using namespace vh;

static inline bool has(uint32_t h, uint32_t x) { return (h & x) != 0; }

// ================================================================ mapreduce
// Fixed reduction tree shared with the checker: blocks of 2048 values; thread
// t of 256 folds v[t + 256 j] (j = 0..7) starting from the identity, then
// a[t] = a[t] op a[t + s] for s = 128, 64, ..., 1; block results are reduced
// by the same kernel until one value remains.  Removed agents are the
// identity at their slot.
namespace {

struct MapSpec {
    uint32_t kind, offset, fdt, pad;
    int64_t value;
    int32_t op, dtype;
};

__device__ __forceinline__ bool is_f(int dt) { return dt == VH_DT_F64 || dt == VH_DT_F32; }

__device__ __forceinline__ uint64_t to_bits(double f) { return (uint64_t)__double_as_longlong(f); }
__device__ __forceinline__ double from_bits(uint64_t b) { return __longlong_as_double((long long)b); }

__device__ uint64_t identity(int op, int dt) {
    bool f = is_f(dt);
    switch (op) {
    case VH_OP_ADD: case VH_OP_OR: return f ? to_bits(0.0) : 0ull;
    case VH_OP_MUL: return f ? to_bits(1.0) : 1ull;
    case VH_OP_MAX: return f ? to_bits(-__builtin_inf()) : (uint64_t)INT64_MIN;
    case VH_OP_MIN: return f ? to_bits(__builtin_inf()) : (uint64_t)INT64_MAX;
    case VH_OP_AND: return dt == VH_DT_BOOL ? 1ull : ~0ull;
    }
    return 0;
}
// Julia max/min on floats: NaN wins, -0.0 < +0.0
__device__ double jl_fmax(double a, double b) {
    if (a != a || b != b) return a != a ? a : b;
    if (a == b) return signbit(a) ? b : a;
    return a > b ? a : b;
}
__device__ double jl_fmin(double a, double b) {
    if (a != a || b != b) return a != a ? a : b;
    if (a == b) return signbit(a) ? a : b;
    return a < b ? a : b;
}
__device__ uint64_t op_apply(int op, int dt, uint64_t x, uint64_t y) {
    if (is_f(dt)) {
        double a = from_bits(x), b = from_bits(y), r = 0.0;
        switch (op) {
        case VH_OP_ADD: r = a + b; break;
        case VH_OP_MUL: r = a * b; break;
        case VH_OP_MAX: r = jl_fmax(a, b); break;
        case VH_OP_MIN: r = jl_fmin(a, b); break;
        }
        return to_bits(r);
    }
    int64_t a = (int64_t)x, b = (int64_t)y;
    uint64_t r = 0;
    switch (op) {
    case VH_OP_ADD: r = x + y; break;
    case VH_OP_MUL: r = x * y; break;
    case VH_OP_MAX: r = (uint64_t)(a > b ? a : b); break;
    case VH_OP_MIN: r = (uint64_t)(a < b ? a : b); break;
    case VH_OP_AND: r = x & y; break;
    case VH_OP_OR: r = x | y; break;
    }
    if (dt == VH_DT_BOOL) r = r != 0;
    return r;
}
__device__ double field_f(const uint8_t* p, uint32_t dt) {
    if (dt == VH_DT_F64) { double v; memcpy(&v, p, 8); return v; }
    float v; memcpy(&v, p, 4); return v;
}
__device__ int64_t field_i(const uint8_t* p, uint32_t dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_U64: { int64_t v; memcpy(&v, p, 8); return v; }
    case VH_DT_I32: { int32_t v; memcpy(&v, p, 4); return v; }
    case VH_DT_U32: { uint32_t v; memcpy(&v, p, 4); return v; }
    case VH_DT_U16: { uint16_t v; memcpy(&v, p, 2); return v; }
    case VH_DT_U8: case VH_DT_BOOL: return p[0];
    case VH_DT_F64: case VH_DT_F32: return (int64_t)field_f(p, dt);
    }
    return 0;
}
__device__ uint64_t map_val(const MapSpec& m, const uint8_t* st) {
    bool f = is_f(m.dtype);
    if (m.kind == VH_MAP_ONE) return f ? to_bits(1.0) : 1ull;
    const uint8_t* p = st + m.offset;
    if (m.kind == VH_MAP_FIELD_EQ) {
        bool eq = is_f(m.fdt) ? field_f(p, m.fdt) == (double)m.value : field_i(p, m.fdt) == m.value;
        return f ? to_bits(eq ? 1.0 : 0.0) : (uint64_t)eq;
    }
    if (f) return to_bits(is_f(m.fdt) ? field_f(p, m.fdt) : (double)field_i(p, m.fdt));
    return (uint64_t)field_i(p, m.fdt);
}

constexpr int MB = 256, MI = 8, MTILE = MB * MI;

__device__ uint64_t tree_block(uint64_t x, uint64_t* lds, int op, int dt) {
    lds[threadIdx.x] = x;
    __syncthreads();
    for (int s = MB / 2; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s) lds[threadIdx.x] = op_apply(op, dt, lds[threadIdx.x], lds[threadIdx.x + s]);
        __syncthreads();
    }
    return lds[0];
}

// level 0 over records (agent states with died flags, or edge states)
__global__ __launch_bounds__(MB) void k_mr_records(const uint8_t* __restrict__ st, uint32_t bytes,
                                                   const uint8_t* __restrict__ died, uint64_t n,
                                                   MapSpec ms, uint64_t* __restrict__ out) {
    __shared__ uint64_t lds[MB];
    uint64_t id = identity(ms.op, ms.dtype);
    uint64_t acc = id;
    for (int j = 0; j < MI; j++) {
        uint64_t e = (uint64_t)blockIdx.x * MTILE + threadIdx.x + (uint64_t)MB * j;
        if (e < n) {
            uint64_t v = (died && died[e]) ? id : map_val(ms, st + e * bytes);
            acc = op_apply(ms.op, ms.dtype, acc, v);
        }
    }
    uint64_t r = tree_block(acc, lds, ms.op, ms.dtype);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}
__global__ __launch_bounds__(MB) void k_mr_values(const uint64_t* __restrict__ in, uint64_t n, int op,
                                                  int dt, uint64_t* __restrict__ out) {
    __shared__ uint64_t lds[MB];
    uint64_t acc = identity(op, dt);
    for (int j = 0; j < MI; j++) {
        uint64_t e = (uint64_t)blockIdx.x * MTILE + threadIdx.x + (uint64_t)MB * j;
        if (e < n) acc = op_apply(op, dt, acc, in[e]);
    }
    uint64_t r = tree_block(acc, lds, op, dt);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}
__global__ void k_identity(uint64_t* out, int op, int dt) { *out = identity(op, dt); }

// ---- computed maps (vh_mapreduce_expr): a postfix program per record;
// values are 8 bits + a tag (0 Int64, 1 Float64, 2 Bool)
struct ExprProg {
    vh_expr_op op[VH_EXPR_MAX_OPS];
    uint32_t n;
};
constexpr uint8_t TI = 0, TF = 1, TB = 2;

// a float to Int64: truncated; NaN 0, out of range saturated
__device__ __forceinline__ int64_t f2i(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775808.0) return INT64_MAX;
    if (x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
__device__ __forceinline__ double ev_f(uint8_t tag, uint64_t v) {
    return tag == TF ? from_bits(v) : (double)(int64_t)v;
}
__device__ uint64_t expr_val(const ExprProg& p, const uint8_t* st, int dt, uint8_t& tag_out) {
    uint64_t sv[VH_EXPR_MAX_DEPTH];
    uint8_t sg[VH_EXPR_MAX_DEPTH];
    int sp = 0;
    for (uint32_t k = 0; k < p.n; k++) {
        const uint32_t c = p.op[k].code;
        const uint64_t arg = p.op[k].arg;
        if (c == VH_EX_FIELD) {
            const uint32_t fdt = p.op[k].dtype;
            const uint8_t* q = st + arg;
            if (is_f(fdt)) {
                sv[sp] = to_bits(field_f(q, fdt));
                sg[sp] = TF;
            } else {
                sv[sp] = (uint64_t)field_i(q, fdt);
                sg[sp] = fdt == VH_DT_BOOL ? TB : TI;
                if (fdt == VH_DT_BOOL) sv[sp] = sv[sp] != 0;
            }
            sp++;
        } else if (c == VH_EX_CONST_I || c == VH_EX_CONST_F || c == VH_EX_CONST_B) {
            sv[sp] = c == VH_EX_CONST_B ? (arg != 0) : arg;
            sg[sp] = c == VH_EX_CONST_I ? TI : (c == VH_EX_CONST_F ? TF : TB);
            sp++;
        } else if (c == VH_EX_NEG || c == VH_EX_ABS || c == VH_EX_NOT) {
            const uint64_t a = sv[sp - 1];
            const uint8_t g = sg[sp - 1];
            if (c == VH_EX_NOT) {
                sv[sp - 1] = g == TF ? (uint64_t)(from_bits(a) == 0.0) : (uint64_t)(a == 0);
                sg[sp - 1] = TB;
            } else if (g == TF) {
                const double x = from_bits(a);
                sv[sp - 1] = to_bits(c == VH_EX_NEG ? -x : fabs(x));
            } else {
                const int64_t x = (int64_t)a;
                sv[sp - 1] = (uint64_t)(c == VH_EX_NEG ? (int64_t)(0ull - (uint64_t)x) : (x < 0 ? (int64_t)(0ull - (uint64_t)x) : x));
                sg[sp - 1] = TI;
            }
        } else if (c == VH_EX_SELECT) {
            const uint64_t b = sv[sp - 1], a = sv[sp - 2], cc = sv[sp - 3];
            const uint8_t gb = sg[sp - 1], ga = sg[sp - 2], gc = sg[sp - 3];
            const bool t = gc == TF ? from_bits(cc) != 0.0 : cc != 0;
            sp -= 2;
            if (ga == TF || gb == TF) {
                sv[sp - 1] = to_bits(t ? ev_f(ga, a) : ev_f(gb, b));
                sg[sp - 1] = TF;
            } else {
                sv[sp - 1] = t ? a : b;
                sg[sp - 1] = (ga == TB && gb == TB) ? TB : TI;
            }
        } else {   // binary
            const uint64_t b = sv[sp - 1], a = sv[sp - 2];
            const uint8_t gb = sg[sp - 1], ga = sg[sp - 2];
            sp--;
            const bool fl = ga == TF || gb == TF;
            uint64_t r = 0;
            uint8_t g = TI;
            if (c >= VH_EX_EQ && c <= VH_EX_GE) {
                bool t = false;
                if (fl) {
                    const double x = ev_f(ga, a), y = ev_f(gb, b);
                    t = c == VH_EX_EQ ? x == y : c == VH_EX_NE ? x != y : c == VH_EX_LT ? x < y
                      : c == VH_EX_LE ? x <= y : c == VH_EX_GT ? x > y : x >= y;
                } else {
                    const int64_t x = (int64_t)a, y = (int64_t)b;
                    t = c == VH_EX_EQ ? x == y : c == VH_EX_NE ? x != y : c == VH_EX_LT ? x < y
                      : c == VH_EX_LE ? x <= y : c == VH_EX_GT ? x > y : x >= y;
                }
                r = t;
                g = TB;
            } else if (c == VH_EX_AND || c == VH_EX_OR) {
                const bool x = ga == TF ? from_bits(a) != 0.0 : a != 0;
                const bool y = gb == TF ? from_bits(b) != 0.0 : b != 0;
                r = c == VH_EX_AND ? (x && y) : (x || y);
                g = TB;
            } else if (fl || c == VH_EX_DIV) {
                const double x = ev_f(ga, a), y = ev_f(gb, b);
                double z = 0.0;
                switch (c) {
                case VH_EX_ADD: z = x + y; break;
                case VH_EX_SUB: z = x - y; break;
                case VH_EX_MUL: z = x * y; break;
                case VH_EX_DIV: z = x / y; break;
                case VH_EX_MIN: z = jl_fmin(x, y); break;
                case VH_EX_MAX: z = jl_fmax(x, y); break;
                }
                r = to_bits(z);
                g = TF;
            } else {
                const int64_t x = (int64_t)a, y = (int64_t)b;
                switch (c) {
                case VH_EX_ADD: r = a + b; break;
                case VH_EX_SUB: r = a - b; break;
                case VH_EX_MUL: r = a * b; break;
                case VH_EX_MIN: r = (uint64_t)(x < y ? x : y); break;
                case VH_EX_MAX: r = (uint64_t)(x > y ? x : y); break;
                }
                g = TI;
            }
            sv[sp - 1] = r;
            sg[sp - 1] = g;
        }
    }
    tag_out = sg[0];
    const uint64_t v = sv[0];
    const uint8_t g = sg[0];
    if (is_f(dt)) return to_bits(ev_f(g, v));
    if (dt == VH_DT_BOOL) return g == TF ? (uint64_t)(from_bits(v) != 0.0) : (uint64_t)(v != 0);
    return g == TF ? (uint64_t)f2i(from_bits(v)) : v;
}

__global__ __launch_bounds__(MB) void k_mr_expr(const uint8_t* __restrict__ st, uint32_t bytes,
                                                const uint8_t* __restrict__ died, uint64_t n, ExprProg prog,
                                                int op, int dt, uint64_t* __restrict__ out) {
    __shared__ uint64_t lds[MB];
    const uint64_t id = identity(op, dt);
    uint64_t acc = id;
    for (int j = 0; j < MI; j++) {
        const uint64_t e = (uint64_t)blockIdx.x * MTILE + threadIdx.x + (uint64_t)MB * j;
        if (e < n) {
            uint8_t tg;
            const uint64_t v = (died && died[e]) ? id : expr_val(prog, st + e * bytes, dt, tg);
            acc = op_apply(op, dt, acc, v);
        }
    }
    const uint64_t r = tree_block(acc, lds, op, dt);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

}  // namespace

// host restatement of op_apply for the final combination with init
static uint64_t h_op(int op, int dt, uint64_t x, uint64_t y) {
    bool f = dt == VH_DT_F64 || dt == VH_DT_F32;
    if (f) {
        double a, b, r = 0;
        memcpy(&a, &x, 8);
        memcpy(&b, &y, 8);
        auto fmax_ = [](double p, double q) {
            if (p != p || q != q) return p != p ? p : q;
            if (p == q) return std::signbit(p) ? q : p;
            return p > q ? p : q;
        };
        auto fmin_ = [](double p, double q) {
            if (p != p || q != q) return p != p ? p : q;
            if (p == q) return std::signbit(p) ? p : q;
            return p < q ? p : q;
        };
        switch (op) {
        case VH_OP_ADD: r = a + b; break;
        case VH_OP_MUL: r = a * b; break;
        case VH_OP_MAX: r = fmax_(a, b); break;
        case VH_OP_MIN: r = fmin_(a, b); break;
        }
        uint64_t o;
        memcpy(&o, &r, 8);
        return o;
    }
    int64_t a = (int64_t)x, b = (int64_t)y;
    uint64_t r = 0;
    switch (op) {
    case VH_OP_ADD: r = x + y; break;
    case VH_OP_MUL: r = x * y; break;
    case VH_OP_MAX: r = (uint64_t)(a > b ? a : b); break;
    case VH_OP_MIN: r = (uint64_t)(a < b ? a : b); break;
    case VH_OP_AND: r = x & y; break;
    case VH_OP_OR: r = x | y; break;
    }
    if (dt == VH_DT_BOOL) r = r != 0;
    return r;
}

// val4empty src/Helpers.jl:44-81
static int val4empty(int op, int dt, uint64_t* out) {
    bool f = dt == VH_DT_F64 || dt == VH_DT_F32;
    double d;
    switch (op) {
    case VH_OP_ADD: d = 0.0; *out = f ? 0 : 0; if (f) memcpy(out, &d, 8); return VH_OK;
    case VH_OP_MUL: d = 1.0; if (f) memcpy(out, &d, 8); else *out = 1; return VH_OK;
    case VH_OP_MAX:
        d = -HUGE_VAL;
        if (f) memcpy(out, &d, 8); else *out = (uint64_t)(dt == VH_DT_BOOL ? -1 : -INT64_MAX);
        return VH_OK;
    case VH_OP_MIN:
        d = HUGE_VAL;
        if (f) memcpy(out, &d, 8); else *out = (uint64_t)(dt == VH_DT_BOOL ? 1 : INT64_MAX);
        return VH_OK;
    case VH_OP_AND:
        if (f) return set_error(VH_ERR_ASSERT, "The & operator is only supported for integer and boolean types");
        *out = dt == VH_DT_BOOL ? 1 : (uint64_t)INT64_MAX;
        return VH_OK;
    case VH_OP_OR:
        if (f) return set_error(VH_ERR_ASSERT, "The | operator is only supported for integer and boolean types");
        *out = 0;
        return VH_OK;
    }
    return set_error(VH_ERR_ASSERT, "Can not derive the init value for the operator. You must add "
                                    "this information via the `init` keyword.");
}

static int reduce_levels(Sim& s, uint64_t* buf, uint64_t n, int op, int dt, uint64_t* result) {
    hipStream_t st = s.stream;
    while (n > 1) {
        uint64_t nb = (n + MTILE - 1) / MTILE;
        uint64_t* o = (uint64_t*)s.scratch.take(nb * 8 + 8);
        if (!o) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_mr_values, dim3((unsigned)nb), dim3(MB), 0, st, buf, n, op, dt, o);
        buf = o;
        n = nb;
    }
    HIPCHECK(hipMemcpyAsync(result, buf, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return VH_OK;
}

// mapreduce src/AgentMethods.jl:533-565 (kind 0) / src/EdgeMethods.jl:972-1001 (kind 1)
extern "C" int vh_mapreduce(vh_sim_t h, int kind, uint8_t t, const vh_map_desc* f, int op, int dtype,
                            const void* init, void* out) {
    Sim& s = h->s;
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "mapreduce needs an initialized simulation");
    if (op < VH_OP_ADD || op > VH_OP_OR) return set_error(VH_ERR_ASSERT, "unknown operator");
    uint64_t acc;
    if (init) memcpy(&acc, init, 8);
    else RC(val4empty(op, dtype, &acc));
    MapSpec ms;
    memset(&ms, 0, sizeof ms);
    ms.kind = f->kind;
    ms.offset = f->offset;
    ms.fdt = f->field_dtype;
    ms.value = f->value;
    ms.op = op;
    ms.dtype = dtype;
    s.scratch.reset(s.stream);
    const uint8_t* recs;
    const uint8_t* died = nullptr;
    uint64_t n;
    uint32_t bytes;
    if (kind == 0) {
        if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
        AgentStore& A = s.a[t];
        recs = A.rstate();
        bytes = A.bytes;
        n = A.nextid - 1;
        if (A.mortal) died = A.died.as<uint8_t>();
    } else {
        if (t >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
        EdgeStore& E = s.e[t];
        if (has(E.hints, VH_HINT_STATELESS))
            return set_error(VH_ERR_ASSERT, "mapreduce is not defined for the hint combination");
        recs = E.st.as<uint8_t>();
        bytes = E.bytes;
        n = E.nedges;
    }
    uint64_t tot;
    if (n == 0) {
        uint64_t* o = (uint64_t*)s.scratch.take(8);
        hipLaunchKernelGGL(k_identity, dim3(1), dim3(1), 0, s.stream, o, op, dtype);
        RC(reduce_levels(s, o, 1, op, dtype, &tot));
    } else {
        uint64_t nb = (n + MTILE - 1) / MTILE;
        uint64_t* o = (uint64_t*)s.scratch.take(nb * 8 + 8);
        if (!o) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_mr_records, dim3((unsigned)nb), dim3(MB), 0, s.stream, recs, bytes, died, n,
                           ms, o);
        RC(reduce_levels(s, o, nb, op, dtype, &tot));
    }
    uint64_t r = h_op(op, dtype, tot, acc);
    memcpy(out, &r, 8);
    return VH_OK;
}

static uint32_t field_bytes(uint32_t dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_F64: case VH_DT_U64: return 8;
    case VH_DT_F32: case VH_DT_U32: case VH_DT_I32: return 4;
    case VH_DT_U16: return 2;
    case VH_DT_BOOL: case VH_DT_U8: return 1;
    }
    return 0;
}

// an expression program: known codes, fields inside `bytes`, stack depth
// within limits and one value at the end
static int check_expr(const vh_expr_op* prog, uint32_t nops, uint32_t bytes) {
    if (!prog || nops == 0 || nops > VH_EXPR_MAX_OPS)
        return set_error(VH_ERR_ARG, "mapreduce program: 1 to %d operations", VH_EXPR_MAX_OPS);
    int d = 0;
    for (uint32_t k = 0; k < nops; k++) {
        const uint32_t c = prog[k].code;
        int pop = 0, push = 1;
        switch (c) {
        case VH_EX_FIELD: {
            const uint32_t fb = field_bytes(prog[k].dtype);
            if (!fb || prog[k].arg + fb > bytes)
                return set_error(VH_ERR_ARG, "mapreduce program: field %u outside the state", k);
            break;
        }
        case VH_EX_CONST_I: case VH_EX_CONST_F: case VH_EX_CONST_B: break;
        case VH_EX_NEG: case VH_EX_ABS: case VH_EX_NOT: pop = 1; break;
        case VH_EX_ADD: case VH_EX_SUB: case VH_EX_MUL: case VH_EX_DIV: case VH_EX_MIN: case VH_EX_MAX:
        case VH_EX_EQ: case VH_EX_NE: case VH_EX_LT: case VH_EX_LE: case VH_EX_GT: case VH_EX_GE:
        case VH_EX_AND: case VH_EX_OR: pop = 2; break;
        case VH_EX_SELECT: pop = 3; break;
        default: return set_error(VH_ERR_ARG, "mapreduce program: unknown operation %u", c);
        }
        if (d < pop) return set_error(VH_ERR_ARG, "mapreduce program: operation %u lacks operands", k);
        d += push - pop;
        if (d > VH_EXPR_MAX_DEPTH) return set_error(VH_ERR_ARG, "mapreduce program: stack deeper than %d",
                                                    VH_EXPR_MAX_DEPTH);
    }
    if (d != 1) return set_error(VH_ERR_ARG, "mapreduce program: leaves %d values", d);
    return VH_OK;
}

// mapreduce with a computed f (vh_mapreduce_expr): vh_mapreduce's reduction
// tree with the program evaluated per record
extern "C" int vh_mapreduce_expr(vh_sim_t h, int kind, uint8_t t, const vh_expr_op* prog, uint32_t nops, int op,
                                 int dtype, const void* init, void* out) {
    if (!h || !out) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "mapreduce needs an initialized simulation");
    if (op < VH_OP_ADD || op > VH_OP_OR) return set_error(VH_ERR_ASSERT, "unknown operator");
    uint64_t acc;
    if (init) memcpy(&acc, init, 8);
    else RC(val4empty(op, dtype, &acc));
    const uint8_t* recs;
    const uint8_t* died = nullptr;
    uint64_t n;
    uint32_t bytes;
    if (kind == 0) {
        if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
        AgentStore& A = s.a[t];
        recs = A.rstate();
        bytes = A.bytes;
        n = A.nextid - 1;
        if (A.mortal) died = A.died.as<uint8_t>();
    } else {
        if (t >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
        EdgeStore& E = s.e[t];
        if (has(E.hints, VH_HINT_STATELESS))
            return set_error(VH_ERR_ASSERT, "mapreduce is not defined for the hint combination");
        recs = E.st.as<uint8_t>();
        bytes = E.bytes;
        n = E.nedges;
    }
    RC(check_expr(prog, nops, bytes));
    ExprProg P;
    memset(&P, 0, sizeof P);
    memcpy(P.op, prog, nops * sizeof(vh_expr_op));
    P.n = nops;
    s.scratch.reset(s.stream);
    uint64_t tot;
    if (n == 0) {
        uint64_t* o = (uint64_t*)s.scratch.take(8);
        if (!o) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_identity, dim3(1), dim3(1), 0, s.stream, o, op, dtype);
        RC(reduce_levels(s, o, 1, op, dtype, &tot));
    } else {
        const uint64_t nb = (n + MTILE - 1) / MTILE;
        uint64_t* o = (uint64_t*)s.scratch.take(nb * 8 + 8);
        if (!o) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_mr_expr, dim3((unsigned)nb), dim3(MB), 0, s.stream, recs, bytes, died, n, P, op,
                           dtype, o);
        RC(reduce_levels(s, o, nb, op, dtype, &tot));
    }
    const uint64_t r = h_op(op, dtype, tot, acc);
    memcpy(out, &r, 8);
    return VH_OK;
}

// ========================================== mapreduce without allocation
// vh_mapreduce_noalloc: vh_mapreduce's kernels and reduction tree with the
// block partials in `scratch`, a file-scope device array, instead of the
// simulation's scratch arena (which grows with hipMalloc). Level 1 holds at
// most MR_L1 block results (MR_L1 * MTILE = 2^28 records), levels 2 and 3 at
// most 64 and 1. Each GPU holds its own copy of the array until the process
// ends; its address comes from hipGetSymbolAddress once per device (the
// array's name is not a device address in host code). Every partial is
// written before it is read, so nothing needs clearing; calls that overlap on
// one GPU share the array and race.
constexpr uint64_t MR_L1 = 131072;
constexpr uint64_t MR_SCRATCH_WORDS = MR_L1 + 256;
constexpr int MR_MAX_DEVICES = 64;
static __device__ __align__(256) uint64_t scratch[MR_SCRATCH_WORDS];

static std::mutex g_scratch_mu;
static uint64_t* g_scratch_ptr[MR_MAX_DEVICES] = {};

// the device address of `scratch` on `dev` (looked up once per device, with
// that device current)
static int scratch_on(int dev, uint64_t** out) {
    if (dev < 0 || dev >= MR_MAX_DEVICES) return set_error(VH_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPCHECK(hipGetDevice(&cur));
        if (cur != dev) HIPCHECK(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) HIPCHECK(hipSetDevice(cur));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (uint64_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return VH_OK;
}

extern "C" int vh_mapreduce_noalloc(vh_sim_t h, int kind, uint8_t t, const vh_map_desc* f, int op, int dtype,
                                    const void* init, void* out) {
    if (!h || !f || !out) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "mapreduce needs an initialized simulation");
    if (op < VH_OP_ADD || op > VH_OP_OR) return set_error(VH_ERR_ASSERT, "unknown operator");
    uint64_t acc;
    if (init) memcpy(&acc, init, 8);
    else RC(val4empty(op, dtype, &acc));
    MapSpec ms;
    memset(&ms, 0, sizeof ms);
    ms.kind = f->kind;
    ms.offset = f->offset;
    ms.fdt = f->field_dtype;
    ms.value = f->value;
    ms.op = op;
    ms.dtype = dtype;
    const uint8_t* recs;
    const uint8_t* died = nullptr;
    uint64_t n;
    uint32_t bytes;
    if (kind == 0) {
        if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
        AgentStore& A = s.a[t];
        recs = A.rstate();
        bytes = A.bytes;
        n = A.nextid - 1;
        if (A.mortal) died = A.died.as<uint8_t>();
    } else {
        if (t >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
        EdgeStore& E = s.e[t];
        if (has(E.hints, VH_HINT_STATELESS))
            return set_error(VH_ERR_ASSERT, "mapreduce is not defined for the hint combination");
        recs = E.st.as<uint8_t>();
        bytes = E.bytes;
        n = E.nedges;
    }
    if (n > MR_L1 * MTILE)
        return set_error(VH_ERR_CAPACITY, "mapreduce_noalloc: %llu records exceed its maximum of %llu",
                         (unsigned long long)n, (unsigned long long)(MR_L1 * MTILE));
    uint64_t* area;
    RC(scratch_on(s.device, &area));
    hipStream_t st = s.stream;
    uint64_t* buf = area;
    uint64_t m;
    if (n == 0) {
        hipLaunchKernelGGL(k_identity, dim3(1), dim3(1), 0, st, buf, op, dtype);
        m = 1;
    } else {
        m = (n + MTILE - 1) / MTILE;   // <= MR_L1
        hipLaunchKernelGGL(k_mr_records, dim3((unsigned)m), dim3(MB), 0, st, recs, bytes, died, n, ms, buf);
    }
    // levels 2 (<= 64 values, at MR_L1) and 3 (1 value, at MR_L1 + 128)
    uint64_t* next = area + MR_L1;
    while (m > 1) {
        const uint64_t nb = (m + MTILE - 1) / MTILE;
        hipLaunchKernelGGL(k_mr_values, dim3((unsigned)nb), dim3(MB), 0, st, buf, m, op, dtype, next);
        buf = next;
        next += 128;
        m = nb;
    }
    HIPCHECK(hipGetLastError());
    uint64_t tot = 0;
    HIPCHECK(hipMemcpyAsync(&tot, buf, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const uint64_t r = h_op(op, dtype, tot, acc);
    memcpy(out, &r, 8);
    return VH_OK;
}

// num_agents src/Agent.jl:324-343
extern "C" int vh_num_agents(vh_sim_t h, uint8_t t, uint64_t* out) {
    Sim& s = h->s;
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    AgentStore& A = s.a[t];
    if (!A.mortal || !s.initialized) {
        *out = A.nextid - 1;
        return VH_OK;
    }
    vh_map_desc one;
    memset(&one, 0, sizeof one);
    one.kind = VH_MAP_ONE;
    int64_t z = 0;
    return vh_mapreduce(h, 0, t, &one, VH_OP_ADD, VH_DT_I64, &z, out);
}

// num_edges(sim, T) src/Edge.jl:373-389
extern "C" int vh_num_edges(vh_sim_t h, uint8_t et, uint64_t* out) {
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    *out = s.initialized ? s.e[et].nedges : s.e[et].h_to.size();
    return VH_OK;
}

extern "C" int vh_read_agents(vh_sim_t h, uint8_t t, void* states, uint8_t* died, uint64_t cap,
                              uint64_t* nslots) {
    Sim& s = h->s;
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    AgentStore& A = s.a[t];
    uint64_t n = s.initialized ? std::min(A.nslots, A.nextid - 1) : A.nextid - 1;
    *nslots = n;
    if (cap < n) return set_error(VH_ERR_CAPACITY, "capacity %llu < %llu", (unsigned long long)cap,
                                  (unsigned long long)n);
    if (!s.initialized) {
        if (states && A.bytes) memcpy(states, A.h_state.data(), n * A.bytes);
        if (died) memset(died, 0, n);
        return VH_OK;
    }
    // with an I/O stream the copies wait for this simulation's last apply!
    // only, not for other work queued on a shared compute stream
    hipStream_t q = s.stream;
    if (s.io) {
        HIPCHECK(hipEventSynchronize(s.last));
        q = s.io;
    }
    if (states && A.bytes && n)
        HIPCHECK(hipMemcpyAsync(states, A.rstate(), n * A.bytes, hipMemcpyDeviceToHost, q));
    if (died) {
        if (A.mortal && n) HIPCHECK(hipMemcpyAsync(died, A.died.p, n, hipMemcpyDeviceToHost, q));
        else memset(died, 0, n);
    }
    HIPCHECK(hipStreamSynchronize(q));
    return VH_OK;
}

struct HostCsr {
    std::vector<uint32_t> off;
    std::vector<uint64_t> src;
    std::vector<uint8_t> st;
};
static int fetch_csr(Sim& s, EdgeStore& E, HostCsr& H) {
    H.off.assign(E.rows + 1, 0);
    if (E.rows) HIPCHECK(hipMemcpyAsync(H.off.data(), E.off.p, (E.rows + 1) * 4, hipMemcpyDeviceToHost, s.stream));
    H.src.assign(E.nedges, 0);
    bool ign = has(E.hints, VH_HINT_IGNOREFROM);
    std::vector<uint32_t> s32;
    if (!ign && E.nedges) {
        if (E.src64) {
            HIPCHECK(hipMemcpyAsync(H.src.data(), E.src.p, E.nedges * 8, hipMemcpyDeviceToHost, s.stream));
        } else {
            s32.resize(E.nedges);
            HIPCHECK(hipMemcpyAsync(s32.data(), E.src.p, E.nedges * 4, hipMemcpyDeviceToHost, s.stream));
        }
    }
    H.st.assign(E.nedges * E.bytes, 0);
    if (E.bytes && E.nedges)
        HIPCHECK(hipMemcpyAsync(H.st.data(), E.st.p, E.nedges * E.bytes, hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    if (!s32.empty())
        for (uint64_t i = 0; i < E.nedges; i++) H.src[i] = vh_make_id(E.src_type, s32[i]);
    return VH_OK;
}

// all_edges src/EdgeMethods.jl:1004-1027: (to, from, state) in container order
extern "C" int vh_read_edges(vh_sim_t h, uint8_t et, uint64_t* to, uint64_t* from, void* states,
                             uint64_t cap, uint64_t* n) {
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    EdgeStore& E = s.e[et];
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "read_edges needs an initialized simulation");
    *n = E.nedges;
    if (cap < E.nedges) return set_error(VH_ERR_CAPACITY, "capacity");
    HostCsr H;
    RC(fetch_csr(s, E, H));
    uint64_t k = 0;
    for (int t = 1; t <= VH_MAXA; t++)
        for (uint64_t nr = 1; nr <= E.nrows[t]; nr++) {
            uint64_t r = E.rowbase[t] + nr - 1;
            for (uint32_t j = H.off[r]; j < H.off[r + 1]; j++, k++) {
                if (to) to[k] = vh_make_id(t, nr);
                if (from) from[k] = H.src[j];
                if (states && E.bytes) memcpy((uint8_t*)states + k * E.bytes, H.st.data() + (uint64_t)j * E.bytes, E.bytes);
            }
        }
    return VH_OK;
}

// edges / neighborids / edgestates / num_edges of one target
extern "C" int vh_edges_of(vh_sim_t h, uint8_t et, uint64_t to, uint64_t* from, void* states,
                           uint64_t cap, uint64_t* n) {
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    EdgeStore& E = s.e[et];
    *n = 0;
    uint32_t t = vh_type_of(to);
    uint64_t nr = vh_nr_of(to);
    if (has(E.hints, VH_HINT_SINGLETYPE) && (int)t != E.target)
        return set_error(VH_ERR_ASSERT, "The :SingleType hint is set for the edge type, but the agent "
                                        "has another type");
    if (!E.rows || t < 1 || t > VH_MAXA || nr == 0 || nr > E.nrows[t]) return VH_OK;
    uint64_t r = E.rowbase[t] + nr - 1;
    uint32_t o[2];
    HIPCHECK(hipMemcpyAsync(o, E.off.as<uint32_t>() + r, 8, hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    uint32_t cnt = o[1] - o[0];
    *n = cnt;
    if (!cnt) return VH_OK;
    if (cap < cnt) return set_error(VH_ERR_CAPACITY, "capacity");
    if (from) {
        if (has(E.hints, VH_HINT_IGNOREFROM)) {
            memset(from, 0, cnt * 8);
        } else if (E.src64) {
            HIPCHECK(hipMemcpyAsync(from, E.src.as<uint64_t>() + o[0], cnt * 8, hipMemcpyDeviceToHost, s.stream));
        } else {
            std::vector<uint32_t> v(cnt);
            HIPCHECK(hipMemcpyAsync(v.data(), E.src.as<uint32_t>() + o[0], cnt * 4, hipMemcpyDeviceToHost, s.stream));
            HIPCHECK(hipStreamSynchronize(s.stream));
            for (uint32_t i = 0; i < cnt; i++) from[i] = vh_make_id(E.src_type, v[i]);
        }
    }
    if (states && E.bytes)
        HIPCHECK(hipMemcpyAsync(states, E.st.as<uint8_t>() + (uint64_t)o[0] * E.bytes, (uint64_t)cnt * E.bytes,
                                hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    return VH_OK;
}

// agentstate outside a transition src/AgentMethods.jl:91-123
extern "C" int vh_agentstate(vh_sim_t h, uint64_t id, uint8_t t, void* out) {
    Sim& s = h->s;
    if (vh_type_of(id) != t || t < 1 || t > s.m->na)
        return set_error(VH_ERR_ASSERT, "The id of the agent does not match the given type");
    AgentStore& A = s.a[t];
    uint64_t nr = vh_nr_of(id);
    if (nr == 0 || nr > A.nslots) return set_error(VH_ERR_ASSERT, "agent nr out of range");
    if (A.mortal) {
        uint8_t d = 0;
        HIPCHECK(hipMemcpyAsync(&d, A.died.as<uint8_t>() + nr - 1, 1, hipMemcpyDeviceToHost, s.stream));
        HIPCHECK(hipStreamSynchronize(s.stream));
        if (d) return set_error(VH_ERR_ASSERT, "agentstate was requested for agent %llu that has been removed",
                                (unsigned long long)id);
    }
    if (A.bytes)
        HIPCHECK(hipMemcpyAsync(out, A.rstate() + (nr - 1) * A.bytes, A.bytes, hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    return VH_OK;
}

static int copy_buf(DevBuf& dst, const DevBuf& src, size_t bytes, hipStream_t st) {
    if (!src.p) return VH_OK;
    RC(dst.ensure(std::max<size_t>(src.cap, 256), st));
    if (bytes) HIPCHECK(hipMemcpyAsync(dst.p, src.p, std::min(bytes, src.cap), hipMemcpyDeviceToDevice, st));
    return VH_OK;
}

// copy_simulation src/Simulation.jl:500-510
extern "C" int vh_sim_copy(vh_sim_t h, vh_sim_t* out) {
    Sim& s = h->s;
    HIPCHECK(hipStreamSynchronize(s.stream));
    vh_sim* c = nullptr;
    vh_model types;
    types.m = *s.m;
    RC(vh_sim_create(&types, &c));
    Sim& d = c->s;
    d.initialized = s.initialized;
    d.num_transitions = s.num_transitions;
    for (int t = 1; t <= s.m->na; t++) {
        AgentStore& A = s.a[t];
        AgentStore& B = d.a[t];
        B.h_state = A.h_state;
        B.nextid = A.nextid;
        B.nslots = A.nslots;
        B.cap = A.cap;
        B.rd = A.rd;
        B.nreuse = A.nreuse;
        B.last_change = A.last_change;
        size_t sb = A.cap * A.bytes;
        RC(copy_buf(B.state[0], A.state[0], sb, d.stream));
        RC(copy_buf(B.state[1], A.state[1], sb, d.stream));
        RC(copy_buf(B.died, A.died, A.cap, d.stream));
        RC(copy_buf(B.died_now, A.died_now, A.cap, d.stream));
        RC(copy_buf(B.reuse, A.reuse, A.nreuse * 4, d.stream));
    }
    memcpy(d.vals, s.vals, sizeof d.vals);
    memcpy(d.vdt, s.vdt, sizeof d.vdt);
    d.vset[0] = s.vset[0];
    d.vset[1] = s.vset[1];
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        EdgeStore& F = d.e[et];
        F.h_from = E.h_from;
        F.h_to = E.h_to;
        F.h_st = E.h_st;
        F.h_first = E.h_first;
        // device-generated in-edges requested before finish_init!, and the
        // CSR version the stencil fast path keys on
        F.gen = E.gen;
        F.ba_m = E.ba_m;
        F.gen_mean_deg = E.gen_mean_deg;
        F.gen_seed = E.gen_seed;
        F.gen_wseed = E.gen_wseed;
        F.stencil = E.stencil;
        F.sspec = E.sspec;
        F.stencil_version = E.stencil_version;
        F.version = E.version;
        F.last_change = E.last_change;
        memcpy(F.rowbase, E.rowbase, sizeof F.rowbase);
        memcpy(F.nrows, E.nrows, sizeof F.nrows);
        F.rows = E.rows;
        F.nedges = E.nedges;
        F.src64 = E.src64;
        F.src_type = E.src_type;
        RC(copy_buf(F.off, E.off, (E.rows + 1) * 4, d.stream));
        RC(copy_buf(F.src, E.src, E.nedges * (E.src64 ? 8 : 4), d.stream));
        RC(copy_buf(F.st, E.st, E.nedges * E.bytes, d.stream));
    }
    HIPCHECK(hipStreamSynchronize(d.stream));
    *out = c;
    return VH_OK;
}

extern "C" int vh_device_bytes(vh_sim_t h, uint64_t* out) {
    Sim& s = h->s;
    uint64_t b = s.scratch.bytes_held() + s.d_ctx.cap + s.d_err.cap + s.d_stats.cap;
    for (int t = 0; t <= VH_MAXA; t++) {
        AgentStore& A = s.a[t];
        b += A.state[0].cap + A.state[1].cap + A.died.cap + A.died_now.cap + A.reuse.cap;
    }
    for (int et = 0; et < VH_MAXE; et++) {
        s.e[et].each_buf([&b](DevBuf& d) { b += d.cap; });
    }
    *out = b;
    return VH_OK;
}
