// vh_bodies.hip — transition bodies as HIP device code.
//
// Generic bodies run thread-per-slot through k_generic (the agent loop of
// src/AgentMethods.jl:193-245 with the write policy of :159-181); the model
// bodies (SIR, weighted bounded confidence, dynamic contacts) have their own
// kernels over the CSR.
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "vh_engine.h"

using namespace vh;

namespace {

// ----------------------------------------------------------------- bodies
template <int TR>
struct Body;

template <>
struct Body<VH_TR_NOOP> {
    __device__ static bool run(const DCtx*, const uint8_t*, uint64_t, Emitter&, uint8_t*) { return true; }
};
// create_sum_state_neighbors test/core.jl:71-83
template <>
struct Body<VH_TR_SUM_NEIGHBORS> {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter&, uint8_t* out) {
        int et = c->role_e[0];
        uint32_t b, e;
        int64_t sum = 0;
        if (d_row(c, et, id, b, e)) {
            const DEdge& E = c->e[et];
            if (VH_HAS(E.hints, VH_HINT_IGNOREFROM)) {
                vh_flag(c, VH_DERR_HINT);
                return true;
            }
            for (uint32_t k = b; k < e; k++) {
                const uint8_t* ns = d_state(c, vh_src_id(E, k));
                if (!ns) return true;
                sum += ld_i64(ns);
            }
        }
        st_i64(out, sum);
        return true;
    }
};
__device__ bool pred(const DCtx* c, const uint8_t* st, uint64_t id) {
    const vh_test_params* p = (const vh_test_params*)c->params;
    int64_t foo = c->a[vh_type_of(id)].bytes >= 8 ? ld_i64(st) : 0;
    switch (p->cmp) {
    case VH_CMP_ALWAYS: return true;
    case VH_CMP_NEVER: return false;
    case VH_CMP_LT: return foo < p->value;
    case VH_CMP_GE: return foo >= p->value;
    case VH_CMP_EQ: return foo == p->value;
    case VH_CMP_ODD: return (foo % 2) != 0;
    case VH_CMP_EVEN: return (foo % 2) == 0;
    case VH_CMP_NUM_EDGES_EQ: {
        uint32_t b, e;
        d_row(c, c->role_e[0], id, b, e);
        return (int64_t)(e - b) == p->value;
    }
    }
    return false;
}
template <>
struct Body<VH_TR_DIE_IF> {
    __device__ static bool run(const DCtx* c, const uint8_t* st, uint64_t id, Emitter&, uint8_t*) {
        return !pred(c, st, id);
    }
};
template <>
struct Body<VH_TR_ADD_SELF_EDGE> {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter& em, uint8_t*) {
        const vh_test_params* p = (const vh_test_params*)c->params;
        em.edge(id, id, p->edge_state);
        return true;
    }
};
// add_edges!(sim, id, edges(sim, id, ET)) test/edges.jl:357-359
template <>
struct Body<VH_TR_READD_EDGES> {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter& em, uint8_t*) {
        int et = c->role_e[0];
        uint32_t b, e;
        if (!d_row(c, et, id, b, e)) return true;
        const DEdge& E = c->e[et];
        for (uint32_t k = b; k < e; k++) {
            uint64_t from = VH_HAS(E.hints, VH_HINT_IGNOREFROM) ? 0 : vh_src_id(E, k);
            em.edge(from, id, E.st + (uint64_t)k * E.bytes);
        }
        return true;
    }
};
// test/independent.jl:55-68
template <>
struct Body<VH_TR_NEIGHBOR_FOO_EDGES> {
    __device__ static bool run(const DCtx* c, const uint8_t* st, uint64_t id, Emitter& em, uint8_t*) {
        const vh_test_params* p = (const vh_test_params*)c->params;
        int64_t foo = ld_i64(st);
        if (foo == p->value) return false;
        int et = c->role_e[0];
        uint32_t b, e;
        if (d_row(c, et, id, b, e)) {
            const DEdge& E = c->e[et];
            uint8_t es[8];
            st_i64(es, foo);
            for (uint32_t k = b; k < e; k++) {
                uint64_t nid = vh_src_id(E, k);
                em.edge(nid, id, es);
                em.edge(id, nid, es);
            }
        }
        return true;
    }
};
// test/independent.jl:165-176 (value births of T(foo + 10j)); cmp selects
// the edges: 0 both directions with the newborn id as state (independent.jl),
// 1 newborn -> caller only (test/addexisting.jl:41-44)
template <>
struct Body<VH_TR_SPAWN> {
    __device__ static bool run(const DCtx* c, const uint8_t* st, uint64_t id, Emitter& em, uint8_t*) {
        const vh_test_params* p = (const vh_test_params*)c->params;
        int at = c->role_a[0];
        uint8_t ns[VH_MAX_STATE];
        uint32_t bytes = at > 0 ? c->a[at].bytes : 0u;
        uint32_t cbytes = c->a[vh_type_of(id)].bytes;
        for (int64_t j = 1; j <= p->value; j++) {
            for (uint32_t k = 0; k < bytes; k++) ns[k] = k < cbytes ? st[k] : 0;
            if (bytes >= 8 && cbytes >= 8) st_i64(ns, ld_i64(st) + 10 * j);
            uint64_t nid = em.agent(ns);
            if (c->role_e[0] >= 0) {
                uint8_t es[8];
                st_i64(es, (int64_t)nid);
                em.edge(nid, id, es);
                if (p->cmp == 0) em.edge(id, nid, es);
            }
        }
        return true;
    }
};
template <>
struct Body<VH_TR_COUNT_EDGES> {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter&, uint8_t* out) {
        uint32_t b, e;
        d_row(c, c->role_e[0], id, b, e);
        st_i64(out, (int64_t)(e - b));
        return true;
    }
};

// diffuse test/raster.jl:38-41 (diffuse3D :143-146): the Bool field at
// `field_offset` becomes true when it is, or when any in-neighbour's is
template <>
struct Body<VH_TR_DIFFUSE> {
    __device__ static bool run(const DCtx* c, const uint8_t* st, uint64_t id, Emitter&, uint8_t* out) {
        const vh_raster_params* p = (const vh_raster_params*)c->params;
        const uint32_t f = p->field_offset;
        uint8_t v = st[f] ? 1 : 0;
        int et = c->role_e[0];
        uint32_t b, e;
        if (!v && d_row(c, et, id, b, e)) {
            const DEdge& E = c->e[et];
            for (uint32_t k = b; k < e && !v; k++) {
                const uint8_t* ns = d_state(c, vh_src_id(E, k));
                if (!ns) return true;
                v = ns[f] ? 1 : 0;
            }
        }
        out[f] = v;
        return true;
    }
};
// value_on_pos test/raster.jl:279-285: v = the Int64 at offset 0 of the
// first in-neighbour over E; move_to!(sim, raster, id, (v, ..., v), E(), E())
// with distance 0 (src/Raster.jl:437-463): edge cell -> id, then id -> cell;
// the new state is v at offset 0
template <>
struct Body<VH_TR_MOVE_TO> {
    __device__ static bool run(const DCtx* c, const uint8_t*, uint64_t id, Emitter& em, uint8_t* out) {
        const vh_raster_params* p = (const vh_raster_params*)c->params;
        int et = c->role_e[0];
        uint32_t b, e;
        if (!d_row(c, et, id, b, e)) {   // first(nothing) fails in the reference
            vh_flag(c, VH_DERR_TARGET);
            return true;
        }
        const uint8_t* ns = d_state(c, vh_src_id(c->e[et], b));
        if (!ns) return true;
        const int64_t v = ld_i64(ns);
        uint64_t lin = 0, stride = 1;
        for (uint32_t d = 0; d < p->ndims && d < 4; d++) {
            if (v < 1 || v > p->dims[d]) {   // raster[pos] out of bounds
                vh_flag(c, VH_DERR_TARGET);
                return true;
            }
            lin += (uint64_t)(v - 1) * stride;
            stride *= (uint64_t)p->dims[d];
        }
        const uint64_t cell = vh_make_id(p->cell_type, p->first_nr + lin);
        if (p->flags & 1u) em.edge(cell, id, p->edge_state);
        if (p->flags & 2u) em.edge(id, cell, p->edge_state);
        st_i64(out, v);
        return true;
    }
};

// remove_edges! bodies of test/mpi/test_edgetypes.jl:295-445 (enum
// vh_remove_mode); `idx` is the Int64 at offset 0 of the caller's state
template <>
struct Body<VH_TR_REMOVE_EDGES> {
    // first (or second) in-neighbour of `id` over `et`; false when it has fewer
    __device__ static bool nth_neighbor(const DCtx* c, int et, uint64_t id, uint32_t n, uint64_t& out) {
        uint32_t b, e;
        if (!d_row(c, et, id, b, e) || e - b <= n) return false;
        const DEdge& E = c->e[et];
        if (VH_HAS(E.hints, VH_HINT_IGNOREFROM)) {
            vh_flag(c, VH_DERR_HINT);
            return false;
        }
        out = vh_src_id(E, b + n);
        return true;
    }
    __device__ static bool run(const DCtx* c, const uint8_t* st, uint64_t id, Emitter& em, uint8_t*) {
        const vh_test_params* p = (const vh_test_params*)c->params;
        const int64_t idx = ld_i64(st);
        const int et = c->role_e[0], es = c->role_e[1];
        uint64_t nb = 0, nb2 = 0;
        switch ((int)p->value) {
        case VH_RM_OWN:
            if (idx % 2 == 0) em.remove(0, id, true);
            break;
        case VH_RM_NEIGHBOR:
            if (idx % 2 == 0) {
                if (!nth_neighbor(c, et, id, 0, nb)) vh_flag(c, VH_DERR_TARGET);
                else em.remove(0, nb, true);
            }
            break;
        case VH_RM_READD:
            if (!nth_neighbor(c, et, id, 0, nb)) {
                vh_flag(c, VH_DERR_TARGET);
                break;
            }
            em.remove(nb, id, false);
            em.edge(nb, id, p->edge_state);
            break;
        case VH_RM_FROM_TWO: {
            const bool single = VH_HAS(c->e[et].hints, VH_HINT_SINGLEEDGE);
            if (!nth_neighbor(c, et, id, 0, nb) || (!single && !nth_neighbor(c, et, id, 1, nb2))) {
                vh_flag(c, VH_DERR_TARGET);
                break;
            }
            em.remove(nb, id, false);
            if (!single) em.remove(nb2, id, false);
            break;
        }
        case VH_RM_TO_NEIGHBOR:
            if (idx % 2 == 0) {
                if (!nth_neighbor(c, et, id, 0, nb)) vh_flag(c, VH_DERR_TARGET);
                else em.remove(id, nb, false);
            }
            break;
        case VH_RM_SEND_ID:   // the log is bound to SendId (role_edge[1])
            if (!nth_neighbor(c, et, id, 0, nb)) vh_flag(c, VH_DERR_TARGET);
            else em.edge(id, nb, p->edge_state);
            break;
        case VH_RM_TO_SENDID:
            if (idx < 100) {
                if (!nth_neighbor(c, es, id, 0, nb)) vh_flag(c, VH_DERR_TARGET);
                else em.remove(id, nb, false);
            }
            break;
        case VH_RM_ZERO_FROM:
            em.remove(0, id, false);
            break;
        }
        return true;
    }
};

template <int TR>
__global__ __launch_bounds__(256) void k_generic(const DCtx* __restrict__ c, int ci, int t, int pass,
                                                 int with_edge) {
    vh_generic_call<Body<TR>>(c, ci, t, pass, with_edge);
}

// ----------------------------------------------------------------- SIR
// state word: health byte 0, pad byte 1, days bytes 2-3
// Exposure pass: every infected agent (read state) sets the bit of each of
// its out-neighbours in `expo` (cleared before). A susceptible agent without
// the bit has no infected in-neighbour, so "some in-edge with an infected
// source and a draw below p" is false without drawing: k_sir skips its draws.
// The pass runs while the infected are few: k_sir counts the infected agents
// it writes into 64 counters per step parity (one atomic per block), and
// k_sir_decide, at the start of the next step, sums them into the decision
// word that the pass and k_sir both read, so both take the same decision.
// Words of `ctl`: [0, 128) counters (parity * 64 + j), 128 + parity: the
// counters of that parity are complete, 130: the decision.
constexpr uint32_t SIR_CTL_WORDS = 131;

__global__ void k_sir_decide(uint32_t* __restrict__ ctl, uint64_t step, uint64_t n) {
    __shared__ uint32_t s_sum[64];
    const uint32_t j = threadIdx.x, cur = (uint32_t)(step & 1), nxt = cur ^ 1u;
    s_sum[j] = ctl[cur * 64 + j];
    __syncthreads();
    if (j == 0) {
        uint64_t total = 0;
        for (int q = 0; q < 64; q++) total += s_sum[q];
        ctl[130] = ctl[128 + cur] && total * 32 <= n ? 1u : 0u;
        ctl[128 + nxt] = 1u;   // this step's k_sir fills the other set
    }
    ctl[nxt * 64 + j] = 0u;
}
__device__ __forceinline__ bool sir_expose_on(const uint32_t* ctl) { return ctl && ctl[130] != 0u; }
__global__ __launch_bounds__(256) void k_sir_expose(const uint32_t* __restrict__ rs,
                                                    const uint32_t* __restrict__ out_off,
                                                    const uint32_t* __restrict__ out_dst, uint64_t n,
                                                    const uint32_t* __restrict__ ctl,
                                                    uint32_t* __restrict__ expo) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !sir_expose_on(ctl) || (rs[i] & 0xFFu) != VH_I) return;
    for (uint32_t k = out_off[i]; k < out_off[i + 1]; k++) {
        const uint32_t v = out_dst[k];
        atomicOr(expo + (v >> 5), 1u << (v & 31u));
    }
}

// row of every in-edge (transpose input) and its source index as key
__global__ void k_edge_rows(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, uint64_t rows,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    for (uint32_t k = off[r]; k < off[r + 1]; k++) {
        keys[k] = src[k] - 1u;
        vals[k] = (uint32_t)r;
    }
}
// off[v] = first index whose key >= v, v in [0, n]
__global__ void k_sorted_offsets(const uint32_t* __restrict__ keys, uint64_t ne, uint64_t n,
                                 uint32_t* __restrict__ off) {
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    uint64_t lo = 0, hi = ne;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    off[v] = (uint32_t)lo;
}

__global__ __launch_bounds__(256) void k_sir(const uint32_t* __restrict__ off,
                                             const uint32_t* __restrict__ src,
                                             const uint32_t* __restrict__ rs,
                                             uint32_t* __restrict__ ws, uint64_t n, uint64_t nrows,
                                             uint64_t rowbase, uint32_t t, double p, uint64_t seed,
                                             uint32_t rec, uint64_t step, const uint32_t* __restrict__ expo,
                                             uint32_t* __restrict__ ctl) {
    __shared__ uint32_t s_inf[4];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < n;
    uint32_t s = in ? rs[i] : 0u;
    uint32_t h = s & 0xFFu;
    const bool skip = in && h == VH_S && sir_expose_on(ctl) && !((expo[i >> 5] >> (i & 31u)) & 1u);
    if (in && h == VH_S && !skip) {
        if (i < nrows) {
            const uint64_t pre = vh_hash_pre(seed, vh_make_id(t, i + 1), step);
            const uint64_t pt = vh_u01_below(p);
            uint32_t b = off[rowbase + i], e = off[rowbase + i + 1];
            // "some in-edge k has an infected source and u_k < p" is
            // order-free: the draw (ALU) first, the source's state (a random
            // gather) only for the few draws below p
            if (vh_any_infecting(pre, pt, b, e, [&](uint32_t k) { return (rs[src[k] - 1] & 0xFFu) == VH_I; }))
                s = (s & 0xFF00u) | VH_I;
        }
    } else if (in && h == VH_I) {
        uint32_t days = ((s >> 16) + 1u) & 0xFFFFu;
        s = (s & 0xFF00u) | (days << 16) | (days >= rec ? VH_R : VH_I);
    }
    if (in) ws[i] = s;
    if (ctl) {   // infected agents of the written state, for the next step
        const uint64_t m = __ballot(in && (s & 0xFFu) == VH_I);
        if ((threadIdx.x & 63u) == 0) s_inf[threadIdx.x >> 6] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t c = s_inf[0] + s_inf[1] + s_inf[2] + s_inf[3];
            if (c) atomicAdd(ctl + ((step + 1) & 1) * 64 + (blockIdx.x & 63u), c);
        }
    }
}

// ----------------------------------------------------------------- opinion
// Weighted bounded confidence: x_i <- sum_{j: |x_j - x_i| <= eps} w x_j /
// sum w over the in-edges; x_i when none is accepted. Canonical order of both
// sums: degree <= 64 sequential in CSR (insertion) order; degree > 64 lane l
// of 64 adds k = l, l + 64, ... in order, then the xor butterfly
// p[l] + p[l ^ o], o = 32..1. A rejected neighbour adds +0.0, which is
// bit-identical to skipping it (the sums start at +0.0 and every term is
// >= +0.0). Agents of degree <= 64 are light (thread per agent in a 256-agent
// block), the others heavy (one wave each).
template <class T>
__device__ __forceinline__ void op_accum(T xj, float w, T xi, T eps, T& num, T& den) {
    const bool acc = fabs(xj - xi) <= eps;
    num = num + (acc ? (T)w * xj : (T)0);
    den = den + (acc ? (T)w : (T)0);
}

// inclusive sum over the lanes of a wave up to each lane (DPP: shifts within
// rows of 16 lanes, then the row totals broadcast to the rows after them)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return v;
}

// the value of lane l ^ O (O = 1, 2: DPP quad permutations; 4, 8, 16:
// ds_swizzle in bitmask mode; 32: ds_bpermute)
template <int O>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v) {
    if constexpr (O == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    else if constexpr (O == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    else if constexpr (O <= 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (O << 10) | 0x1F);
    else return (uint32_t)__shfl_xor((int)v, O, 64);
}
template <int O, class T>
__device__ __forceinline__ T lane_xor_t(T v) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint64_t r = (uint64_t)lane_xor<O>((uint32_t)u) | (uint64_t)lane_xor<O>((uint32_t)(u >> 32)) << 32;
        return __builtin_bit_cast(T, r);
    } else {
        return __builtin_bit_cast(T, lane_xor<O>(__builtin_bit_cast(uint32_t, v)));
    }
}
// p[l] + p[l ^ o] for o = 32, 16, ..., 1: every lane ends with the sum in
// the canonical order of the wave-per-agent path
template <class T>
__device__ __forceinline__ T xor_butterfly(T v) {
    v = v + lane_xor_t<32>(v);
    v = v + lane_xor_t<16>(v);
    v = v + lane_xor_t<8>(v);
    v = v + lane_xor_t<4>(v);
    v = v + lane_xor_t<2>(v);
    v = v + lane_xor_t<1>(v);
    return v;
}

// block exclusive scan of the light degrees: the thread's first slot in the
// block's light edges and their total
__device__ __forceinline__ uint32_t op_light_scan(uint32_t ld, uint32_t* s_wsum, uint32_t& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(ld);
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    uint32_t pre = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint32_t v = s_wsum[q];
        pre += q < wv ? v : 0u;
        total += v;
    }
    return pre + incl - ld;
}

// -- one pass (state vectors below the plan threshold). The light edges of a
// block are processed in tiles of OP_TILE slots: all threads load sources and
// weights coalesced and gather OP_U neighbour states each, staging state and
// weight by slot in LDS; then every thread adds its own agent's slots.
constexpr int OP_TILE = 2048;
constexpr int OP_U = 8;   // gathers in flight per thread

template <class T>
__device__ __forceinline__ void opinion_block(const uint32_t* __restrict__ off,
                                              const uint32_t* __restrict__ src,
                                              const T* __restrict__ x, const float* __restrict__ w,
                                              T* __restrict__ xo, uint64_t n, uint64_t nrows,
                                              uint64_t rowbase, T eps, uint32_t blk) {
    __shared__ T s_x[OP_TILE];
    __shared__ float s_w[OP_TILE];
    __shared__ uint8_t s_own[OP_TILE];
    __shared__ uint32_t s_b[256], s_p[256], s_wsum[4];
    const int tid = threadIdx.x;
    const uint64_t i = (uint64_t)blk * 256 + tid;
    uint32_t b = 0, d = 0;
    T xi = (T)0;
    if (i < n) {
        xi = (*(x + i));
        if (i < nrows) {
            b = (*(off + rowbase + i));
            d = (*(off + rowbase + i + 1)) - b;
        }
    }
    const bool light = i < n && d <= 64u;
    const uint32_t ld = light ? d : 0u;
    uint32_t total;
    const uint32_t P = op_light_scan(ld, s_wsum, total);
    s_b[tid] = b;
    s_p[tid] = P;
    T num = (T)0, den = (T)0;
    for (uint32_t t0 = 0; t0 < total; t0 += OP_TILE) {
        const uint32_t lo = P > t0 ? P : t0;
        const uint32_t hi = (P + ld) < (t0 + OP_TILE) ? (P + ld) : (t0 + OP_TILE);
        for (uint32_t p = lo; p < hi; p++) s_own[p - t0] = (uint8_t)tid;
        __syncthreads();
        const uint32_t cnt = (total - t0) < (uint32_t)OP_TILE ? (total - t0) : (uint32_t)OP_TILE;
        // batches of OP_U slots per thread: all index loads, then all gathers
        for (uint32_t q0 = 0; q0 < cnt; q0 += 256u * OP_U) {
            uint32_t kk[OP_U], sj[OP_U];
            float wk[OP_U];
            T xj[OP_U];
#pragma unroll
            for (int u = 0; u < OP_U; u++) {
                const uint32_t q = q0 + (uint32_t)u * 256u + tid;
                const uint32_t a = q < cnt ? s_own[q] : 0u;
                kk[u] = q < cnt ? s_b[a] + (t0 + q - s_p[a]) : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int u = 0; u < OP_U; u++) {
                sj[u] = kk[u] != 0xFFFFFFFFu ? (*(src + kk[u])) : 1u;
                wk[u] = kk[u] != 0xFFFFFFFFu ? (*(w + kk[u])) : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < OP_U; u++) xj[u] = kk[u] != 0xFFFFFFFFu ? x[sj[u] - 1] : (T)0;
#pragma unroll
            for (int u = 0; u < OP_U; u++) {
                const uint32_t q = q0 + (uint32_t)u * 256u + tid;
                if (q < cnt) {
                    s_x[q] = xj[u];
                    s_w[q] = wk[u];
                }
            }
        }
        __syncthreads();
        for (uint32_t p = lo; p < hi; p++) op_accum(s_x[p - t0], s_w[p - t0], xi, eps, num, den);
        __syncthreads();
    }
    if (light) *(xo + i) = den > (T)0 ? num / den : xi;
}

template <class T>
__device__ __forceinline__ void opinion_heavy(const uint32_t* __restrict__ off,
                                              const uint32_t* __restrict__ src,
                                              const T* __restrict__ x, const float* __restrict__ w,
                                              T* __restrict__ xo, uint64_t r, uint64_t rowbase, T eps) {
    const int lane = threadIdx.x & 63;
    const uint64_t i = r - rowbase;
    const T xi = x[i];
    const uint32_t b = off[r], e = off[r + 1];
    T num = (T)0, den = (T)0;
    // lane l adds k = l, l+64, ... in order; four strides in flight per lane
    uint32_t k = b + lane;
    for (; k + 192 < e; k += 256) {
        const uint32_t s0 = src[k], s1 = src[k + 64], s2 = src[k + 128], s3 = src[k + 192];
        const float w0 = w[k], w1 = w[k + 64], w2 = w[k + 128], w3 = w[k + 192];
        const T x0 = x[s0 - 1], x1 = x[s1 - 1], x2 = x[s2 - 1], x3 = x[s3 - 1];
        op_accum(x0, w0, xi, eps, num, den);
        op_accum(x1, w1, xi, eps, num, den);
        op_accum(x2, w2, xi, eps, num, den);
        op_accum(x3, w3, xi, eps, num, den);
    }
    for (; k < e; k += 64) op_accum(x[src[k] - 1], w[k], xi, eps, num, den);
    num = xor_butterfly(num);
    den = xor_butterfly(den);
    if (lane == 0) xo[i] = den > (T)0 ? num / den : xi;
}

// blocks [0, nh) run one wave per heavy agent (the longest work starts
// first), the others are light blocks
template <class T>
__global__ __launch_bounds__(256) void k_opinion(const uint32_t* __restrict__ off,
                                                 const uint32_t* __restrict__ src,
                                                 const T* __restrict__ x, const float* __restrict__ w,
                                                 T* __restrict__ xo, uint64_t n, uint64_t nrows,
                                                 uint64_t rowbase, T eps, const uint32_t* __restrict__ heavy,
                                                 uint64_t nheavy, uint32_t nh) {
    if (blockIdx.x >= nh) {
        opinion_block<T>(off, src, x, w, xo, n, nrows, rowbase, eps, blockIdx.x - nh);
    } else {
        const uint64_t wv = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        if (wv < nheavy) opinion_heavy<T>(off, src, x, w, xo, heavy[wv], rowbase, eps);
    }
}

// Many simulations' one-pass opinion steps in one launch: block b belongs to
// the member with the largest blk0 <= b (its heavy-wave blocks first, then
// its light blocks), every member with its own CSR, states and eps.
template <class T>
__global__ __launch_bounds__(256) void k_opinion_multi(const OpMember* __restrict__ mem, uint32_t nm) {
    uint32_t lo = 0, hi = nm;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (mem[mid].blk0 <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const OpMember& m = mem[lo];
    const uint32_t local = blockIdx.x - m.blk0;
    if (local >= m.nh) {
        opinion_block<T>(m.off, m.src, (const T*)m.x, m.w, (T*)m.xo, m.n, m.nrows, m.rowbase, (T)m.eps,
                         local - m.nh);
    } else {
        const uint64_t wv = ((uint64_t)local * blockDim.x + threadIdx.x) >> 6;
        if (wv < m.nheavy) opinion_heavy<T>(m.off, m.src, (const T*)m.x, m.w, (T*)m.xo, m.heavy[wv], m.rowbase,
                                           (T)m.eps);
    }
}

// ----------------------------------------------------------------- gather plan
// For a large state vector the neighbour states of a static edge store are
// gathered in two passes through a plan built once per edge store.
//
// Windows: the sources are cut into windows of `wsz` consecutive agents, at
// most 160 KB of states (one CU's LDS) and at most WIN_RUNS windows. Blocks:
// the agents are cut into blocks of consecutive agents with fewer than TB
// staged in-edges, TB being the slots a block stages in LDS; an agent's
// in-edges are staged when its in-degree is at most WIN_HMAX (rows above go
// through chunk lists). An agent costs its staged in-degree + 1; block b
// holds the agents whose exclusive cost prefix lies in [b S, (b + 1) S),
// S = TB - WIN_HMAX - 1. The plan orders the CSR slots by (window, block),
// the chunk-list rows' slots after the staged ones of their window, keeping
// CSR order within: the staged in-edges of a block from one window are one
// run of the plan.
//
// Pass 1 (k_win_gather) streams the plan: each block of the launch takes a
// contiguous share, stages each window it meets in LDS and writes
// y[j] = x[src(j)] for its share, so the random reads are LDS reads and HBM
// sees sequential streams (the plan's local source indices, y, the windows).
//
// Pass 2 (k_win_sum) takes the agent blocks from a queue. A block's staged
// entries are listed by (window, slot): entry e holds its slot (bit 15 set on
// the first entry of a run) and its weight, and its value is
// y[prun[r] + e - pfirst[r]] for its run r, counted from the run starts
// before its 64-entry chunk (cbase) and a ballot of the start bits. Values and
// weights are staged in LDS by slot; then each agent of in-degree <= 64 adds
// its slots in order, and a wave each agent of in-degree 65..WIN_HMAX (lane l
// its slots l, l + 64, ..., then the xor butterfly). Agents above WIN_HMAX go
// first, one wave each, through their chunk lists. Most of the next block's
// values, and its list, runs and agents, load while a block is summed. The
// values are exactly x[src[k]] and the sums keep the canonical order, so the
// results are the one-pass kernel's.
template <class T>
struct WinCfg {
    static constexpr uint32_t WSZ = 163840 / sizeof(T);             // states per window at most
    static constexpr uint32_t TB = sizeof(T) == 8 ? 12288 : 16384;  // staged slots per block at most
    static constexpr uint32_t QU = TB / 1024;                       // list entries per thread
    static constexpr uint32_t QA = sizeof(T) == 8 ? 8 : 12;         // of them, loaded during the previous sums
    static constexpr uint32_t CPB = TB / 64;                        // 64-entry chunks per block
    static constexpr uint32_t CHUNK = sizeof(T) == 8 ? 256 : 384;   // heavy-row slots staged per wave
};
constexpr uint32_t WIN_RUNS = 1024;   // windows at most (runs of a block)
constexpr uint32_t WIN_HMAX = 512;    // in-degrees staged within blocks (larger: chunk lists)
constexpr uint32_t WIN_SLOT = 0x3FFFu, WIN_START = 0x8000u;

// A heavy agent (one wave) from its chunk lists: the in-edges of heavy agent
// h are numbered from hbase[h] in CSR order, in chunks of CHUNK slots, each
// chunk listed in plan order (hty: plan position, hslot: slot in the chunk).
// The wave copies a chunk's values into its LDS region `buf` by slot, then
// lane l adds slots l, l + 64, ... in order: the canonical lane-strided
// partials, then the xor butterfly.
template <class T>
__device__ __forceinline__ void opinion_heavy_plan(const T* __restrict__ x, const T* __restrict__ y,
                                                   const float* __restrict__ w, T* __restrict__ xo,
                                                   const uint4 hi, T eps, const uint32_t* __restrict__ hty,
                                                   const uint16_t* __restrict__ hslot, T* buf) {
    constexpr uint32_t CH = WinCfg<T>::CHUNK;
    constexpr int HU = CH / 64;
    const uint32_t lane = threadIdx.x & 63;
    // hi: row, CSR start, degree, first chunk-list entry
    const uint64_t r = hi.x;
    const uint32_t b = hi.y, d = hi.z, hb = hi.w;
    const T xi = x[r];
    T num = (T)0, den = (T)0;
    for (uint32_t c0 = 0; c0 < d; c0 += CH) {
        const uint32_t cnt = d - c0 < CH ? d - c0 : CH;
        float wq[HU];
        uint32_t jj[HU], sl[HU];
        T v[HU];
#pragma unroll
        for (int u = 0; u < HU; u++) {
            const uint32_t q = lane + 64u * (uint32_t)u;
            wq[u] = 0.0f;
            jj[u] = 0u;
            sl[u] = 0xFFFFFFFFu;
            if (q < cnt) {
                wq[u] = w[b + c0 + q];
                jj[u] = hty[hb + c0 + q];
                sl[u] = hslot[hb + c0 + q];
            }
        }
#pragma unroll
        for (int u = 0; u < HU; u++) v[u] = sl[u] != 0xFFFFFFFFu ? y[jj[u]] : (T)0;
#pragma unroll
        for (int u = 0; u < HU; u++)
            if (sl[u] != 0xFFFFFFFFu) buf[sl[u]] = v[u];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int u = 0; u < HU; u++) {
            const uint32_t q = lane + 64u * (uint32_t)u;
            if (q < cnt) op_accum(buf[q], wq[u], xi, eps, num, den);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    num = xor_butterfly(num);
    den = xor_butterfly(den);
    if (lane == 0) xo[r] = den > (T)0 ? num / den : xi;
}


// Pass 1: block l of the grid (one per CU) takes plan entries
// [ne l / G, ne (l + 1) / G). For each window its share meets: the window's
// states go to LDS (16-byte loads), then wave v, step u writes the 256
// entries from c = g0 + 256 (v + 16 u) in two halves, lane l entries
// c + 2l, c + 2l + 1 and c + 128 + 2l, c + 129 + 2l (one 4-byte load of their
// two source indices, one 2-state store: each store instruction writes one
// contiguous range; measured 236 us against 308 us with 4 entries per lane
// at 10^8 entries), for U steps u at a time: the loads of all U issued
// before the stores (10^7 agents, 10^8 entries, ms per opinion step, FP64 /
// FP32: U = 2 0.689-0.692 / 0.515-0.520, U = 4 0.670-0.674 / 0.472-0.477,
// U = 6 0.660-0.672 / 0.455-0.461, U = 8 0.664-0.673 / 0.456-0.460).
typedef uint32_t pu32x4 __attribute__((ext_vector_type(4)));

template <class T>
struct Pair;
template <>
struct Pair<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<float> {
    typedef float type __attribute__((ext_vector_type(2)));
};

template <class T>
__global__ __launch_bounds__(1024) void k_win_gather(const uint16_t* __restrict__ psl, const T* __restrict__ x,
                                                     T* __restrict__ y, const uint32_t* __restrict__ wstart,
                                                     uint32_t W, uint32_t wsz, uint64_t nsrc, uint64_t ne) {
    typedef typename Pair<T>::type T2;
    constexpr uint32_t PER16 = 16 / sizeof(T);
    constexpr int U = 6;
    __shared__ __align__(16) T win[WinCfg<T>::WSZ];
    const uint32_t G = gridDim.x, l = blockIdx.x;
    const uint64_t j0 = ne * l / G, j1 = ne * (l + 1) / G;
    if (j0 >= j1) return;   // whole block
    uint32_t lo = 0, hi = W;   // the window of j0: wstart[lo] <= j0 < wstart[lo + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (wstart[mid] <= j0) lo = mid;
        else hi = mid;
    }
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t w = lo;
    uint64_t j = j0;
    while (j < j1) {
        const uint64_t we = wstart[w + 1];
        if (we <= j) {   // empty window
            w++;
            continue;
        }
        const uint64_t s0 = (uint64_t)w * wsz;
        const uint32_t wn = (uint32_t)((nsrc - s0) < wsz ? nsrc - s0 : wsz);
        for (uint32_t q = threadIdx.x * PER16; q < wn; q += 1024 * PER16) {
            if (q + PER16 <= wn) {
                *(pu32x4*)(win + q) = __builtin_nontemporal_load((const pu32x4*)(x + s0 + q));
            } else {
                for (uint32_t t = q; t < wn; t++) win[t] = x[s0 + t];
            }
        }
        __syncthreads();
        const uint64_t je = we < j1 ? we : j1;
        for (uint64_t g0 = j & ~3ull; g0 < je; g0 += 4096ull * U) {
            uint32_t p0[U], p1[U];
            bool f0[U], f1[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t ga = g0 + 256ull * (wv + 16 * u) + 2 * lane, gb = ga + 128;
                f0[u] = ga >= j && ga + 2 <= je;
                f1[u] = gb >= j && gb + 2 <= je;
                p0[u] = f0[u] ? __builtin_nontemporal_load((const uint32_t*)(psl + ga)) : 0u;
                p1[u] = f1[u] ? __builtin_nontemporal_load((const uint32_t*)(psl + gb)) : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t ga = g0 + 256ull * (wv + 16 * u) + 2 * lane, gb = ga + 128;
                if (f0[u]) {
                    const T2 v = {win[p0[u] & 0xFFFFu], win[p0[u] >> 16]};
                    *(T2*)(y + ga) = v;
                } else {
                    for (uint64_t t = ga; t < ga + 2; t++)
                        if (t >= j && t < je) y[t] = win[psl[t]];
                }
                if (f1[u]) {
                    const T2 v = {win[p1[u] & 0xFFFFu], win[p1[u] >> 16]};
                    *(T2*)(y + gb) = v;
                } else {
                    for (uint64_t t = gb; t < gb + 2; t++)
                        if (t >= j && t < je) y[t] = win[psl[t]];
                }
            }
        }
        __syncthreads();
        j = je;
        w++;
    }
}

template <class T>
struct WinLDS {
    T val[WinCfg<T>::TB];          // values by slot (heavy waves: CHUNK each)
    float wgt[WinCfg<T>::TB];      // weights by slot
    uint32_t rp[2][WIN_RUNS];      // runs of this block and the next: plan position, list index
    uint16_t rf[2][WIN_RUNS];
    uint32_t wsum[WinCfg<T>::TB / 1024 + 1][16];
    uint32_t nextq[2];             // queue items taken ahead
    uint32_t nmed;                 // agents of in-degree 65..WIN_HMAX of the block (< TB / 64)
    uint32_t mpk[WinCfg<T>::TB / 64], magent[WinCfg<T>::TB / 64];   // slot | in-degree << 16
    T mx[WinCfg<T>::TB / 64];
};

// The round-0 and round-1 agents of a block (a0 + tid, a0 + 1024 + tid):
// in-degrees and states, loaded with the block's list
template <class T>
struct WinAgents {
    uint32_t o0, o1, o2, o3;   // raw loads; selected where used
    T x0, x1;
};

template <class T>
__device__ __forceinline__ WinAgents<T> win_agents(const uint32_t* __restrict__ off, const T* __restrict__ x,
                                                   uint64_t nrows, uint32_t a0, uint32_t a1, uint32_t tid) {
    WinAgents<T> g;
    const uint32_t i0 = a0 + tid, i1 = i0 + 1024;
    const uint32_t j0 = i0 < a1 ? i0 : a0, j1 = i1 < a1 ? i1 : a0;
    const uint32_t r0 = j0 < nrows ? j0 : 0u, r1 = j1 < nrows ? j1 : 0u;
    g.x0 = x[j0];
    g.x1 = x[j1];
    g.o0 = off[r0];
    g.o1 = off[r0 + 1];
    g.o2 = off[r1];
    g.o3 = off[r1 + 1];
    return g;
}

// The agents of a block, in rounds of 1024 (agent a0 + 1024 r + tid): one
// exclusive scan of the staged in-degrees over all rounds (one barrier, which
// also publishes the staged values); each agent of in-degree <= 64 adds its
// slots in order, the others of the block are listed and summed by a wave
// each (lane-strided partials, xor butterfly). Rounds 0 and 1 use the agents
// loaded early (g); later rounds load theirs when LD (without LD the block
// has at most two rounds, and no global load is issued). after_scan() runs
// right after the scan's barrier.
template <class T, bool LD, class F>
__device__ __forceinline__ void win_sum_agents(WinLDS<T>& sm, const uint32_t* __restrict__ off,
                                               const T* __restrict__ x, T* __restrict__ xo, uint64_t nrows, T eps,
                                               uint32_t a0, uint32_t a1, const WinAgents<T>& g, F after_scan) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t K = (a1 - a0 + 1023) / 1024;
    uint32_t inclk[2] = {0u, 0u}, ldk[2] = {0u, 0u};
    for (uint32_t r = 0; r < K; r++) {
        const uint32_t i = a0 + r * 1024 + tid;
        uint32_t d = 0;
        if (i < a1 && i < nrows) d = r == 0 ? g.o1 - g.o0 : r == 1 ? g.o3 - g.o2 : LD ? off[i + 1] - off[i] : 0u;
        const uint32_t ld = d <= WIN_HMAX ? d : 0u;
        const uint32_t incl = wave_incl_scan(ld);
        if (lane == 63) sm.wsum[r][wv] = incl;
        if (r < 2) {
            inclk[r] = incl;
            ldk[r] = ld;
        }
    }
    if (tid == 0) sm.nmed = 0;
    __syncthreads();
    after_scan();
    uint32_t base = 0;
    for (uint32_t r = 0; r < K; r++) {
        const uint32_t i = a0 + r * 1024 + tid;
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const uint32_t v = sm.wsum[r][q];
            pre += q < wv ? v : 0u;
            tot += v;
        }
        uint32_t d, incl, ld;
        T xi;
        if (!LD || r < 2) {
            d = (i < a1 && i < nrows) ? (r == 0 ? g.o1 - g.o0 : g.o3 - g.o2) : 0u;
            incl = inclk[r];
            ld = ldk[r];
            xi = i < a1 ? (r == 0 ? g.x0 : g.x1) : (T)0;
        } else {
            d = (i < a1 && i < nrows) ? off[i + 1] - off[i] : 0u;
            ld = d <= WIN_HMAX ? d : 0u;
            incl = wave_incl_scan(ld);
            xi = i < a1 ? x[i] : (T)0;
        }
        const uint32_t P = base + pre + incl - ld;
        if (i < a1 && d <= 64u) {
            T num = (T)0, den = (T)0;
            uint32_t q = 0;
            for (; q + 4 <= d; q += 4) {
                const T y0 = sm.val[P + q], y1 = sm.val[P + q + 1], y2 = sm.val[P + q + 2], y3 = sm.val[P + q + 3];
                const float w0 = sm.wgt[P + q], w1 = sm.wgt[P + q + 1], w2 = sm.wgt[P + q + 2],
                            w3 = sm.wgt[P + q + 3];
                op_accum(y0, w0, xi, eps, num, den);
                op_accum(y1, w1, xi, eps, num, den);
                op_accum(y2, w2, xi, eps, num, den);
                op_accum(y3, w3, xi, eps, num, den);
            }
            for (; q < d; q++) op_accum(sm.val[P + q], sm.wgt[P + q], xi, eps, num, den);
            xo[i] = den > (T)0 ? num / den : xi;
        } else if (i < a1 && d <= WIN_HMAX) {
            const uint32_t m = atomicAdd(&sm.nmed, 1u);   // < TB / 64 (each costs > 65 of < TB)
            sm.mpk[m] = P | d << 16;
            sm.magent[m] = i;
            sm.mx[m] = xi;
        }
        base += tot;
    }
    __syncthreads();
    const uint32_t nmed = sm.nmed;
    for (uint32_t m = (uint32_t)wv; m < nmed; m += 16) {
        const uint32_t P = sm.mpk[m] & 0xFFFFu, d = sm.mpk[m] >> 16;
        const T xi = sm.mx[m];
        T num = (T)0, den = (T)0;
        for (uint32_t q = (uint32_t)lane; q < d; q += 64) op_accum(sm.val[P + q], sm.wgt[P + q], xi, eps, num, den);
        num = xor_butterfly(num);
        den = xor_butterfly(den);
        if (lane == 0) xo[sm.magent[m]] = den > (T)0 ? num / den : xi;
    }
}

// Pass 2: one block of 1024 per CU. The waves first take the heavy agents
// (wave g of the grid agents g, g + 16 G, ...), then the blocks take light
// blocks from a queue (queue[0]: next light block, queue[1]: blocks done;
// the last block to finish sets both to 0 for the next launch). Each thread
// stages QU list entries of a block; the values of its first QA entries load
// while the previous block is summed (in registers: the block's LDS is in
// use), the other QU - QA after the sums. Measured slower at 10^7 agents:
// all QU loaded during the sums (past the 128 VGPRs of a 1024-thread block:
// spills), 512 threads with 256 VGPRs and all QU during the sums, two
// half-size blocks in LDS, each agent's first slot from the plan instead of
// the block scan; QA = 6 or 10 of 12 (FP64) was no faster than 8.
template <class T>
__global__ __launch_bounds__(1024) void k_win_sum(const uint32_t* __restrict__ off, const T* __restrict__ x,
                                                  const T* __restrict__ y, const float* __restrict__ w,
                                                  T* __restrict__ xo, uint64_t nrows, T eps,
                                                  const uint4* __restrict__ hinfo, uint64_t nheavy,
                                                  const uint32_t* __restrict__ hty,
                                                  const uint16_t* __restrict__ hslot,
                                                  const uint32_t* __restrict__ ablk,
                                                  const uint32_t* __restrict__ lbase,
                                                  const uint16_t* __restrict__ tpk, const float* __restrict__ twt,
                                                  const uint32_t* __restrict__ rbase,
                                                  const uint32_t* __restrict__ prun,
                                                  const uint16_t* __restrict__ pfirst,
                                                  const uint16_t* __restrict__ cbase, uint32_t nb,
                                                  uint32_t* __restrict__ queue) {
    constexpr uint32_t QU = WinCfg<T>::QU, QA = WinCfg<T>::QA, QB = QU - QA;
    static_assert(QA % 2 == 0 && QB % 2 == 0 && QB > 0, "chunk run counts are read in 16-bit pairs");
    __shared__ __align__(16) unsigned char smraw[sizeof(WinLDS<T>)];
    WinLDS<T>& sm = *reinterpret_cast<WinLDS<T>*>(smraw);
    const uint32_t tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint64_t lemask = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    for (uint64_t h = (uint64_t)blockIdx.x * 16 + (uint32_t)wv; h < nheavy; h += (uint64_t)gridDim.x * 16)
        opinion_heavy_plan<T>(x, y, w, xo, hinfo[h], eps, hty, hslot, sm.val + wv * WinCfg<T>::CHUNK);
    // every load is branch-free (lanes past a list read a valid dummy, an
    // inactive list reads as empty): a load under a branch makes the compiler
    // wait for all loads in flight where the branch joins; and loads complete
    // in issue order, so the sums issue no global load (blocks of at most two
    // rounds of agents, the others take a path that does)
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    uint32_t ltid = tid;
    auto load_list = [&](uint32_t b, bool act, uint32_t* pk) __attribute__((always_inline)) {
        const uint32_t lb = lbase[b], L = act ? lbase[b + 1] - lb : 0u;
#pragma unroll
        for (uint32_t u = 0; u < QU; u++) {
            const uint32_t e = ltid + 1024 * u;
            pk[u] = tpk[lb + (e < L ? e : 0u)];   // (past the list: masked where used)
        }
    };
    auto load_runs = [&](uint32_t b, uint32_t& a, uint32_t& f) __attribute__((always_inline)) {
        const uint32_t rb = rbase[b], R = rbase[b + 1] - rb;
        const uint32_t r = rb + (ltid < R ? ltid : 0u);   // (R == 0: the slack after the runs)
        a = prun[r];
        f = pfirst[r];
    };
    // values and weights of entries u in [U0, U0 + QH) of block b (its runs in
    // sm.rp[c]; the wave's chunk run counts by scalar loads)
    auto load_vals = [&](auto NN, uint32_t b, bool act, uint32_t c, uint32_t U0, const uint32_t* pk, T* v,
                         float* wt) __attribute__((always_inline)) {
        constexpr uint32_t QH = decltype(NN)::value;
        const uint32_t lb = lbase[b], L = act ? lbase[b + 1] - lb : 0u;
        const uint32_t* cw = (const uint32_t*)(cbase + ((uint64_t)b * 256 + (uint32_t)wv * 16));
        uint32_t cwv[QH / 2];
#pragma unroll
        for (uint32_t q = 0; q < QH / 2; q++) cwv[q] = cw[U0 / 2 + q];
        // the run table reads of all QH entries first (unconditional: run 0
        // for lanes past the list), then the global loads
        uint32_t rpv[QH], rfv[QH];
#pragma unroll
        for (uint32_t k = 0; k < QH; k++) {
            const uint32_t u = U0 + k;
            const uint32_t e = ltid + 1024 * u;
            const bool ok = e < L;
            const uint64_t m = __ballot(ok && (pk[u] & WIN_START));
            const uint32_t cb = (cwv[k >> 1] >> ((k & 1) * 16)) & 0xFFFFu;
            const uint32_t run = ok ? cb + __popcll(m & lemask) - 1u : 0u;
            rpv[k] = sm.rp[c][run];
            rfv[k] = sm.rf[c][run];
        }
#pragma unroll
        for (uint32_t k = 0; k < QH; k++) {
            const uint32_t e = ltid + 1024 * (U0 + k);
            const bool ok = e < L;
            const uint32_t pos = ok ? rpv[k] + (e - rfv[k]) : 0u;
            v[k] = y[pos];
            wt[k] = twt[lb + (ok ? e : 0u)];
        }
    };
    auto scatter = [&](auto NN, uint32_t L, uint32_t U0, const uint32_t* pk, const T* v, const float* wt)
                       __attribute__((always_inline)) {
        constexpr uint32_t QH = decltype(NN)::value;
#pragma unroll
        for (uint32_t k = 0; k < QH; k++)
            if (ltid + 1024 * (U0 + k) < L) {
                sm.val[pk[U0 + k] & WIN_SLOT] = v[k];
                sm.wgt[pk[U0 + k] & WIN_SLOT] = wt[k];
            }
    };
    if (tid == 0) {
        sm.nextq[0] = atomicAdd(queue, 1u);
        sm.nextq[1] = atomicAdd(queue, 1u);
    }
    __syncthreads();   // (also ends the heavy agents' use of sm.val)
    const uint32_t q0 = __builtin_amdgcn_readfirstlane(sm.nextq[0]), q1 = __builtin_amdgcn_readfirstlane(sm.nextq[1]);
    uint32_t b = q0 < nb ? q0 : NONE, bn = q1 < nb ? q1 : NONE;
    __syncthreads();
    if (b != NONE) {
        uint32_t pk[QU];
        T vA[QA];
        float wA[QA];
        WinAgents<T> ag;
        int cur = 0, par = 0;
        {
            uint32_t ra, rfv;
            load_runs(b, ra, rfv);
            load_list(b, true, pk);
            ag = win_agents<T>(off, x, nrows, ablk[b], ablk[b + 1], ltid);
            sm.rp[0][tid] = ra;
            sm.rf[0][tid] = (uint16_t)rfv;
            __syncthreads();
            load_vals(std::integral_constant<uint32_t, QA>(), b, true, 0, 0, pk, vA, wA);
        }
        // per block: the rest of its values; the next block's list, runs and
        // agents; all its values staged; the next queue item; the sums, whose
        // scan barrier publishes the staged values and the next runs, after
        // which the first QA values of the next block are issued (in flight
        // during the sums); a barrier
        for (;;) {
            asm volatile("" : "+v"(ltid));
            const bool hasn = bn != NONE;
            const uint32_t bnc = hasn ? bn : b;
            const uint32_t L = lbase[b + 1] - lbase[b];
            T vB[QB];
            float wB[QB];
            load_vals(std::integral_constant<uint32_t, QB>(), b, true, cur, QA, pk, vB, wB);
            uint32_t pkn[QU];
            uint32_t ra, rfv;
            load_list(bnc, hasn, pkn);
            load_runs(bnc, ra, rfv);
            const WinAgents<T> agn = win_agents<T>(off, x, nrows, ablk[bnc], ablk[bnc + 1], ltid);
            scatter(std::integral_constant<uint32_t, QA>(), L, 0, pk, vA, wA);
            scatter(std::integral_constant<uint32_t, QB>(), L, QA, pk, vB, wB);
            sm.rp[cur ^ 1][tid] = ra;
            sm.rf[cur ^ 1][tid] = (uint16_t)rfv;
            if (tid == 0) sm.nextq[par] = atomicAdd(queue, 1u);
            auto after_scan = [&]() __attribute__((always_inline)) {
                load_vals(std::integral_constant<uint32_t, QA>(), bnc, hasn, cur ^ 1, 0, pkn, vA, wA);
            };
            {
                // a block of at most two rounds sums without a global load,
                // so the next block's values stay in flight
                const uint32_t a0 = ablk[b], a1 = ablk[b + 1];
                if (a1 - a0 <= 2048u)
                    win_sum_agents<T, false>(sm, off, x, xo, nrows, eps, a0, a1, ag, after_scan);
                else {
                    win_sum_agents<T, true>(sm, off, x, xo, nrows, eps, a0, a1, ag, after_scan);
                    __builtin_amdgcn_s_waitcnt(0);   // (its loads have drained the counter anyway)
                }
            }
            __syncthreads();
            if (!hasn) break;
            const uint32_t it = __builtin_amdgcn_readfirstlane(sm.nextq[par]);
            b = bn;
            bn = it < nb ? it : NONE;
#pragma unroll
            for (uint32_t u = 0; u < QU; u++) pk[u] = pkn[u];
            ag = agn;
            cur ^= 1;
            par ^= 1;
        }
    }
    if (tid == 0 && atomicAdd(queue + 1, 1u) == gridDim.x - 1) {
        queue[0] = 0;
        queue[1] = 0;
    }
}

// consistency of a built plan, one thread per block: its agents, list and
// runs in range, every run's first entry and only those marked, each run
// inside y, each chunk's run count; any failure sets *bad (the plan is then
// refused: pass 2 trusts these bounds)
__global__ void k_win_check(const uint32_t* __restrict__ ablk, const uint32_t* __restrict__ lbase,
                            const uint32_t* __restrict__ rbase, const uint32_t* __restrict__ prun,
                            const uint16_t* __restrict__ pfirst, const uint16_t* __restrict__ tpk,
                            const uint16_t* __restrict__ cbase, uint32_t nb, uint64_t n, uint64_t nl,
                            uint64_t nruns, uint64_t ne, uint32_t TB, uint32_t cpb, uint32_t* __restrict__ bad) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const uint32_t a0 = ablk[b], a1 = ablk[b + 1], lb = lbase[b], lb1 = lbase[b + 1];
    const uint32_t rb = rbase[b], rb1 = rbase[b + 1];
    bool ok = a0 < a1 && a1 <= n && lb <= lb1 && lb1 <= nl && lb1 - lb < TB && rb <= rb1 && rb1 <= nruns &&
              rb1 - rb <= WIN_RUNS && (lb1 > lb) == (rb1 > rb);
    if (ok) {
        const uint32_t L = lb1 - lb;
        uint32_t prev = 0;
        for (uint32_t r = rb; r < rb1 && ok; r++) {
            const uint32_t f = pfirst[r], f1 = r + 1 < rb1 ? pfirst[r + 1] : L;
            ok = (r == rb ? f == 0 : f > prev) && f1 > f && f1 <= L && (uint64_t)prun[r] + (f1 - f) <= ne;
            for (uint32_t e = f; e < f1 && ok; e++) {
                const uint32_t pk = tpk[lb + e];
                ok = ((pk & WIN_START) != 0) == (e == f) && (pk & WIN_SLOT) < L;
            }
            prev = f;
        }
        uint32_t r = rb;
        for (uint32_t c = 0; c < cpb && 64u * c < L && ok; c++) {
            while (r < rb1 && pfirst[r] < 64u * c) r++;
            ok = cbase[b * 256 + (c % 16) * 16 + c / 16] == r - rb;
        }
    }
    if (!ok) atomicOr(bad, 1u);
}

// plan build: staged in-degree + 1 of each agent (rows past nrows: 1)
__global__ void k_win_cost(const uint32_t* __restrict__ off, uint64_t n, uint64_t nrows,
                           uint32_t* __restrict__ cost) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = i < nrows ? off[i + 1] - off[i] : 0u;
    cost[i] = (d <= WIN_HMAX ? d : 0u) + 1u;
}

// first agent and first list entry of each block (C: exclusive cost prefix,
// n + 1 entries; the thread of agent n writes the ends)
__global__ void k_win_blocks(const uint32_t* __restrict__ C, uint64_t n, uint32_t S, uint32_t nb,
                             uint32_t* __restrict__ ablk, uint32_t* __restrict__ lbase) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        ablk[nb] = (uint32_t)n;
        lbase[nb] = C[n] - (uint32_t)n;
        return;
    }
    const uint32_t b = C[i] / S;
    if (i == 0 || C[i - 1] / S != b) {
        ablk[b] = (uint32_t)i;
        lbase[b] = C[i] - (uint32_t)i;
    }
}

// plan sort key of each CSR slot: window * (nb + 1) + (chunk-list row ? nb : block)
__global__ void k_win_keys(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, uint64_t nrows,
                           const uint32_t* __restrict__ C, uint32_t S, uint32_t nb, uint32_t wsz,
                           uint32_t* __restrict__ keys) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const uint32_t b = off[r], e = off[r + 1];
    const uint32_t col = e - b > WIN_HMAX ? nb : C[r] / S;
    for (uint32_t k = b; k < e; k++) keys[k] = ((src[k] - 1u) / wsz) * (nb + 1u) + col;
}

// per plan entry j: local source index, inverse permutation; the first plan
// entry of every window (wstart, W + 1 entries)
__global__ void k_win_fill(const uint32_t* __restrict__ src, const uint32_t* __restrict__ keys,
                           const uint32_t* __restrict__ perm, uint64_t ne, uint32_t nb, uint32_t wsz, uint32_t W,
                           uint16_t* __restrict__ psl, uint32_t* __restrict__ pinv, uint32_t* __restrict__ wstart) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ne) return;
    const uint32_t k = perm[j], w = keys[j] / (nb + 1u);
    psl[j] = (uint16_t)((src[k] - 1u) - w * wsz);
    pinv[k] = (uint32_t)j;
    const int64_t wp = j ? (int64_t)(keys[j - 1] / (nb + 1u)) : -1;
    for (int64_t q = wp + 1; q <= (int64_t)w; q++) wstart[q] = (uint32_t)j;
    if (j == ne - 1)
        for (uint32_t q = w + 1; q <= W; q++) wstart[q] = (uint32_t)ne;
}

// list sort key of each staged in-edge (numbered by the staged prefix
// C[r] - r + position): block * W + window; its CSR slot
__global__ void k_win_lkeys(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, uint64_t nrows,
                            const uint32_t* __restrict__ C, uint32_t S, uint32_t W, uint32_t wsz,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ kidx) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const uint32_t b = off[r], d = off[r + 1] - b;
    if (d > WIN_HMAX) return;
    const uint32_t blk = C[r] / S, l0 = C[r] - (uint32_t)r;
    for (uint32_t q = 0; q < d; q++) {
        keys[l0 + q] = blk * W + (src[b + q] - 1u) / wsz;
        kidx[l0 + q] = b + q;
    }
}

// 1 at the first list entry of each run (sorted keys)
__global__ void k_win_starts(const uint32_t* __restrict__ keys, uint64_t nl, uint32_t* __restrict__ flag) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nl) return;
    flag[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1u : 0u;
}

// packed lists: slot | run start, weight in list order; per run its plan
// position and list index in the block (ridx: exclusive run count)
__global__ void k_win_pack(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ order,
                           const uint32_t* __restrict__ kidx, const uint32_t* __restrict__ flag,
                           const uint32_t* __restrict__ ridx, const uint32_t* __restrict__ pinv,
                           const uint32_t* __restrict__ lbase, uint32_t W, uint64_t nl, const float* __restrict__ w,
                           uint16_t* __restrict__ tpk, float* __restrict__ twt, uint32_t* __restrict__ prun,
                           uint16_t* __restrict__ pfirst) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nl) return;
    const uint32_t idx = order[e], blk = keys[e] / W, k = kidx[idx];
    const uint32_t st = flag[e];
    tpk[e] = (uint16_t)((idx - lbase[blk]) | (st ? WIN_START : 0u));
    twt[e] = w ? w[k] : 0.0f;
    if (st) {
        prun[ridx[e]] = pinv[k];
        pfirst[ridx[e]] = (uint16_t)(e - lbase[blk]);
    }
}

// first run of each block (nb + 1 entries) and each chunk's run count, at
// cbase[b 256 + (c % 16) 16 + c / 16]: the run counts of wave w's chunks
// (c = w, w + 16, ...) are 32 contiguous bytes
__global__ void k_win_rbase(const uint32_t* __restrict__ lbase, const uint32_t* __restrict__ ridx, uint32_t nb,
                            uint32_t cpb, uint32_t* __restrict__ rbase, uint16_t* __restrict__ cbase) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t b = (uint32_t)(t / cpb), c = (uint32_t)(t % cpb);
    if (b > nb) return;
    if (b == nb) {
        if (c == 0) rbase[nb] = ridx[lbase[nb]];
        return;
    }
    const uint32_t lb = lbase[b], L = lbase[b + 1] - lb;
    if (c == 0) rbase[b] = ridx[lb];
    cbase[(uint64_t)b * 256 + (c % 16) * 16 + c / 16] = (uint16_t)(64u * c < L ? ridx[lb + 64u * c] - ridx[lb] : 0u);
}

// heavy chunk lists: plan position and chunk slot per entry
__global__ void k_heavy_fill(const uint32_t* __restrict__ order, const uint32_t* __restrict__ kidx,
                             const uint16_t* __restrict__ tsl, const uint32_t* __restrict__ pinv, uint64_t nl,
                             uint32_t* __restrict__ hty, uint16_t* __restrict__ hslot) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nl) return;
    const uint32_t idx = order[e];
    hty[e] = pinv[kidx[idx]];
    hslot[e] = tsl[idx];
}

// in-edge count and chunk count of each heavy row
__global__ void k_heavy_sizes(const uint32_t* __restrict__ off, const uint32_t* __restrict__ heavy, uint64_t nheavy,
                              uint32_t chunk, uint32_t* __restrict__ deg, uint32_t* __restrict__ chunks) {
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nheavy) return;
    const uint32_t r = heavy[h], d = off[r + 1] - off[r];
    deg[h] = d;
    chunks[h] = (d + chunk - 1) / chunk;
}

// one wave per heavy row: sort key (chunk, source window), CSR slot and
// slot within the chunk of each of its in-edges (numbered from hbase[h])
__global__ __launch_bounds__(256) void k_heavy_keys(const uint32_t* __restrict__ off,
                                                    const uint32_t* __restrict__ src,
                                                    const uint32_t* __restrict__ heavy, uint64_t nheavy,
                                                    const uint32_t* __restrict__ hbase,
                                                    const uint32_t* __restrict__ cbase, uint32_t psize,
                                                    uint32_t pbits, uint32_t chunk, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ kidx, uint16_t* __restrict__ tsl) {
    const uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (h >= nheavy) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t r = heavy[h], b = off[r], d = off[r + 1] - b, hb = hbase[h], cb = cbase[h];
    for (uint32_t t = lane; t < d; t += 64) {
        keys[hb + t] = ((cb + t / chunk) << pbits) | ((src[b + t] - 1u) / psize);
        kidx[hb + t] = b + t;
        tsl[hb + t] = (uint16_t)(t % chunk);
    }
}

// row, CSR start, degree and first chunk-list entry of each heavy row
__global__ void k_heavy_info(const uint32_t* __restrict__ off, const uint32_t* __restrict__ heavy, uint64_t nheavy,
                             const uint32_t* __restrict__ hbase, uint4* __restrict__ hinfo) {
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nheavy) return;
    const uint32_t r = heavy[h], b = off[r];
    hinfo[h] = make_uint4(r, b, off[r + 1] - b, hbase[h]);
}

__global__ void k_heavy_rows(const uint32_t* __restrict__ off, uint64_t rows, uint64_t rowbase,
                             uint64_t nrows, uint32_t* __restrict__ list, uint32_t* __restrict__ cnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows) return;
    const uint64_t r = rowbase + i;
    if (off[r + 1] - off[r] > 64u) list[atomicAdd(cnt, 1u)] = (uint32_t)r;
}

// ----------------------------------------------------------------- contact
// state: health byte 0, susceptibility float at bytes 4-7.
// The step runs on per-slot flag bitmaps (64 slots per word): `sp` marks
// slots that are dead in the read state or die in this step, `dead` the
// former, `inf` slots infected in the read state. A partner draw reads one
// bit of `sp` and, for the ~4% of partners it marks, one of `dead`. At 10^7
// slots each map is 1.25 MB, so the two hot ones stay in one XCD's 4 MB L2
// while the draws and neighbour checks gather from them (the state vector
// itself is 80 MB and is streamed past the cache).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bool cbit(const uint64_t* __restrict__ b, uint64_t i) {
    return (b[i >> 6] >> (i & 63)) & 1ull;
}

// one thread per slot, whole waves per word (grid = nwords * 64 threads)
__global__ __launch_bounds__(256) void k_contact_flags(const uint64_t* __restrict__ rs,
                                                       const uint8_t* __restrict__ died, uint64_t n,
                                                       uint32_t t, vh_contact_params p, uint64_t step,
                                                       ContactBits B) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t w = i >> 6;
    if (w >= B.nwords) return;   // uniform per wave
    bool in = i < n;
    bool dead = !in || died[i] != 0;
    bool inf = in && (__builtin_nontemporal_load(rs + i) & 0xFFu) == VH_I;
    bool dn = !dead && vh_u01(vh_hash(p.seed, vh_make_id(t, i + 1), step, VH_KEY_DEATH)) < p.p_death;
    uint64_t md = __ballot(dead), mi = __ballot(inf), mn = __ballot(dn);
    if ((threadIdx.x & 63u) == 0) {
        B.sp[w] = md | mn;
        B.dead[w] = md;
        B.inf[w] = mi;
        B.dcnt[w] = (uint32_t)__popcll(mn);
    }
}

// Health update, death flag, partner draws. The kept partners of every
// slot (accepted draws whose partner does not die in this step, in draw
// order) go to a fixed-stride log, row i at log[i * n_contacts], written
// through LDS as one contiguous block range. Dynamic LDS: 256 * n_contacts
// words when `staged`.
__global__ __launch_bounds__(256) void k_contact_main(
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ src, const uint64_t* __restrict__ rs,
    uint64_t* __restrict__ ws, uint8_t* __restrict__ died_now, uint64_t n, uint64_t nrows,
    uint64_t rowbase, uint32_t t, vh_contact_params p, uint64_t step, ContactBits B,
    uint32_t* __restrict__ kept, uint32_t* __restrict__ log, uint32_t staged) {
    extern __shared__ u32x4 lds_v[];   // 16-byte aligned
    uint32_t* lds = (uint32_t*)lds_v;
    const uint32_t C = p.n_contacts;
    const uint32_t tid = threadIdx.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * 256;
    const uint64_t i = i0 + tid;

    bool live = i < n && !cbit(B.dead, i);
    bool survive = false;
    uint32_t nk = 0;
    if (live) {
        const uint64_t id = vh_make_id(t, i + 1);
        const uint64_t pre = vh_hash_pre(p.seed, id, step);
        const uint64_t s = __builtin_nontemporal_load(rs + i);
        uint32_t h = (uint32_t)(s & 0xFFu);
        if (h == VH_S) {
            if (off && i < nrows) {
                float susc = __uint_as_float((uint32_t)(s >> 32));
                double pinf = p.beta * (double)susc;
                const uint64_t pt = vh_u01_below(pinf);
                uint32_t b = off[rowbase + i], e = off[rowbase + i + 1];
                // "some in-edge k has an infected source and u_k < pinf" is
                // order-free: test the draw first (ALU) and look up the
                // source's flag only for the few draws below pinf
                if (vh_any_infecting(pre, pt, b, e, [&](uint32_t k) { return cbit(B.inf, src[k] - 1); }))
                    h = VH_I;
            }
        } else if (h == VH_I) {
            if (vh_u01(vh_hash_fin(pre, VH_KEY_RECOVER)) < p.gamma) h = VH_R;
        }
        if (cbit(B.sp, i))   // live and marked: dies in this step
            died_now[i] = 1;
        else {
            survive = true;
            __builtin_nontemporal_store((s & ~(uint64_t)0xFFu) | h, ws + i);
        }
    }
    if (survive) {
        const uint64_t pre = vh_hash_pre(p.seed, vh_make_id(t, i + 1), step);
        uint32_t* row = staged ? lds + (uint64_t)tid * C : log + i * C;
        uint64_t key = VH_KEY_PARTNER;
        for (uint32_t cc = 0; cc < C; cc++, key += p.max_redraw) {
            for (uint32_t r = 0; r < p.max_redraw; r++) {
                uint64_t hh = vh_hash_fin(pre, key + r);
                uint64_t j = 1 + (((hh >> 32) * n) >> 32);
                if (j == i + 1) continue;
                if (cbit(B.sp, j - 1)) {
                    if (cbit(B.dead, j - 1)) continue;   // dead in the read state: redraw
                    break;                               // dies in this step: not kept
                }
                row[nk++] = (uint32_t)j;
                break;
            }
        }
    }
    if (i < n) kept[i] = nk;
    if (staged) {
        __syncthreads();
        uint64_t rows_here = n - i0 < 256 ? n - i0 : 256;
        uint64_t len = rows_here * C;
        uint32_t* dst = log + i0 * C;   // 16-byte aligned: i0 is a multiple of 256
        uint64_t len4 = len >> 2;
        for (uint64_t q = tid; q < len4; q += 256)
            __builtin_nontemporal_store(*(const u32x4*)(lds + 4 * q), (u32x4*)(dst + 4 * q));
        for (uint64_t q = 4 * len4 + tid; q < len; q += 256) dst[q] = lds[q];
    }
}

// newborns of this step's deaths, in slot order of the dying callers: the
// top of the reuse stack first, then fresh slots (src/AgentMethods.jl:37-89)
__global__ void k_contact_births(uint64_t n, uint32_t t, vh_contact_params p, uint64_t step,
                                 ContactBits B, const uint32_t* __restrict__ dpos,
                                 const uint32_t* __restrict__ reuse, uint64_t nreuse, uint64_t nextid,
                                 uint64_t* __restrict__ ws, uint8_t* __restrict__ died) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t word = B.sp[i >> 6] & ~B.dead[i >> 6];   // dying in this step
    uint32_t lane = (uint32_t)(i & 63);
    if (!((word >> lane) & 1ull)) return;
    uint64_t g = dpos[i >> 6] + (uint64_t)__popcll(word & ((1ull << lane) - 1ull));
    uint64_t slot = g < nreuse ? reuse[nreuse - 1 - g] : nextid + (g - nreuse);
    float ns = 0.5f + 0.5f * vh_u01f(vh_hash(p.seed, vh_make_id(t, i + 1), step, VH_KEY_BIRTH));
    ws[slot - 1] = (uint64_t)VH_S | ((uint64_t)__float_as_uint(ns) << 32);
    died[slot - 1] = 0;
}

// this step's deaths onto the reuse stack in slot order (src/AgentMethods.jl:314-345)
__global__ void k_contact_deaths(uint64_t n, ContactBits B, const uint32_t* __restrict__ dpos,
                                 uint32_t* __restrict__ reuse, uint64_t base, uint8_t* __restrict__ died) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t word = B.sp[i >> 6] & ~B.dead[i >> 6];   // dying in this step
    uint32_t lane = (uint32_t)(i & 63);
    if (!((word >> lane) & 1ull)) return;
    uint64_t g = dpos[i >> 6] + (uint64_t)__popcll(word & ((1ull << lane) - 1ull));
    reuse[base + g] = (uint32_t)(i + 1);
    died[i] = 1;
}

// rows past the last caller (fresh newborn slots) are empty
__global__ void k_contact_extend(uint32_t* __restrict__ off, uint64_t from_row, uint64_t rows) {
    uint64_t r = from_row + 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= rows) off[r] = off[from_row];
}

// The rebuilt contact CSR (row i = caller slot i, kept partners in draw
// order): the block's log rows are read as one contiguous range, compacted
// in LDS and written as the block's contiguous CSR range. Dynamic LDS:
// 2 * 256 * n_contacts words when `staged`.
__global__ __launch_bounds__(256) void k_contact_edges(uint64_t n, uint32_t C,
                                                       const uint32_t* __restrict__ kept,
                                                       const uint32_t* __restrict__ log,
                                                       const uint32_t* __restrict__ off,
                                                       uint32_t* __restrict__ src, uint32_t staged) {
    extern __shared__ u32x4 lds_v[];   // 16-byte aligned
    uint32_t* lds = (uint32_t*)lds_v;
    const uint32_t tid = threadIdx.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * 256;
    const uint64_t i = i0 + tid;
    if (!staged) {   // uniform
        if (i >= n) return;
        uint32_t k = kept[i], o = off[i];
        for (uint32_t r = 0; r < k; r++) src[o + r] = log[i * C + r];
        return;
    }
    uint32_t* lin = lds;
    uint32_t* lout = lds + 256 * C;
    const uint64_t rows_here = n - i0 < 256 ? n - i0 : 256;
    const uint64_t len = rows_here * C;
    const uint32_t* lsrc = log + i0 * C;   // 16-byte aligned: i0 is a multiple of 256
    const uint64_t len4 = len >> 2;
    for (uint64_t q = tid; q < len4; q += 256)
        *(u32x4*)(lin + 4 * q) = __builtin_nontemporal_load((const u32x4*)(lsrc + 4 * q));
    for (uint64_t q = 4 * len4 + tid; q < len; q += 256) lin[q] = lsrc[q];
    const uint32_t base = off[i0], end = off[i0 + rows_here];
    __syncthreads();
    if (i < n) {
        uint32_t k = kept[i], o = off[i] - base;
        for (uint32_t r = 0; r < k; r++) lout[o + r] = lin[tid * C + r];
    }
    __syncthreads();
    for (uint32_t q = tid; q < end - base; q += 256) __builtin_nontemporal_store(lout[q], src + base + q);
}

// ----------------------------------------------------------------- raster stencils
// Cell i of a raster (column-major) has coordinates c; its in-edges over a
// stencil come from org = c - off[k] (wrapped per dimension when periodic,
// else only when inside the raster), one per stencil offset k, ordered as
// connect_raster_neighbors! inserts them: by org's index, then by k.
__device__ inline void cell_coords(const StencilSpec& sp, uint64_t i, int64_t* c) {
    for (uint32_t d = 0; d < sp.ndims; d++) {
        c[d] = (int64_t)(i % (uint64_t)sp.dims[d]);
        i /= (uint64_t)sp.dims[d];
    }
}
__device__ inline bool stencil_org(const StencilSpec& sp, const int64_t* c, uint32_t k, uint64_t& org) {
    uint64_t lin = 0, stride = 1;
    for (uint32_t d = 0; d < sp.ndims; d++) {
        int64_t o = c[d] - (int64_t)sp.off[k][d];
        if (o < 0 || o >= sp.dims[d]) {
            if (!sp.periodic) return false;
            o %= sp.dims[d];
            if (o < 0) o += sp.dims[d];
        }
        lin += (uint64_t)o * stride;
        stride *= (uint64_t)sp.dims[d];
    }
    org = lin;
    return true;
}
__global__ void k_stencil_degree(StencilSpec sp, uint64_t ncells, uint64_t rowbase, uint32_t* __restrict__ deg) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    int64_t c[4];
    cell_coords(sp, i, c);
    uint32_t n = 0;
    uint64_t org;
    for (uint32_t k = 0; k < sp.nst; k++) n += stencil_org(sp, c, k, org) ? 1u : 0u;
    deg[rowbase + i] = n;
}
__global__ void k_stencil_fill(StencilSpec sp, uint64_t ncells, uint64_t rowbase,
                               const uint32_t* __restrict__ off, void* __restrict__ src, uint32_t src64,
                               uint8_t* __restrict__ st, uint32_t bytes) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    int64_t c[4];
    cell_coords(sp, i, c);
    // keys (org << 8 | k), insertion-sorted; offsets are visited last to
    // first, so inside the raster (org falling with k) each key appends
    uint64_t key[VH_STENCIL_MAX];
    uint32_t n = 0;
    for (int k = (int)sp.nst - 1; k >= 0; k--) {
        uint64_t org;
        if (!stencil_org(sp, c, (uint32_t)k, org)) continue;
        const uint64_t kk = (org << 8) | (uint64_t)k;
        uint32_t j = n++;
        while (j > 0 && key[j - 1] > kk) {
            key[j] = key[j - 1];
            j--;
        }
        key[j] = kk;
    }
    const uint64_t o = off[rowbase + i];
    for (uint32_t j = 0; j < n; j++) {
        const uint64_t nr = sp.first_nr + (key[j] >> 8);
        if (src64) ((uint64_t*)src)[o + j] = vh_make_id(sp.cell_type, nr);
        else ((uint32_t*)src)[o + j] = (uint32_t)nr;
        for (uint32_t b = 0; b < bytes; b++) st[(o + j) * bytes + b] = sp.state[b];
    }
}

// diffuse (test/raster.jl diffuse) of every slot of the cell type: a slot is
// active when it is or some in-neighbour is; the in-neighbours of cell c are
// the cells c - offset of the stencil (wrapped when periodic), the sources
// k_stencil_fill gives its CSR row. Other slots of the type have no in-edges.
__global__ __launch_bounds__(256) void k_diffuse_stencil(StencilSpec sp, const uint8_t* __restrict__ rs,
                                                         uint8_t* __restrict__ ws, uint64_t nslots, uint64_t ncells,
                                                         uint32_t bytes, uint32_t f) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    const uint8_t* r = rs + i * bytes;
    uint8_t* w = ws + i * bytes;
    for (uint32_t k = 0; k < bytes; k++) w[k] = r[k];
    uint8_t v = r[f] ? 1 : 0;
    const uint64_t nr = i + 1;
    if (!v && nr >= sp.first_nr && nr - sp.first_nr < ncells) {
        if (ncells <= 0x7FFFFFFFull) {
            // 32-bit coordinates; a wrap needs one add or subtract while the
            // offset is at most the extent (else the general path below)
            int32_t c[4], dim[4];
            uint32_t rest = (uint32_t)(nr - sp.first_nr);
            for (uint32_t d = 0; d < sp.ndims; d++) {
                dim[d] = (int32_t)sp.dims[d];
                c[d] = (int32_t)(rest % (uint32_t)dim[d]);
                rest /= (uint32_t)dim[d];
            }
            for (uint32_t k = 0; k < sp.nst && !v; k++) {
                uint32_t lin = 0, stride = 1;
                bool in = true;
                for (uint32_t d = 0; d < sp.ndims; d++) {
                    int32_t o = c[d] - (int32_t)sp.off[k][d];
                    if (o < 0 || o >= dim[d]) {
                        if (!sp.periodic) {
                            in = false;
                            break;
                        }
                        o = o < 0 ? o + dim[d] : o - dim[d];
                        if (o < 0 || o >= dim[d]) {
                            o %= dim[d];
                            if (o < 0) o += dim[d];
                        }
                    }
                    lin += (uint32_t)o * stride;
                    stride *= (uint32_t)dim[d];
                }
                if (in && rs[(sp.first_nr - 1 + lin) * bytes + f]) v = 1;
            }
        } else {
            int64_t c[4];
            cell_coords(sp, nr - sp.first_nr, c);
            for (uint32_t k = 0; k < sp.nst && !v; k++) {
                uint64_t org;
                if (stencil_org(sp, c, k, org) && rs[(sp.first_nr - 1 + org) * bytes + f]) v = 1;
            }
        }
    }
    w[f] = v;
}

}  // namespace

// ====================================================================== launchers
namespace vh {

int launch_generic(int tr, const DCtx* d_ctx, int ci, int call_t, uint64_t nslots, int pass,
                   int with_edge, hipStream_t st) {
    if (!nslots) return VH_OK;
    if (tr >= (int)VH_TR_USER) return launch_user_body((uint32_t)tr, d_ctx, ci, call_t, nslots, pass, with_edge, st);
    dim3 g(grid_for(nslots, 256)), b(256);
    switch (tr) {
#define CASE(X)                                                                            \
    case X:                                                                                \
        hipLaunchKernelGGL(k_generic<X>, g, b, 0, st, d_ctx, ci, call_t, pass, with_edge); \
        break;
        CASE(VH_TR_NOOP)
        CASE(VH_TR_SUM_NEIGHBORS)
        CASE(VH_TR_DIE_IF)
        CASE(VH_TR_ADD_SELF_EDGE)
        CASE(VH_TR_READD_EDGES)
        CASE(VH_TR_NEIGHBOR_FOO_EDGES)
        CASE(VH_TR_SPAWN)
        CASE(VH_TR_COUNT_EDGES)
        CASE(VH_TR_DIFFUSE)
        CASE(VH_TR_REMOVE_EDGES)
        CASE(VH_TR_MOVE_TO)
#undef CASE
    default:
        return set_error(VH_ERR_UNREGISTERED, "transition %d has no generic kernel", tr);
    }
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

static int check_model_layout(Sim& s, int t, int et, uint32_t abytes, uint32_t ebytes, const char* who) {
    AgentStore& A = s.a[t];
    EdgeStore& E = s.e[et];
    if (A.bytes != abytes) return set_error(VH_ERR_ARG, "%s: agent state must be %u bytes", who, abytes);
    if (E.bytes != ebytes) return set_error(VH_ERR_ARG, "%s: edge state must be %u bytes", who, ebytes);
    if (VH_HAS(E.hints, VH_HINT_IGNOREFROM))
        return set_error(VH_ERR_ARG, "%s: the edge type must keep source ids", who);
    if (E.rows && E.nedges && (E.src64 || (int)E.src_type != t))
        return set_error(VH_ERR_ARG, "%s: edges must come from the called agent type only "
                                     "(one-agent-type model)", who);
    return VH_OK;
}


// Out-edges of E's rows (one agent type, rows from 0, 1-based sources up to
// nsrc): the transpose, by a stable radix sort of the in-edges by source;
// stamps array of nsrc entries, cleared once
static int build_out_csr(Sim& s, EdgeStore& E, uint64_t nsrc) {
    if (E.out_version == E.version && E.expo.p && E.expo.cap >= ((nsrc + 31) / 32) * 4 + 16) return VH_OK;
    hipStream_t st = s.stream;
    const uint64_t ne = E.nedges, rows = E.rows;
    uint32_t* keys = (uint32_t*)s.scratch.take(ne * 4 + 4);
    uint32_t* vals = (uint32_t*)s.scratch.take(ne * 4 + 4);
    uint32_t* ka = (uint32_t*)s.scratch.take(ne * 4 + 4);
    uint32_t* va = (uint32_t*)s.scratch.take(ne * 4 + 4);
    if (!keys || !vals || !ka || !va) return set_error(VH_ERR_OOM, "scratch");
    hipLaunchKernelGGL(k_edge_rows, dim3(grid_for(rows, 256)), dim3(256), 0, st, E.off.as<uint32_t>(),
                       E.src.as<uint32_t>(), rows, keys, vals);
    uint32_t bits = 1;
    while (bits < 32 && (nsrc >> bits) != 0) bits++;
    if (radix_sort_pairs(&keys, &vals, ka, va, ne, bits, false, s.scratch, st))
        return set_error(VH_ERR_DEVICE, "radix sort failed");
    RC(E.out_off.ensure((nsrc + 1) * 4 + 8, st));
    RC(E.out_dst.ensure(ne * 4 + 8, st));
    hipLaunchKernelGGL(k_sorted_offsets, dim3(grid_for(nsrc + 1, 256)), dim3(256), 0, st, keys, ne, nsrc,
                       E.out_off.as<uint32_t>());
    HIPCHECK(hipMemcpyAsync(E.out_dst.p, vals, ne * 4, hipMemcpyDeviceToDevice, st));
    RC(E.expo.ensure(((nsrc + 31) / 32) * 4 + 16, st));
    RC(E.expo_cnt.ensure(SIR_CTL_WORDS * 4 + 8, st));
    fill_u32(E.expo_cnt.as<uint32_t>(), 0u, SIR_CTL_WORDS, st);   // no complete counts yet
    HIPCHECK(hipGetLastError());
    E.out_version = E.version;
    return VH_OK;
}

int launch_sir(Sim& s, int t, int et, const vh_sir_params& p, uint64_t step) {
    RC(check_model_layout(s, t, et, 4, 0, "SIR"));
    AgentStore& A = s.a[t];
    EdgeStore& E = s.e[et];
    if (!A.writeable) return set_error(VH_ERR_ASSERT, "SIR: the agent type must be in `write`");
    if (!A.nslots) return VH_OK;
    // the exposure pass needs the in-edges of `t` only (rows from 0, sources
    // of `t`); it runs for at least sir_expose_min agents (default 65536:
    // below, its three extra launches per step cost more than the draws it
    // saves) unless the options turn it off
    uint32_t* expo = nullptr;
    uint32_t* cnt = nullptr;
    const uint64_t nrows = E.rows ? E.nrows[t] : 0ull;
    if (nrows && E.nedges && !E.src64 && E.src_type == (uint32_t)t && E.rowbase[t] == 0 && nrows == E.rows &&
        A.nslots < 0xFFFFFFFFull && g_opts.sir_expose != 0 && (int64_t)A.nslots >= g_opts.sir_expose_min) {
        RC(build_out_csr(s, E, A.nslots));
        expo = E.expo.as<uint32_t>();
        cnt = E.expo_cnt.as<uint32_t>();
        fill_u32(expo, 0, (A.nslots + 31) / 32, s.stream);
        hipLaunchKernelGGL(k_sir_decide, dim3(1), dim3(64), 0, s.stream, cnt, step, A.nslots);
        hipLaunchKernelGGL(k_sir_expose, dim3(grid_for(A.nslots, 256)), dim3(256), 0, s.stream,
                           (const uint32_t*)A.rstate(), E.out_off.as<uint32_t>(), E.out_dst.as<uint32_t>(),
                           A.nslots, (const uint32_t*)cnt, expo);
    }
    hipLaunchKernelGGL(k_sir, dim3(grid_for(A.nslots, 256)), dim3(256), 0, s.stream,
                       E.off.as<uint32_t>(), E.src.as<uint32_t>(), (const uint32_t*)A.rstate(),
                       (uint32_t*)A.wstate(), A.nslots, nrows, E.rowbase[t],
                       (uint32_t)t, p.p_inf, p.seed, p.recovery, step, (const uint32_t*)expo, cnt);
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

// Rows of degree > 64 (heavy), ascending; built once per CSR version
int build_heavy_bins(Sim& s, EdgeStore& E, int t) {
    if (E.bins_version == E.version) return VH_OK;
    const uint64_t nrows = E.nrows[t];
    RC(E.heavy.ensure(nrows * 4 + 8, s.stream));
    uint32_t* cnt = (uint32_t*)s.d_stats.p;
    fill_u32(cnt, 0, 1, s.stream);
    if (nrows)
        hipLaunchKernelGGL(k_heavy_rows, dim3(grid_for(nrows, 256)), dim3(256), 0, s.stream,
                           E.off.as<uint32_t>(), E.rows, E.rowbase[t], nrows, E.heavy.as<uint32_t>(), cnt);
    uint32_t h = 0;
    HIPCHECK(hipMemcpyAsync(&h, cnt, 4, hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    if (h > 1) {   // the atomic compaction's order varies between runs: sort it
        std::vector<uint32_t> rows(h);
        HIPCHECK(hipMemcpy(rows.data(), E.heavy.p, (size_t)h * 4, hipMemcpyDeviceToHost));
        std::sort(rows.begin(), rows.end());
        HIPCHECK(hipMemcpy(E.heavy.p, rows.data(), (size_t)h * 4, hipMemcpyHostToDevice));
    }
    E.nheavy = h;
    E.bins_version = E.version;
    return VH_OK;
}

// The gather plan of the in-edges of E (one target type at row 0, n agents,
// sources 1-based numbers <= n < 2^32) for states of `tb` bytes and windows
// of `wsz` sources: the plan's local source indices (psrc) and windows
// (wstart); the blocks (ablk, lbase), their light lists (tpk, twt), runs
// (rbase, prun, pfirst) and chunk run counts (cbase); the heavy rows' chunk
// lists (hbase, hty, hslot, hinfo). E.plan_ok stays false when a sort key
// would need more than 32 bits (the one-pass kernel runs).
int build_gather_plan(Sim& s, EdgeStore& E, uint32_t wsz, uint64_t n, uint32_t tb) {
    if (E.plan_version == E.version && E.plan_wsz == wsz && E.plan_n == n && E.plan_tb == tb) return VH_OK;
    hipStream_t st = s.stream;
    E.plan_ok = false;
    E.plan_version = E.version;
    E.plan_wsz = wsz;
    E.plan_n = n;
    E.plan_tb = tb;
    const uint64_t ne = E.nedges, nrows = E.rows;
    const uint32_t TB = tb == 8 ? WinCfg<double>::TB : WinCfg<float>::TB;
    const uint32_t CPB = TB / 64, S = TB - WIN_HMAX - 1;
    const uint32_t chunk = tb == 8 ? WinCfg<double>::CHUNK : WinCfg<float>::CHUNK;
    const uint64_t W = (n + wsz - 1) / wsz;
    if (!ne || ne + n >= 0xFFFFFFFFull || W > WIN_RUNS || nrows > n) return VH_OK;
    const uint32_t* off = E.off.as<uint32_t>();
    const uint32_t* src = E.src.as<uint32_t>();
    // blocks
    uint32_t* C = (uint32_t*)s.scratch.take((n + 1) * 4);
    uint32_t* cost = (uint32_t*)s.scratch.take(n * 4 + 4);
    if (!C || !cost) return set_error(VH_ERR_OOM, "scratch");
    hipLaunchKernelGGL(k_win_cost, dim3(grid_for(n, 256)), dim3(256), 0, st, off, n, nrows, cost);
    if (scan_excl_u32(cost, C, n, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
    uint32_t cl[2] = {0, 0};
    HIPCHECK(hipMemcpyAsync(cl, C + n - 1, 8, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const uint32_t nb = cl[0] / S + 1;
    const uint64_t nl = (uint64_t)cl[1] - n;
    if (W * (nb + 1ull) > 0xFFFFFFFFull) return VH_OK;
    // the build's pieces in one block (the sorts take many), released at the
    // next reset when large
    const uint64_t nh_e = ne - nl;
    if (s.scratch.room(24 * ne + 28 * nl + 22 * nh_e + 32 * (uint64_t)nb + ne / 4 + ((uint64_t)256 << 20)))
        return set_error(VH_ERR_OOM, "scratch for the gather plan");
    s.scratch.shrink = true;
    uint32_t kbits = 1, lbits = 1;
    while ((1ull << kbits) < W * (nb + 1ull)) kbits++;
    while ((1ull << lbits) < W * (uint64_t)nb) lbits++;
    RC(E.ablk.ensure(((uint64_t)nb + 1) * 4 + 8, st));
    RC(E.lbase.ensure(((uint64_t)nb + 1) * 4 + 8, st));
    hipLaunchKernelGGL(k_win_blocks, dim3(grid_for(n + 1, 256)), dim3(256), 0, st, C, n, S, nb,
                       E.ablk.as<uint32_t>(), E.lbase.as<uint32_t>());
    // plan order: stable sort of the CSR slots by (window, block or heavy)
    RC(E.psrc.ensure(ne * 2 + 16, st));
    RC(E.wstart.ensure((W + 1) * 4 + 8, st));
    uint32_t* pinv = (uint32_t*)s.scratch.take(ne * 4 + 4);
    {
        uint32_t* keys = (uint32_t*)s.scratch.take(ne * 4);
        uint32_t* vals = (uint32_t*)s.scratch.take(ne * 4);
        uint32_t* ka = (uint32_t*)s.scratch.take(ne * 4);
        uint32_t* va = (uint32_t*)s.scratch.take(ne * 4);
        if (!pinv || !keys || !vals || !ka || !va) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_win_keys, dim3(grid_for(nrows, 256)), dim3(256), 0, st, off, src, nrows, C, S, nb, wsz,
                           keys);
        if (radix_sort_pairs(&keys, &vals, ka, va, ne, kbits, true, s.scratch, st))
            return set_error(VH_ERR_DEVICE, "plan sort");
        hipLaunchKernelGGL(k_win_fill, dim3(grid_for(ne, 256)), dim3(256), 0, st, src, keys, vals, ne, nb, wsz,
                           (uint32_t)W, E.psrc.as<uint16_t>(), pinv, E.wstart.as<uint32_t>());
        HIPCHECK(hipGetLastError());
    }
    // light lists: each block's light in-edges by (window, slot)
    RC(E.tpk.ensure(nl * 2 + 8, st));
    RC(E.twt.ensure(nl * 4 + 8, st));
    RC(E.rbase.ensure(((uint64_t)nb + 1) * 4 + 8, st));
    RC(E.cbase.ensure((uint64_t)nb * 256 * 2 + 64, st));
    const float* w4 = E.bytes == 4 ? E.st.as<float>() : nullptr;
    {
        uint32_t* ridx = (uint32_t*)s.scratch.take((nl + 1) * 4);
        if (!ridx) return set_error(VH_ERR_OOM, "scratch");
        uint32_t nruns = 0;
        if (nl) {
            uint32_t* keys = (uint32_t*)s.scratch.take(nl * 4 + 4);
            uint32_t* vals = (uint32_t*)s.scratch.take(nl * 4 + 4);
            uint32_t* ka = (uint32_t*)s.scratch.take(nl * 4 + 4);
            uint32_t* va = (uint32_t*)s.scratch.take(nl * 4 + 4);
            uint32_t* kidx = (uint32_t*)s.scratch.take(nl * 4 + 4);
            uint32_t* flag = (uint32_t*)s.scratch.take(nl * 4 + 4);
            if (!keys || !vals || !ka || !va || !kidx || !flag) return set_error(VH_ERR_OOM, "scratch");
            hipLaunchKernelGGL(k_win_lkeys, dim3(grid_for(nrows, 256)), dim3(256), 0, st, off, src, nrows, C, S,
                               (uint32_t)W, wsz, keys, kidx);
            if (radix_sort_pairs(&keys, &vals, ka, va, nl, lbits, true, s.scratch, st))
                return set_error(VH_ERR_DEVICE, "list sort");
            hipLaunchKernelGGL(k_win_starts, dim3(grid_for(nl, 256)), dim3(256), 0, st, keys, nl, flag);
            if (scan_excl_u32(flag, ridx, nl, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
            HIPCHECK(hipMemcpyAsync(&nruns, ridx + nl, 4, hipMemcpyDeviceToHost, st));
            HIPCHECK(hipStreamSynchronize(st));
            RC(E.prun.ensure((uint64_t)nruns * 4 + 8, st));
            RC(E.pfirst.ensure((uint64_t)nruns * 2 + 8, st));
            hipLaunchKernelGGL(k_win_pack, dim3(grid_for(nl, 256)), dim3(256), 0, st, keys, vals, kidx, flag, ridx,
                               pinv, E.lbase.as<uint32_t>(), (uint32_t)W, nl, w4, E.tpk.as<uint16_t>(),
                               E.twt.as<float>(), E.prun.as<uint32_t>(), E.pfirst.as<uint16_t>());
        } else {
            fill_u32(ridx, 0, 1, st);
            RC(E.prun.ensure(8, st));
            RC(E.pfirst.ensure(8, st));
        }
        hipLaunchKernelGGL(k_win_rbase, dim3(grid_for(((uint64_t)nb + 1) * CPB, 256)), dim3(256), 0, st,
                           E.lbase.as<uint32_t>(), ridx, nb, CPB, E.rbase.as<uint32_t>(), E.cbase.as<uint16_t>());
        HIPCHECK(hipGetLastError());
    }
    RC(E.wq.ensure(16, st));
    fill_u32(E.wq.as<uint32_t>(), 0, 3, st);   // queue (2) and the check flag
    {
        uint32_t nruns = 0, bad = 0;
        HIPCHECK(hipMemcpyAsync(&nruns, E.rbase.as<uint32_t>() + nb, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        hipLaunchKernelGGL(k_win_check, dim3(grid_for(nb, 64)), dim3(64), 0, st, E.ablk.as<uint32_t>(),
                           E.lbase.as<uint32_t>(), E.rbase.as<uint32_t>(), E.prun.as<uint32_t>(),
                           E.pfirst.as<uint16_t>(), E.tpk.as<uint16_t>(), E.cbase.as<uint16_t>(), nb, n, nl,
                           (uint64_t)nruns, ne, TB, CPB, E.wq.as<uint32_t>() + 2);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(&bad, E.wq.as<uint32_t>() + 2, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        if (bad) return set_error(VH_ERR_DEVICE, "gather plan inconsistent");
    }
    // rows of in-degree above WIN_HMAX (from E.heavy, ascending): chunk lists
    // sorted by (chunk, window)
    uint64_t nheavy = 0;
    if (E.nheavy) {
        uint32_t* hd = (uint32_t*)s.scratch.take(E.nheavy * 4 + 4);
        uint32_t* hc = (uint32_t*)s.scratch.take(E.nheavy * 4 + 4);
        if (!hd || !hc) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_heavy_sizes, dim3(grid_for(E.nheavy, 256)), dim3(256), 0, st, off,
                           E.heavy.as<uint32_t>(), E.nheavy, chunk, hd, hc);
        std::vector<uint32_t> rows(E.nheavy), deg(E.nheavy);
        HIPCHECK(hipMemcpyAsync(rows.data(), E.heavy.p, E.nheavy * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(deg.data(), hd, E.nheavy * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<uint32_t> huge;
        for (uint64_t h = 0; h < E.nheavy; h++)
            if (deg[h] > WIN_HMAX) huge.push_back(rows[h]);
        nheavy = huge.size();
        if (nheavy) {
            RC(E.huge.ensure(nheavy * 4 + 8, st));
            HIPCHECK(hipMemcpyAsync(E.huge.p, huge.data(), nheavy * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipStreamSynchronize(st));
        }
    }
    E.plan_nhuge = nheavy;
    if (nheavy) {
        uint32_t* hd = (uint32_t*)s.scratch.take(nheavy * 4 + 4);
        uint32_t* hc = (uint32_t*)s.scratch.take(nheavy * 4 + 4);
        uint32_t* hcb = (uint32_t*)s.scratch.take(nheavy * 4 + 8);
        if (!hd || !hc || !hcb) return set_error(VH_ERR_OOM, "scratch");
        RC(E.hbase.ensure(nheavy * 4 + 8, st));
        hipLaunchKernelGGL(k_heavy_sizes, dim3(grid_for(nheavy, 256)), dim3(256), 0, st, off,
                           E.huge.as<uint32_t>(), nheavy, chunk, hd, hc);
        if (scan_excl_u32(hd, E.hbase.as<uint32_t>(), nheavy, s.scratch, st) ||
            scan_excl_u32(hc, hcb, nheavy, s.scratch, st))
            return set_error(VH_ERR_DEVICE, "scan");
        uint32_t htot[2] = {0, 0};
        HIPCHECK(hipMemcpyAsync(&htot[0], E.hbase.as<uint32_t>() + nheavy, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipMemcpyAsync(&htot[1], hcb + nheavy, 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        const uint32_t nhe = htot[0], nch = htot[1];
        uint32_t cbits = 1, wbits = 1;
        while ((1ull << cbits) < nch) cbits++;
        while ((1ull << wbits) < W) wbits++;
        if (cbits + wbits > 32 || nl + nhe != ne) return VH_OK;
        RC(E.hty.ensure((uint64_t)nhe * 4 + 8, st));
        RC(E.hslot.ensure((uint64_t)nhe * 2 + 8, st));
        uint32_t* keys = (uint32_t*)s.scratch.take((uint64_t)nhe * 4);
        uint32_t* vals = (uint32_t*)s.scratch.take((uint64_t)nhe * 4);
        uint32_t* ka = (uint32_t*)s.scratch.take((uint64_t)nhe * 4);
        uint32_t* va = (uint32_t*)s.scratch.take((uint64_t)nhe * 4);
        uint32_t* kidx = (uint32_t*)s.scratch.take((uint64_t)nhe * 4);
        uint16_t* tsl = (uint16_t*)s.scratch.take((uint64_t)nhe * 2);
        if (!keys || !vals || !ka || !va || !kidx || !tsl) return set_error(VH_ERR_OOM, "scratch");
        hipLaunchKernelGGL(k_heavy_keys, dim3(grid_for(nheavy * 64, 256)), dim3(256), 0, st, off, src,
                           E.huge.as<uint32_t>(), nheavy, E.hbase.as<uint32_t>(), hcb, wsz, wbits, chunk, keys,
                           kidx, tsl);
        if (radix_sort_pairs(&keys, &vals, ka, va, nhe, cbits + wbits, true, s.scratch, st))
            return set_error(VH_ERR_DEVICE, "chunk sort");
        hipLaunchKernelGGL(k_heavy_fill, dim3(grid_for(nhe, 256)), dim3(256), 0, st, vals, kidx, tsl, pinv,
                           (uint64_t)nhe, E.hty.as<uint32_t>(), E.hslot.as<uint16_t>());
        RC(E.hinfo.ensure(nheavy * 16 + 16, st));
        hipLaunchKernelGGL(k_heavy_info, dim3(grid_for(nheavy, 256)), dim3(256), 0, st, off,
                           E.huge.as<uint32_t>(), nheavy, E.hbase.as<uint32_t>(), E.hinfo.as<uint4>());
        HIPCHECK(hipGetLastError());
        // longest rows first: pass 2's waves take them round-robin, so every
        // wave's first row is among the longest
        std::vector<uint4> hi(nheavy);
        HIPCHECK(hipMemcpyAsync(hi.data(), E.hinfo.p, nheavy * 16, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        std::stable_sort(hi.begin(), hi.end(), [](const uint4& a, const uint4& b) { return a.z > b.z; });
        HIPCHECK(hipMemcpyAsync(E.hinfo.p, hi.data(), nheavy * 16, hipMemcpyHostToDevice, st));
        HIPCHECK(hipStreamSynchronize(st));
    } else if (nl != ne) {
        return VH_OK;
    }
    E.plan_nb = nb;
    E.plan_w = (uint32_t)W;
    E.plan_ok = true;
    return VH_OK;
}

// compute units of the current device (the persistent plan kernels run one
// block per CU)
static uint32_t device_cus() {
    static uint32_t cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) {
            (void)hipGetLastError();
            v = 256;
        }
        cus[dev] = (uint32_t)v;
    }
    return cus[dev];
}

// the degree bins and, when the step takes the two-pass gather, its plan and
// gathered-value buffer for an opinion step of agent type t over edge type et
// (kept while the edges and the state type stay the same); `plan` tells which
static int opinion_plan(Sim& s, int t, int et, bool f32, bool& plan) {
    AgentStore& A = s.a[t];
    EdgeStore& E = s.e[et];
    const uint32_t tb = f32 ? 4 : 8;
    const uint64_t nrows = E.rows ? E.nrows[t] : 0;
    if (E.rows) RC(build_heavy_bins(s, E, t));
    // two-pass gather through a plan from a state vector of at least
    // opinion_plan_bytes (< 0: never), with source windows of
    // opinion_part_bytes (0 or more than 160 KB: 160 KB; widened to make at
    // most WIN_RUNS windows, a multiple of 4 sources)
    const int64_t plan_min = g_opts.opinion_plan_bytes;
    plan = plan_min >= 0 && E.rows && E.nedges && !E.src64 && E.src_type == (uint32_t)t &&
           E.rowbase[t] == 0 && nrows == E.rows && A.nslots < 0xFFFFFFFFull &&
           (uint64_t)A.nslots * tb >= (uint64_t)plan_min;
    if (plan) {
        const uint64_t cap = f32 ? WinCfg<float>::WSZ : WinCfg<double>::WSZ;
        uint64_t wsz = g_opts.opinion_part_bytes <= 0 ? cap : (uint64_t)g_opts.opinion_part_bytes / tb;
        if (wsz > cap) wsz = cap;
        const uint64_t wmin = (A.nslots + WIN_RUNS - 1) / WIN_RUNS;
        if (wsz < wmin) wsz = wmin;
        wsz = (wsz + 3) / 4 * 4;
        if (wsz > cap) plan = false;   // more sources than WIN_RUNS windows hold
        if (plan) {
            RC(build_gather_plan(s, E, (uint32_t)wsz, A.nslots, tb));
            plan = E.plan_ok;
        }
    }
    if (plan) RC(E.pval.ensure(E.nedges * tb + 16, s.stream));
    return VH_OK;
}

int launch_opinion(Sim& s, int t, int et, const vh_opinion_params& p, bool f32, uint64_t step) {
    (void)step;
    RC(check_model_layout(s, t, et, f32 ? 4 : 8, 4, "opinion"));
    AgentStore& A = s.a[t];
    EdgeStore& E = s.e[et];
    if (!A.writeable) return set_error(VH_ERR_ASSERT, "opinion: the agent type must be in `write`");
    if (!A.nslots) return VH_OK;
    const uint64_t nrows = E.rows ? E.nrows[t] : 0;
    bool plan = false;
    RC(opinion_plan(s, t, et, f32, plan));
    const uint64_t nheavy = E.rows ? E.nheavy : 0;
    const uint32_t nlight = grid_for(A.nslots, 256);
    const uint32_t nh = (uint32_t)((nheavy + 3) / 4);
    if (!plan && g_opinion_defer && g_opinion_defer->stream == s.stream) {
        OpMember om;
        om.off = E.off.as<uint32_t>();
        om.src = E.src.as<uint32_t>();
        om.x = A.rstate();
        om.w = E.st.as<float>();
        om.xo = A.wstate();
        om.heavy = E.heavy.as<uint32_t>();
        om.n = A.nslots;
        om.nrows = nrows;
        om.rowbase = E.rowbase[t];
        om.nheavy = nheavy;
        om.eps = f32 ? (double)(float)p.eps : p.eps;
        om.nh = nh;
        om.blk0 = nlight + nh;   // (the block count until the launch assigns offsets)
        g_opinion_defer->m[f32 ? 1 : 0].push_back(om);
        return VH_OK;
    }
    if (!plan) {
        dim3 g(nlight + nh), b(256);
        if (f32)
            hipLaunchKernelGGL(k_opinion<float>, g, b, 0, s.stream, E.off.as<uint32_t>(), E.src.as<uint32_t>(),
                               (const float*)A.rstate(), E.st.as<float>(), (float*)A.wstate(), A.nslots, nrows,
                               E.rowbase[t], (float)p.eps, E.heavy.as<uint32_t>(), nheavy, nh);
        else
            hipLaunchKernelGGL(k_opinion<double>, g, b, 0, s.stream, E.off.as<uint32_t>(), E.src.as<uint32_t>(),
                               (const double*)A.rstate(), E.st.as<float>(), (double*)A.wstate(), A.nslots, nrows,
                               E.rowbase[t], p.eps, E.heavy.as<uint32_t>(), nheavy, nh);
        HIPCHECK(hipGetLastError());
        return VH_OK;
    }
    const dim3 g(device_cus()), bb(1024);
    if (f32) {
        hipLaunchKernelGGL(k_win_gather<float>, g, bb, 0, s.stream, E.psrc.as<uint16_t>(), (const float*)A.rstate(),
                           E.pval.as<float>(), E.wstart.as<uint32_t>(), E.plan_w, E.plan_wsz, A.nslots, E.nedges);
        hipLaunchKernelGGL(k_win_sum<float>, g, bb, 0, s.stream, E.off.as<uint32_t>(), (const float*)A.rstate(),
                           E.pval.as<float>(), E.st.as<float>(), (float*)A.wstate(), nrows, (float)p.eps,
                           E.hinfo.as<uint4>(), E.plan_nhuge, E.hty.as<uint32_t>(), E.hslot.as<uint16_t>(),
                           E.ablk.as<uint32_t>(), E.lbase.as<uint32_t>(), E.tpk.as<uint16_t>(), E.twt.as<float>(),
                           E.rbase.as<uint32_t>(), E.prun.as<uint32_t>(), E.pfirst.as<uint16_t>(),
                           E.cbase.as<uint16_t>(), E.plan_nb, E.wq.as<uint32_t>());
    } else {
        hipLaunchKernelGGL(k_win_gather<double>, g, bb, 0, s.stream, E.psrc.as<uint16_t>(),
                           (const double*)A.rstate(), E.pval.as<double>(), E.wstart.as<uint32_t>(), E.plan_w,
                           E.plan_wsz, A.nslots, E.nedges);
        hipLaunchKernelGGL(k_win_sum<double>, g, bb, 0, s.stream, E.off.as<uint32_t>(), (const double*)A.rstate(),
                           E.pval.as<double>(), E.st.as<float>(), (double*)A.wstate(), nrows, p.eps,
                           E.hinfo.as<uint4>(), E.plan_nhuge, E.hty.as<uint32_t>(), E.hslot.as<uint16_t>(),
                           E.ablk.as<uint32_t>(), E.lbase.as<uint32_t>(), E.tpk.as<uint16_t>(), E.twt.as<float>(),
                           E.rbase.as<uint32_t>(), E.prun.as<uint32_t>(), E.pfirst.as<uint16_t>(),
                           E.cbase.as<uint16_t>(), E.plan_nb, E.wq.as<uint32_t>());
    }
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

// the plan of an opinion step built ahead of its first apply!, on the
// simulation's stream, which the call waits for; the build's scratch is
// released when large
extern "C" int vh_opinion_prepare(vh_sim_t h, uint8_t t, uint8_t et, int f32) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    Sim& s = h->s;
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "opinion_prepare after finish_init! only");
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    RC(check_model_layout(s, t, et, f32 ? 4 : 8, 4, "opinion"));
    if (!s.a[t].nslots) return VH_OK;
    s.scratch.reset(s.stream);
    bool plan = false;
    RC(opinion_plan(s, t, et, f32 != 0, plan));
    s.scratch.reset(s.stream);
    HIPCHECK(hipStreamSynchronize(s.stream));
    return VH_OK;
}

thread_local OpinionDefer* g_opinion_defer = nullptr;

int opinion_defer_launch(OpinionDefer& c, hipStream_t st, DevBuf& dev, void** host, size_t* host_cap,
                         size_t* host_off) {
    for (int k = 0; k < 2; k++) {
        std::vector<OpMember>& v = c.m[k];
        if (v.empty()) continue;
        uint64_t blocks = 0;
        for (OpMember& om : v) {
            const uint32_t nb = om.blk0;   // the member's block count
            om.blk0 = (uint32_t)blocks;
            blocks += nb;
        }
        if (blocks > 0x7FFFFFFFull) return set_error(VH_ERR_CAPACITY, "batched opinion launch too large");
        const size_t bytes = v.size() * sizeof(OpMember);
        if (*host_off + bytes > *host_cap) return set_error(VH_ERR_CAPACITY, "batch staging");
        memcpy((uint8_t*)*host + *host_off, v.data(), bytes);
        HIPCHECK(hipMemcpyAsync((uint8_t*)dev.p + *host_off, (uint8_t*)*host + *host_off, bytes,
                                hipMemcpyHostToDevice, st));
        const OpMember* d = (const OpMember*)((uint8_t*)dev.p + *host_off);
        *host_off += (bytes + 255) & ~(size_t)255;
        if (k == 0)
            hipLaunchKernelGGL(k_opinion_multi<double>, dim3((uint32_t)blocks), dim3(256), 0, st, d,
                               (uint32_t)v.size());
        else
            hipLaunchKernelGGL(k_opinion_multi<float>, dim3((uint32_t)blocks), dim3(256), 0, st, d,
                               (uint32_t)v.size());
        HIPCHECK(hipGetLastError());
        v.clear();
    }
    return VH_OK;
}

int launch_diffuse_stencil(Sim& s, int t, int et, uint32_t field) {
    AgentStore& A = s.a[t];
    const StencilSpec& sp = s.e[et].sspec;
    uint64_t ncells = 1;
    for (uint32_t d = 0; d < sp.ndims; d++) ncells *= (uint64_t)sp.dims[d];
    if (!A.nslots) return VH_OK;
    hipLaunchKernelGGL(k_diffuse_stencil, dim3(grid_for(A.nslots, 256)), dim3(256), 0, s.stream, sp,
                       (const uint8_t*)A.rstate(), (uint8_t*)A.wstate(), A.nslots, ncells, A.bytes, field);
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

int stencil_csr(Sim& s, EdgeStore& E, uint64_t rowbase, uint64_t rows) {
    const StencilSpec& sp = E.sspec;
    hipStream_t st = s.stream;
    uint64_t ncells = 1;
    for (uint32_t d = 0; d < sp.ndims; d++) ncells *= (uint64_t)sp.dims[d];
    uint32_t* deg = (uint32_t*)s.scratch.take(rows * 4 + 8);
    if (!deg) return set_error(VH_ERR_OOM, "scratch");
    fill_u32(deg, 0, rows, st);
    hipLaunchKernelGGL(k_stencil_degree, dim3(grid_for(ncells, 256)), dim3(256), 0, st, sp, ncells, rowbase, deg);
    RC(E.off.ensure((rows + 1) * 4 + 8, st));
    if (scan_excl_u32(deg, E.off.as<uint32_t>(), rows, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
    HIPCHECK(hipMemcpyAsync(s.d_stats.p, E.off.as<uint32_t>() + rows, 4, hipMemcpyDeviceToDevice, st));
    RC(readback(s, 1));
    uint64_t ne = (uint32_t)s.h_stats[0];
    RC(E.src.ensure(ne * (E.src64 ? 8 : 4) + 8, st));
    if (E.bytes) RC(E.st.ensure(ne * E.bytes + 8, st));
    hipLaunchKernelGGL(k_stencil_fill, dim3(grid_for(ncells, 256)), dim3(256), 0, st, sp, ncells, rowbase,
                       E.off.as<uint32_t>(), E.src.p, (uint32_t)E.src64, E.st.as<uint8_t>(), E.bytes);
    HIPCHECK(hipGetLastError());
    E.nedges = ne;
    return VH_OK;
}

// One apply! of the contact model: flags, health update with the kept
// partners logged, one readback (deaths, contacts), newborns and deaths on
// the reuse stack, then the contact CSR rebuilt in place from the log (its
// previous entries were last read by k_contact_main). Deaths of this step
// are returned for the edge filtering of other edge types.
int contact_apply(Sim& s, int t, int et, const vh_contact_params& p, uint64_t step, uint64_t* ndeaths) {
    RC(check_model_layout(s, t, et, 8, 0, "contact"));
    AgentStore& A = s.a[t];
    EdgeStore& E = s.e[et];
    hipStream_t st = s.stream;
    const uint64_t n = A.nslots;
    const uint32_t C = p.n_contacts;
    *ndeaths = 0;
    if (n >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "contact: too many slots");
    if ((uint64_t)n * C >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "contact: too many contacts");
    // LDS staging of 256 rows of C contacts (main: one buffer, CSR copy: two)
    const uint32_t staged = C >= 1 && C <= 32;
    const uint64_t nw = (n + 63) / 64;
    ContactBits B;
    B.nwords = nw;
    B.sp = (uint64_t*)s.scratch.take(nw * 8 + 8);
    B.dead = (uint64_t*)s.scratch.take(nw * 8 + 8);
    B.inf = (uint64_t*)s.scratch.take(nw * 8 + 8);
    B.dcnt = (uint32_t*)s.scratch.take(nw * 4 + 8);
    uint32_t* dpos = (uint32_t*)s.scratch.take((nw + 1) * 4 + 8);
    uint32_t* kept = (uint32_t*)s.scratch.take(n * 4 + 8);
    uint32_t* log = (uint32_t*)s.scratch.take(n * C * 4 + 8);
    if (!B.sp || !B.dead || !B.inf || !B.dcnt || !dpos || !kept || !log) return set_error(VH_ERR_OOM, "scratch");
    uint64_t nd = 0, ne = 0;
    if (n) {
        hipLaunchKernelGGL(k_contact_flags, dim3(grid_for(nw * 64, 256)), dim3(256), 0, st,
                           (const uint64_t*)A.rstate(), A.died.as<uint8_t>(), n, (uint32_t)t, p, step, B);
        if (scan_excl_u32(B.dcnt, dpos, nw, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
        hipLaunchKernelGGL(k_contact_main, dim3(grid_for(n, 256)), dim3(256), staged ? 256 * C * 4 : 0, st,
                           E.rows ? E.off.as<uint32_t>() : nullptr, E.src.as<uint32_t>(),
                           (const uint64_t*)A.rstate(), (uint64_t*)A.wstate(), A.died_now.as<uint8_t>(), n,
                           E.rows ? E.nrows[t] : 0ull, E.rowbase[t], (uint32_t)t, p, step, B, kept, log,
                           staged);
        HIPCHECK(hipGetLastError());
        RC(E.off.ensure((n + 1) * 4 + 8, st));
        if (scan_excl_u32(kept, E.off.as<uint32_t>(), n, s.scratch, st))
            return set_error(VH_ERR_DEVICE, "scan");
        HIPCHECK(hipMemcpyAsync((uint32_t*)s.d_stats.p, dpos + nw, 4, hipMemcpyDeviceToDevice, st));
        HIPCHECK(hipMemcpyAsync((uint32_t*)s.d_stats.p + 1, E.off.as<uint32_t>() + n, 4,
                                hipMemcpyDeviceToDevice, st));
        RC(readback(s, 1));
        nd = ((const uint32_t*)s.h_stats)[0];
        ne = ((const uint32_t*)s.h_stats)[1];
    }
    if (nd) {
        uint64_t consumed = std::min<uint64_t>(nd, A.nreuse);
        uint64_t fresh = nd - consumed;
        if (fresh) {
            RC(ensure_slots(s, t, A.nextid - 1 + fresh));
            if (A.nextid - 1 + fresh > n)
                fill_u8(A.died_now.as<uint8_t>() + n, 0, A.nextid - 1 + fresh - n, st);
        }
        hipLaunchKernelGGL(k_contact_births, dim3(grid_for(n, 256)), dim3(256), 0, st, n, (uint32_t)t, p,
                           step, B, dpos, A.reuse.as<uint32_t>(), A.nreuse, A.nextid, (uint64_t*)A.wstate(),
                           A.died.as<uint8_t>());
        A.nreuse -= consumed;
        A.nextid += fresh;
        RC(A.reuse.ensure((A.nreuse + nd) * 4 + 8, st, A.nreuse * 4));
        hipLaunchKernelGGL(k_contact_deaths, dim3(grid_for(n, 256)), dim3(256), 0, st, n, B, dpos,
                           A.reuse.as<uint32_t>(), A.nreuse, A.died.as<uint8_t>());
        A.nreuse += nd;
        HIPCHECK(hipGetLastError());
    }
    // rows = caller slots; rows of dead callers and of newborns are empty
    uint64_t rows = std::max(n, A.nextid - 1);
    RC(E.off.ensure((rows + 1) * 4 + 8, st, (n + 1) * 4));
    if (!n) fill_u32(E.off.as<uint32_t>(), 0, 1, st);
    if (rows > n)
        hipLaunchKernelGGL(k_contact_extend, dim3(grid_for(rows - n, 256)), dim3(256), 0, st,
                           E.off.as<uint32_t>(), n, rows);
    RC(E.src.ensure(ne * 4 + 8, st));
    if (ne)
        hipLaunchKernelGGL(k_contact_edges, dim3(grid_for(n, 256)), dim3(256), staged ? 2 * 256 * C * 4 : 0,
                           st, n, C, kept, log, E.off.as<uint32_t>(), E.src.as<uint32_t>(), staged);
    HIPCHECK(hipGetLastError());
    for (int u = 0; u <= VH_MAXA; u++) {
        E.rowbase[u] = 0;
        E.nrows[u] = 0;
    }
    E.nrows[t] = rows;
    E.rows = rows;
    E.nedges = ne;
    E.src64 = false;
    E.src_type = (uint32_t)t;
    E.version++;
    *ndeaths = nd;
    return VH_OK;
}

}  // namespace vh
