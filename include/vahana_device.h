/* vahana_device.h — the device side of the transition backend for HIP
 * transition bodies: the views of the simulation state a body receives (DCtx),
 * the accessors of the reference's transition API (agentstate, the in-edges of
 * an agent, add_edge!, remove_edges!, add_agent!), the AgentID codec
 * (src/Agent.jl:30-118) and the counter-hash RNG.
 *
 * A body is a struct with
 *     __device__ static bool run(const DCtx* c, const uint8_t* state, uint64_t id,
 *                                Emitter& em, uint8_t* out);
 * called once per living agent of each call type (the agent loop of
 * src/AgentMethods.jl:193-268): `state` is the agent's state, `out` (a copy
 * of it) the state it returns, and `false` means `nothing` (the agent dies).
 * c->params holds the apply! parameter blob, c->role_e / c->role_a the edge
 * and agent types bound to the body's roles. VH_BODY_KERNEL(name, Body)
 * defines the kernel; compiled into a code object (hipcc --genco
 * --offload-arch=gfx950) it is registered at run time with vh_register_body
 * (include/vahana_hip.h), no rebuild of the library needed. Bodies that call
 * em.edge / em.remove / em.agent run twice per apply! (count, then write at
 * the scanned positions), so they must be deterministic.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "vahana_hip.h"

#define VH_MAXA VH_MAX_AGENT_TYPES
#define VH_MAXE VH_MAX_EDGE_TYPES
#define VH_MAX_STATE 64          // largest agent/edge state handled by the generic bodies
#define VH_PARAMS_BYTES 256

// device error word bits (reported as VH_ERR_ASSERT / VH_ERR_CAPACITY)
#define VH_DERR_READ_EDGE   (1u << 0)   // edge type accessed but not in `read`
#define VH_DERR_READ_AGENT  (1u << 1)   // agent type accessed but not in `read`
#define VH_DERR_DIED        (1u << 2)   // agentstate of a removed agent
#define VH_DERR_IMMORTAL    (1u << 3)   // `nothing` returned for an immortal type
#define VH_DERR_WRITE       (1u << 4)   // add_edge!/add_agent! of a type not in `write`
#define VH_DERR_CAP         (1u << 5)   // fixed per-caller emission capacity exceeded
#define VH_DERR_HINT        (1u << 6)   // accessor not defined for the hint combination
#define VH_DERR_TARGET      (1u << 7)   // edge target of the wrong type / invalid id
#define VH_DERR_ADDEX       (1u << 8)   // remove_edges! of a type not in `add_existing`

// ------------------------------------------------------------------ codec
__host__ __device__ __forceinline__ uint64_t vh_make_id(uint64_t type, uint64_t nr) {
    return (type << VH_SHIFT_TYPE) | nr;
}
__host__ __device__ __forceinline__ uint32_t vh_type_of(uint64_t id) {
    return (uint32_t)(id >> VH_SHIFT_TYPE);
}
__host__ __device__ __forceinline__ uint64_t vh_nr_of(uint64_t id) { return id & VH_NR_MASK; }

// ------------------------------------------------------------------ RNG
// SplitMix64 finaliser chained over (seed, a, b, c).
#define VH_GOLDEN 0x9E3779B97F4A7C15ull
__host__ __device__ __forceinline__ uint64_t vh_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// The first three stages depend only on (seed, a, b): a body drawing many
// `c` for one agent and step computes them once (vh_hash_pre) and finishes
// each draw with one stage (vh_hash_fin); the value is vh_hash's.
__host__ __device__ __forceinline__ uint64_t vh_hash_pre(uint64_t s, uint64_t a, uint64_t b) {
    uint64_t h = vh_mix64(s + VH_GOLDEN);
    h = vh_mix64(h + a + VH_GOLDEN);
    return vh_mix64(h + b + VH_GOLDEN);
}
__host__ __device__ __forceinline__ uint64_t vh_hash_fin(uint64_t pre, uint64_t c) {
    return vh_mix64(pre + c + VH_GOLDEN);
}
__host__ __device__ __forceinline__ uint64_t vh_hash(uint64_t s, uint64_t a, uint64_t b,
                                                     uint64_t c) {
    return vh_hash_fin(vh_hash_pre(s, a, b), c);
}
// uniform double in [0,1) from the top 53 bits
__host__ __device__ __forceinline__ double vh_u01(uint64_t h) {
    return (double)(h >> 11) * 0x1.0p-53;
}
// u01(h) < p exactly when (h >> 11) < vh_u01_below(p): u01's value is
// (h >> 11) * 2^-53 without rounding, so the test is (h >> 11) < p * 2^53,
// i.e. below the ceiling of p * 2^53; p <= 0 never passes, p >= 1 and NaN
// always do (`u01(h) >= p` is false for them). An integer compare per draw
// instead of a conversion to double.
__host__ __device__ __forceinline__ uint64_t vh_u01_below(double p) {
    if (!(p > 0.0)) return p != p ? ~0ull : 0ull;
    if (p >= 1.0) return ~0ull;
    return (uint64_t)ceil(p * 0x1.0p53);
}
// uniform float in [0,1) from the top 24 bits
__host__ __device__ __forceinline__ float vh_u01f(uint64_t h) {
    return (float)(h >> 40) * 0x1.0p-24f;
}

// ------------------------------------------------------------------ device views
// One registered agent type.  Slots are 1-based (src/AgentMethods.jl:37-63);
// slot s lives at index s-1 of every per-slot array.
struct DAgent {
    const uint8_t* rstate;   // read buffer (published state)
    uint8_t* wstate;         // write buffer (== rstate for :Independent)
    const uint8_t* died;     // died flags of the read state (mortal types)
    uint8_t* died_now;       // deaths recorded by this step
    uint64_t nslots;         // length(readstate)
    uint32_t bytes;
    uint32_t mortal, indep, readable, writeable;
    // births of this step (two-pass bodies): slot of birth g is
    // reuse[nreuse-1-g] for g < nreuse, else nextid + g - nreuse
    const uint32_t* reuse;
    uint64_t nreuse, nextid;
};

// One registered edge type: a CSR over a row space that concatenates the
// slots of every target agent type (row = rowbase[t] + nr - 1).
struct DEdge {
    const uint32_t* off;     // nrows_total + 1 offsets (null: no edges)
    const void* src;         // source ids: uint32 nr (src64 == 0) or uint64 AgentID
    const uint8_t* st;       // edge states, `bytes` each
    uint64_t rowbase[VH_MAXA + 1];
    uint64_t nrows[VH_MAXA + 1];
    uint32_t bytes, hints, src64, src_type;
    uint32_t readable, writeable, add_existing, pad1;
};

// Per-role emission logs of one apply! (add_edge! / add_agent! inside the body)
struct DEmitEdge {
    uint32_t* cnt;           // per caller (concatenated over call types)
    const uint64_t* offs;    // exclusive scan of cnt (pass 1)
    uint64_t* to;            // log, caller order
    uint64_t* from;
    uint8_t* st;
    int32_t et;              // bound edge type (-1 none)
    uint32_t bytes;
    uint8_t* kind;           // per log entry: VH_LOG_ADD / VH_LOG_CLEAR / VH_LOG_REMOVE_FROM
};
// log entry kinds: add_edge!, remove_edges!(to), remove_edges!(from, to)
#define VH_LOG_ADD 0
#define VH_LOG_CLEAR 1
#define VH_LOG_REMOVE_FROM 2
struct DEmitAgent {
    uint32_t* cnt;
    const uint64_t* offs;
    int32_t at;
    uint32_t pad;
};

struct DCtx {
    DAgent a[VH_MAXA + 1];
    DEdge e[VH_MAXE];
    DEmitEdge ee[2];
    DEmitAgent ea[1];
    uint32_t* err;
    int8_t role_a[4];
    int8_t role_e[4];
    uint64_t cbase[VH_MAX_CALL];   // offset of call type i in the caller arrays
    uint8_t params[VH_PARAMS_BYTES];
};

#define VH_HAS(h, x) (((h) & (x)) != 0u)

// ------------------------------------------------------------------ accessors
__device__ __forceinline__ void vh_flag(const DCtx* c, uint32_t bit) {
    atomicOr(c->err, bit);
}
// row of target `to` in edge type et, or UINT64_MAX when it has no row
__device__ __forceinline__ uint64_t vh_row(const DEdge& E, uint64_t to) {
    uint32_t t = vh_type_of(to);
    uint64_t nr = vh_nr_of(to);
    if (t == 0 || t > VH_MAXA || nr == 0 || nr > E.nrows[t]) return ~0ull;
    return E.rowbase[t] + nr - 1;
}
__device__ __forceinline__ uint64_t vh_src_id(const DEdge& E, uint64_t k) {
    return E.src64 ? ((const uint64_t*)E.src)[k]
                   : vh_make_id(E.src_type, ((const uint32_t*)E.src)[k]);
}

// ------------------------------------------------------------------ body API
// in-edges [b, e) of `to` for edge type et (false: none / not readable)
__device__ __forceinline__ bool d_row(const DCtx* c, int et, uint64_t to, uint32_t& b, uint32_t& e) {
    b = e = 0;
    if (et < 0 || et >= VH_MAXE) {
        vh_flag(c, VH_DERR_HINT);
        return false;
    }
    const DEdge& E = c->e[et];
    if (!E.readable) {
        vh_flag(c, VH_DERR_READ_EDGE);
        return false;
    }
    if (!E.off) return false;
    uint64_t r = vh_row(E, to);
    if (r == ~0ull) return false;
    b = E.off[r];
    e = E.off[r + 1];
    return e > b;
}
// agentstate / agentstate_flexible (src/AgentMethods.jl:91-123)
__device__ __forceinline__ const uint8_t* d_state(const DCtx* c, uint64_t id) {
    uint32_t t = vh_type_of(id);
    if (t == 0 || t > VH_MAXA) {
        vh_flag(c, VH_DERR_TARGET);
        return nullptr;
    }
    const DAgent& A = c->a[t];
    if (!A.readable) {
        vh_flag(c, VH_DERR_READ_AGENT);
        return nullptr;
    }
    uint64_t nr = vh_nr_of(id);
    if (nr == 0 || nr > A.nslots) {
        vh_flag(c, VH_DERR_TARGET);
        return nullptr;
    }
    if (A.mortal && A.died[nr - 1]) {
        vh_flag(c, VH_DERR_DIED);
        return nullptr;
    }
    return A.rstate + (nr - 1) * A.bytes;
}
__device__ __forceinline__ int64_t ld_i64(const uint8_t* p) {
    int64_t v;
    memcpy(&v, p, 8);
    return v;
}
__device__ __forceinline__ void st_i64(uint8_t* p, int64_t v) { memcpy(p, &v, 8); }

// add_edge! / add_agent! inside a body: pass 0 counts, pass 1 writes at the
// position given by the scan of the counts (caller order, then call order)
struct Emitter {
    const DCtx* c;
    uint64_t caller;
    int pass;
    uint32_t ne, nb;

    __device__ void edge(uint64_t from, uint64_t to, const void* st) {
        const DEmitEdge& E = c->ee[0];
        if (E.et < 0 || !c->e[E.et].writeable) {
            vh_flag(c, VH_DERR_WRITE);
            return;
        }
        if (pass == 1) {
            uint64_t p = E.offs[caller] + ne;
            E.to[p] = to;
            E.from[p] = from;
            for (uint32_t k = 0; k < E.bytes; k++) E.st[p * E.bytes + k] = ((const uint8_t*)st)[k];
            if (E.kind) E.kind[p] = VH_LOG_ADD;
        }
        ne++;
    }
    // remove_edges!(sim, to, ET) (all) or remove_edges!(sim, from, to, ET)
    // src/EdgeMethods.jl:527-599, logged in call order with the add_edge!s
    // (_can_remove_edges :101-121: ET in `write` and in `add_existing`)
    __device__ void remove(uint64_t from, uint64_t to, bool all) {
        const DEmitEdge& E = c->ee[0];
        if (E.et < 0 || !c->e[E.et].writeable) {
            vh_flag(c, VH_DERR_WRITE);
            return;
        }
        if (!c->e[E.et].add_existing) {
            vh_flag(c, VH_DERR_ADDEX);
            return;
        }
        if (!all && VH_HAS(c->e[E.et].hints, VH_HINT_IGNOREFROM)) {
            vh_flag(c, VH_DERR_HINT);
            return;
        }
        if (pass == 1) {
            uint64_t p = E.offs[caller] + ne;
            E.to[p] = to;
            E.from[p] = from;
            for (uint32_t k = 0; k < E.bytes; k++) E.st[p * E.bytes + k] = 0;
            E.kind[p] = all ? VH_LOG_CLEAR : VH_LOG_REMOVE_FROM;
        }
        ne++;
    }
    // _get_next_id src/AgentMethods.jl:37-63: LIFO reuse, then fresh slots
    __device__ uint64_t agent(const uint8_t* st) {
        const DEmitAgent& B = c->ea[0];
        if (B.at <= 0 || !c->a[B.at].writeable) {
            vh_flag(c, VH_DERR_WRITE);
            return 0;
        }
        const DAgent& A = c->a[B.at];
        uint64_t id = vh_make_id(B.at, 0);
        if (pass == 1) {
            uint64_t g = B.offs[caller] + nb;
            uint64_t slot = g < A.nreuse ? A.reuse[A.nreuse - 1 - g] : A.nextid + (g - A.nreuse);
            uint8_t* dst = A.wstate + (slot - 1) * A.bytes;
            for (uint32_t k = 0; k < A.bytes; k++) dst[k] = st[k];
            id = vh_make_id(B.at, slot);
        }
        nb++;
        return id;
    }
    __device__ void finish() {
        if (pass != 0) return;
        if (c->ee[0].cnt && ne) c->ee[0].cnt[caller] = ne;
        if (c->ea[0].cnt && nb) c->ea[0].cnt[caller] = nb;
    }
};

// ------------------------------------------------------------------ agent loop
// One call of a body per slot of call type t (thread-per-slot; died slots are
// skipped): transition_with_read! / transition_with_read_with_edge!
// (src/AgentMethods.jl:193-245) with the write policy of :159-181.
template <class B>
__device__ __forceinline__ void vh_generic_call(const DCtx* __restrict__ c, int ci, int t, int pass,
                                                int with_edge) {
    uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    const DAgent& A = c->a[t];
    if (idx > A.nslots) return;
    if (A.mortal && A.died[idx - 1]) return;
    const uint64_t id = vh_make_id(t, idx);
    const uint32_t bytes = A.bytes;
    const uint8_t* rs = A.rstate + (idx - 1) * bytes;
    if (with_edge >= 0) {
        // transition_with_read_with_edge! src/AgentMethods.jl:209-228: only
        // targets of with_edge are called; the others keep their state
        const DEdge& E = c->e[with_edge];
        uint64_t r = vh_row(E, id);
        bool hasedge = E.off && r != ~0ull && E.off[r + 1] > E.off[r];
        if (!hasedge) {
            if (pass == 1 && A.writeable && !A.indep) {
                uint8_t* ws = A.wstate + (idx - 1) * bytes;
                for (uint32_t k = 0; k < bytes; k++) ws[k] = rs[k];
            }
            return;
        }
    }
    uint8_t st[VH_MAX_STATE], out[VH_MAX_STATE];
    for (uint32_t k = 0; k < bytes; k++) st[k] = out[k] = rs[k];
    Emitter em{c, c->cbase[ci] + idx - 1, pass, 0u, 0u};
    bool keep = B::run(c, st, id, em, out);
    if (pass == 0) {
        em.finish();
        return;
    }
    if (!A.writeable) return;
    if (!keep) {
        if (!A.mortal) vh_flag(c, VH_DERR_IMMORTAL);
        else A.died_now[idx - 1] = 1;
        return;
    }
    uint8_t* ws = A.wstate + (idx - 1) * bytes;
    for (uint32_t k = 0; k < bytes; k++) ws[k] = out[k];
}

// the kernel of body B as an unmangled symbol (for vh_register_body)
#define VH_BODY_KERNEL(name, B)                                                                     \
    extern "C" __global__ __launch_bounds__(256) void name(const DCtx* __restrict__ c, int ci, int t, \
                                                           int pass, int with_edge) {               \
        vh_generic_call<B>(c, ci, t, pass, with_edge);                                              \
    }
