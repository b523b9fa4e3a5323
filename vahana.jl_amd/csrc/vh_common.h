// vh_common.h — shared device/host definitions of the MI355X transition backend.
//
// Device views (DCtx) of the simulation state that every transition kernel
// receives, the AgentID codec of src/Agent.jl:30-118 and the counter-hash RNG
// the model bodies draw from.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/vahana_hip.h"

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
// uniform double in [0,1) from the top 53 bits; float from the top 24
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
// Whether some in-edge k in [b, e) has u01(vh_hash_fin(pre, k - b + 1)) < p
// (pt = vh_u01_below(p)) and an infected source (infected(k)). The answer
// does not depend on the order of the in-edges, so the draws are computed
// four at a time and a source is looked up only for a draw below p.
template <class F>
__device__ __forceinline__ bool vh_any_infecting(uint64_t pre, uint64_t pt, uint32_t b, uint32_t e,
                                                 F infected) {
    uint32_t k = b;
    for (; k + 4 <= e; k += 4) {
        const uint64_t c = k - b + 1;
        const bool d0 = (vh_hash_fin(pre, c) >> 11) < pt, d1 = (vh_hash_fin(pre, c + 1) >> 11) < pt;
        const bool d2 = (vh_hash_fin(pre, c + 2) >> 11) < pt, d3 = (vh_hash_fin(pre, c + 3) >> 11) < pt;
        if (d0 | d1 | d2 | d3) {
            if ((d0 && infected(k)) || (d1 && infected(k + 1)) || (d2 && infected(k + 2)) ||
                (d3 && infected(k + 3)))
                return true;
        }
    }
    for (; k < e; k++)
        if ((vh_hash_fin(pre, k - b + 1) >> 11) < pt && infected(k)) return true;
    return false;
}
__host__ __device__ __forceinline__ float vh_u01f(uint64_t h) {
    return (float)(h >> 40) * 0x1.0p-24f;
}
// stream keys of the model bodies (the `c` word of vh_hash)
#define VH_KEY_RECOVER (1ull << 40)
#define VH_KEY_DEATH   (2ull << 40)
#define VH_KEY_BIRTH   (3ull << 40)
#define VH_KEY_PARTNER (4ull << 40)
#define VH_KEY_INIT    (5ull << 40)
#define VH_KEY_WEIGHT  (6ull << 40)
#define VH_KEY_OPINION (7ull << 40)
#define VH_KEY_SUSC    (8ull << 40)

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
