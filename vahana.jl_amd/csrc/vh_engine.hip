// vh_engine.hip — the C ABI: model/simulation registry, finish_init!, the
// phases of apply! (src/Simulation.jl:720-821) as device passes, CSR builds,
// death/birth bookkeeping and queries.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <unordered_map>

#include "vh_engine.h"

using namespace vh;

namespace vh {

static thread_local char g_err[1024];
static int g_asserts = 1;
vh_options g_opts = {16ll << 20, 0, 1, 0, 65536};

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
int hip_fail(hipError_t e, const char* where) {
    (void)hipGetLastError();   // reported here, not again by a later launch check
    return set_error(e == hipErrorOutOfMemory ? VH_ERR_OOM : VH_ERR_DEVICE, "%s: %s", where,
                     hipGetErrorString(e));
}

int DevBuf::ensure(size_t bytes, hipStream_t st, size_t keep) {
    if (bytes <= cap) return VH_OK;
    size_t ncap = std::max(bytes, cap + cap / 4);
    ncap = (ncap + 255) & ~(size_t)255;
    void* np = nullptr;
    hipError_t e = device_alloc(&np, ncap);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    if (p) {
        if (keep) {
            e = hipMemcpyAsync(np, p, std::min(keep, cap), hipMemcpyDeviceToDevice, st);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync");
        }
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        (void)hipFree(p);
    }
    p = np;
    cap = ncap;
    return VH_OK;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

}  // namespace vh


static inline bool has(uint32_t h, uint32_t x) { return (h & x) != 0; }
static inline bool in_mask(uint32_t m, int i) { return (m >> i) & 1u; }

// ====================================================================== model
extern "C" int vh_model_create(vh_model_t* out) {
    if (!out) return set_error(VH_ERR_ARG, "null out");
    *out = new vh_model();
    return VH_OK;
}
extern "C" int vh_model_destroy(vh_model_t m) {
    delete m;
    return VH_OK;
}
// register_agenttype! src/ModelTypes.jl:81-108
extern "C" int vh_register_agenttype(vh_model_t mh, uint32_t hints, uint32_t bytes,
                                     uint8_t* type_id) {
    Model& m = mh->m;
    if (m.na + 1 > VH_MAXA)
        return set_error(VH_ERR_ASSERT, "Can not add new type, maximal number of types already registered");
    if (hints & ~(VH_HINT_IMMORTAL | VH_HINT_INDEPENDENT))
        return set_error(VH_ERR_ASSERT, "The agent type hint is unknown");
    if (bytes > VH_MAX_STATE) return set_error(VH_ERR_ARG, "agent state larger than %d bytes", VH_MAX_STATE);
    m.na++;
    m.a[m.na].bytes = bytes;
    m.a[m.na].hints = hints;
    *type_id = (uint8_t)m.na;
    return VH_OK;
}
// register_edgetype! src/ModelTypes.jl:158-230
extern "C" int vh_register_edgetype(vh_model_t mh, uint32_t hints, uint32_t bytes, int32_t target,
                                    int64_t size, uint8_t* id) {
    Model& m = mh->m;
    if (m.ne >= VH_MAXE) return set_error(VH_ERR_ASSERT, "too many edge types");
    if (has(hints, VH_HINT_NUMEDGESONLY)) hints |= VH_HINT_STATELESS | VH_HINT_IGNOREFROM;
    if (has(hints, VH_HINT_HASEDGEONLY))
        hints |= VH_HINT_STATELESS | VH_HINT_IGNOREFROM | VH_HINT_SINGLEEDGE;
    if (has(hints, VH_HINT_SINGLETYPE) && target <= 0)
        return set_error(VH_ERR_ASSERT, "the :SingleType hint is set, but the target keyword is missing");
    if (target > 0) hints |= VH_HINT_SINGLETYPE;
    if (target > m.na) return set_error(VH_ERR_UNREGISTERED, "unregistered target type");
    if (has(hints, VH_HINT_SINGLETYPE) && has(hints, VH_HINT_SINGLEEDGE) &&
        !(has(hints, VH_HINT_STATELESS) && has(hints, VH_HINT_IGNOREFROM)))
        return set_error(VH_ERR_ASSERT,
                         "The hints :SingleEdge and :SingleType can be only combined when the type "
                         "has also the hints :Stateless and :IgnoreFrom.");
    if (!has(hints, VH_HINT_STATELESS) && bytes > VH_MAX_STATE)
        return set_error(VH_ERR_ARG, "edge state larger than %d bytes", VH_MAX_STATE);
    Model::ET& e = m.e[m.ne];
    e.bytes = has(hints, VH_HINT_STATELESS) ? 0 : bytes;
    e.hints = hints;
    e.target = target > 0 ? target : 0;
    e.size = size;
    *id = (uint8_t)m.ne++;
    return VH_OK;
}
extern "C" int vh_edgetype_hints(vh_model_t mh, uint8_t et, uint32_t* hints) {
    if (et >= mh->m.ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    *hints = mh->m.e[et].hints;
    return VH_OK;
}
extern "C" const char* vh_last_error(void) { return g_err; }
extern "C" int vh_get_options(vh_options* out) {
    if (!out) return set_error(VH_ERR_ARG, "null argument");
    *out = g_opts;
    return VH_OK;
}
extern "C" int vh_set_options(const vh_options* in) {
    if (!in) return set_error(VH_ERR_ARG, "null argument");
    if (in->opinion_part_bytes < 0) return set_error(VH_ERR_ARG, "opinion_part_bytes must be >= 0");
    if (in->sir_expose != 0 && in->sir_expose != 1) return set_error(VH_ERR_ARG, "sir_expose must be 0 or 1");
    if (in->sir_expose_min < 0) return set_error(VH_ERR_ARG, "sir_expose_min must be >= 0");
    g_opts = *in;
    g_opts.pad = 0;
    return VH_OK;
}
extern "C" int vh_enable_asserts(int enable) {
    g_asserts = enable ? 1 : 0;
    return VH_OK;
}
extern "C" int vh_debug_fail_allocs(uint32_t n) {
    fail_next_allocs(n);
    return VH_OK;
}
extern "C" int vh_debug_set_device_error(vh_sim_t h, uint32_t flags) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    Sim& s = h->s;
    uint32_t e = 0;
    HIPCHECK(hipStreamSynchronize(s.stream));
    HIPCHECK(hipMemcpy(&e, s.d_err.p, 4, hipMemcpyDeviceToHost));
    e |= flags;
    HIPCHECK(hipMemcpy(s.d_err.p, &e, 4, hipMemcpyHostToDevice));
    return VH_OK;
}

// ================================================================== simulation
extern "C" int vh_sim_create(vh_model_t mh, vh_sim_t* out) {
    if (!mh || !out) return set_error(VH_ERR_ARG, "null argument");
    vh_sim* h = new vh_sim();
    Sim& s = h->s;
    s.model = mh->m;
    s.m = &s.model;
    hipError_t e = hipGetDevice(&s.device);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipGetDevice");
    }
    e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipStreamCreate");
    }
    s.own_stream = true;
    e = hipEventCreateWithFlags(&s.last, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s.stream);
        delete h;
        return hip_fail(e, "hipEventCreate");
    }
    for (int t = 1; t <= s.m->na; t++) {
        s.a[t].bytes = s.m->a[t].bytes;
        s.a[t].mortal = !has(s.m->a[t].hints, VH_HINT_IMMORTAL);
        s.a[t].indep = has(s.m->a[t].hints, VH_HINT_INDEPENDENT);
    }
    for (int et = 0; et < s.m->ne; et++) {
        s.e[et].bytes = s.m->e[et].bytes;
        s.e[et].hints = s.m->e[et].hints;
        s.e[et].target = s.m->e[et].target;
        s.e[et].src64 = s.m->na != 1;
        s.e[et].src_type = 1;
    }
    e = hipHostMalloc((void**)&s.h_stats, 64 * sizeof(uint64_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipEventDestroy(s.last);
        (void)hipStreamDestroy(s.stream);
        delete h;
        return hip_fail(e, "hipHostMalloc");
    }
    int r = s.d_stats.ensure(64 * sizeof(uint64_t), s.stream);
    if (!r) r = s.d_err.ensure(sizeof(uint32_t), s.stream);
    if (!r) r = s.d_ctx.ensure(sizeof(DCtx), s.stream);
    if (r) {
        delete h;
        return r;
    }
    fill_u32(s.d_err.as<uint32_t>(), 0, 1, s.stream);
    *out = h;
    return VH_OK;
}

static void sim_release(Sim& s) {
    // on a caller's (possibly shared) stream only this simulation's work
    // must be complete: its last apply!; other work is synchronous
    if (s.stream && s.own_stream) (void)hipStreamSynchronize(s.stream);
    else if (s.last) (void)hipEventSynchronize(s.last);
    for (int t = 0; t <= VH_MAXA; t++) {
        AgentStore& A = s.a[t];
        A.state[0].release(); A.state[1].release();
        A.died.release(); A.died_now.release(); A.reuse.release();
    }
    for (int et = 0; et < VH_MAXE; et++) {
        s.e[et].each_buf([](DevBuf& b) { b.release(); });
    }
    s.d_ctx.release(); s.d_err.release(); s.d_stats.release();
    s.scratch.release();
    s.relabel_map.release();
    if (s.h_stats) (void)hipHostFree(s.h_stats);
    s.h_stats = nullptr;
    if (s.own_stream && s.stream) (void)hipStreamDestroy(s.stream);
    s.stream = nullptr;
    if (s.last) (void)hipEventDestroy(s.last);
    s.last = nullptr;
}

// finish_simulation! src/Simulation.jl:551-585
extern "C" int vh_sim_destroy(vh_sim_t h) {
    if (!h) return VH_OK;
    sim_release(h->s);
    delete h;
    return VH_OK;
}

extern "C" int vh_set_stream(vh_sim_t h, void* stream) {
    Sim& s = h->s;
    HIPCHECK(hipStreamSynchronize(s.stream));
    if (s.own_stream) (void)hipStreamDestroy(s.stream);
    if (stream) {
        s.stream = (hipStream_t)stream;
        s.own_stream = false;
    } else {
        HIPCHECK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        s.own_stream = true;
    }
    return VH_OK;
}
extern "C" int vh_stream_create(void** out) {
    if (!out) return set_error(VH_ERR_ARG, "null argument");
    hipStream_t st = nullptr;
    HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *out = st;
    return VH_OK;
}
extern "C" int vh_stream_destroy(void* stream) {
    if (!stream) return VH_OK;
    HIPCHECK(hipStreamSynchronize((hipStream_t)stream));
    HIPCHECK(hipStreamDestroy((hipStream_t)stream));
    return VH_OK;
}
extern "C" int vh_set_io_stream(vh_sim_t h, void* stream) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    h->s.io = (hipStream_t)stream;
    return VH_OK;
}

extern "C" int vh_sync(vh_sim_t h) {
    HIPCHECK(hipStreamSynchronize(h->s.stream));
    return VH_OK;
}

// read_snapshot! src/HDF5.jl:1594-1646 with _read_agents_restore! :801-850:
// after finish_init!, the died flags of the n slots of type t; the reuse
// stack becomes the died slots in ascending order, as the reference rebuilds it
extern "C" int vh_restore_died(vh_sim_t h, uint8_t t, const uint8_t* died, uint64_t n) {
    if (!h || (n && !died)) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "restore after finish_init! only");
    AgentStore& A = s.a[t];
    if (n != A.nslots) return set_error(VH_ERR_ARG, "restore: %llu slots, the type has %llu",
                                        (unsigned long long)n, (unsigned long long)A.nslots);
    std::vector<uint32_t> free_slots;
    for (uint64_t i = 0; i < n; i++)
        if (died[i]) free_slots.push_back((uint32_t)(i + 1));
    if (!A.mortal) {
        if (!free_slots.empty()) return set_error(VH_ERR_ASSERT, "restore: a died agent of an :Immortal type");
        return VH_OK;
    }
    hipStream_t st = s.stream;
    HIPCHECK(hipStreamSynchronize(st));
    if (n) HIPCHECK(hipMemcpy(A.died.p, died, n, hipMemcpyHostToDevice));
    RC(A.reuse.ensure(free_slots.size() * 4 + 8, st));
    if (!free_slots.empty())
        HIPCHECK(hipMemcpy(A.reuse.p, free_slots.data(), free_slots.size() * 4, hipMemcpyHostToDevice));
    A.nreuse = free_slots.size();
    return VH_OK;
}
extern "C" int vh_set_num_transitions(vh_sim_t h, uint64_t n) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    h->s.num_transitions = n;
    return VH_OK;
}

// last_change of an agent (kind 0) or edge (kind 1) type: the transition
// that last wrote it (src/AgentMethods.jl:433, src/EdgeMethods.jl:682), or
// whose deaths removed edges of it (src/AgentMethods.jl:347-353); 0 after
// initialisation
extern "C" int vh_last_change(vh_sim_t h, int kind, uint8_t type_id, uint64_t* out) {
    if (!h || !out) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (kind == 0) {
        if (type_id < 1 || type_id > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
        *out = s.a[type_id].last_change;
    } else {
        if (type_id >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
        *out = s.e[type_id].last_change;
    }
    return VH_OK;
}
// read_snapshot! restores last_change from the snapshot (src/HDF5.jl:1267,1434-1436)
extern "C" int vh_set_last_change(vh_sim_t h, int kind, uint8_t type_id, uint64_t value) {
    if (!h) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (kind == 0) {
        if (type_id < 1 || type_id > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
        s.a[type_id].last_change = value;
    } else {
        if (type_id >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
        s.e[type_id].last_change = value;
    }
    return VH_OK;
}

extern "C" int vh_num_transitions(vh_sim_t h, uint64_t* out) {
    *out = h->s.num_transitions;
    return VH_OK;
}

// add_agent! during the initialisation phase src/AgentMethods.jl:65-89
extern "C" int vh_add_agents(vh_sim_t h, uint8_t t, const void* states, uint64_t n,
                             uint64_t* ids) {
    Sim& s = h->s;
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    if (s.initialized)
        return set_error(VH_ERR_ASSERT,
                         "You can call add_agent! only in the initialization phase (until "
                         "`finish_init!` is called) or within a transition function called by `apply`.");
    AgentStore& A = s.a[t];
    size_t old = A.h_state.size();
    A.h_state.resize(old + (size_t)n * A.bytes);
    if (n && A.bytes) memcpy(A.h_state.data() + old, states, (size_t)n * A.bytes);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t nr = A.nextid++;
        if (ids) ids[i] = vh_make_id(t, nr);
    }
    if (A.nextid - 1 > VH_NR_MASK) return set_error(VH_ERR_ASSERT, "agent nr space exhausted");
    return VH_OK;
}

// add_edge! during the initialisation phase src/EdgeMethods.jl:388-523
extern "C" int vh_add_edges(vh_sim_t h, uint8_t et, const uint64_t* from, const uint64_t* to,
                            const void* states, uint64_t n) {
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    if (s.initialized)
        return set_error(VH_ERR_ASSERT,
                         "You can call add_edge! only in the initialization phase (until "
                         "`finish_init!` is called) or within a transition function called by `apply`.");
    EdgeStore& E = s.e[et];
    if (E.gen || E.stencil)
        return set_error(VH_ERR_ARG, "the in-edges of this edge type are device-generated");
    const bool ignorefrom = has(E.hints, VH_HINT_IGNOREFROM);
    const bool stateless = has(E.hints, VH_HINT_STATELESS);
    const bool singletype = has(E.hints, VH_HINT_SINGLETYPE);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t tt = vh_type_of(to[i]);
        if (tt < 1 || (int)tt > s.m->na || vh_nr_of(to[i]) == 0 ||
            vh_nr_of(to[i]) >= s.a[tt].nextid || (singletype && (int)tt != E.target))
            return set_error(VH_ERR_ASSERT, "invalid target id %llu for the edge type",
                             (unsigned long long)to[i]);
        if (!ignorefrom) {
            uint32_t ft = vh_type_of(from[i]);
            if (ft < 1 || (int)ft > s.m->na || vh_nr_of(from[i]) == 0 ||
                vh_nr_of(from[i]) >= s.a[ft].nextid)
                return set_error(VH_ERR_ASSERT, "invalid source id %llu", (unsigned long long)from[i]);
        }
    }
    if (E.bytes && !states && n)
        return set_error(VH_ERR_ARG, "add_edge!: the edge type has a state; `states` must not be NULL");
    size_t old = E.h_to.size();
    // _can_add src/EdgeMethods.jl:257-294: a second, different edge to the
    // same target of a :SingleEdge type (not :SingleType, not S+I) asserts.
    // E.h_first maps a target to the index of its first edge in add_edge!
    // order; an index >= old is an entry of this call (from[k - old]).
    if (has(E.hints, VH_HINT_SINGLEEDGE) && !singletype && !(stateless && ignorefrom)) {
        std::vector<uint64_t> added;   // targets first seen in this call (undone on error)
        for (uint64_t i = 0; i < n; i++) {
            auto it = E.h_first.find(to[i]);
            if (it != E.h_first.end()) {
                const size_t k = it->second;
                const bool batch = k >= old;
                const uint64_t kf = batch ? from[k - old] : E.h_from[k];
                const uint8_t* ks = batch ? (const uint8_t*)states + (k - old) * E.bytes
                                          : E.h_st.data() + k * E.bytes;
                bool same = ignorefrom || kf == from[i];
                if (E.bytes && memcmp(ks, (const uint8_t*)states + i * E.bytes, E.bytes) != 0) same = false;
                if (!same) {
                    for (uint64_t tgt : added) E.h_first.erase(tgt);
                    return set_error(VH_ERR_ASSERT,
                                     "An edge has already been added to the agent with the id %llu "
                                     "(and the edgetype hints containing the :SingleEdge hint).",
                                     (unsigned long long)to[i]);
                }
            } else {
                E.h_first.emplace(to[i], old + i);
                added.push_back(to[i]);
            }
        }
    }
    E.h_to.insert(E.h_to.end(), to, to + n);
    E.h_from.resize(old + n);
    if (!ignorefrom) memcpy(E.h_from.data() + old, from, n * 8);
    if (E.bytes) {
        size_t ob = E.h_st.size();
        E.h_st.resize(ob + (size_t)n * E.bytes);
        if (states) memcpy(E.h_st.data() + ob, states, (size_t)n * E.bytes);
    }
    return VH_OK;
}

// ================================================================ CSR kernels
namespace {

struct RowLayout {
    uint64_t base[VH_MAXA + 1];
    uint64_t nrows[VH_MAXA + 1];
};

// keys of log entries in the new row layout
__global__ void k_log_keys(const uint64_t* __restrict__ to, uint64_t n, RowLayout L,
                           uint32_t* __restrict__ keys, uint32_t* err) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t id = to[i];
    uint32_t t = vh_type_of(id);
    uint64_t nr = vh_nr_of(id);
    if (t == 0 || t > VH_MAXA || nr == 0 || nr > L.nrows[t]) {
        atomicOr(err, VH_DERR_TARGET);
        keys[i] = 0;
        return;
    }
    keys[i] = (uint32_t)(L.base[t] + nr - 1);
}

// removal entries of the log (remove_edges!): not edges (key past every
// row), their target row goes to the event list
__global__ void k_log_removals(const uint8_t* __restrict__ kind, uint64_t n, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ ekeys) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool rm = kind[i] != VH_LOG_ADD;
    ekeys[i] = rm ? keys[i] : 0xFFFFFFFFu;
    if (rm) keys[i] = 0xFFFFFFFFu;
}

// sequence of every entry: 0 for the kept container, log position + 1
__global__ void k_entry_seq(uint64_t nbase, uint64_t n, uint32_t* __restrict__ seq) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seq[i] = i < nbase ? 0u : (uint32_t)(i - nbase + 1);
}

// one thread per row: an entry stays unless a removal of its row comes later
// in call order (remove_edges!(to) any source, remove_edges!(from, to) its
// source); with :SingleEdge only the row's last entry is a candidate (the
// container holds the last add_edge!, which a later removal can clear).
// Events: ekeys sorted by row (stable, so log order within a row), eidx the
// log positions.
__global__ void k_remove_filter(const uint32_t* __restrict__ off, const uint64_t* __restrict__ from,
                                const uint32_t* __restrict__ seq, uint64_t rows, uint32_t single,
                                const uint32_t* __restrict__ ekeys, const uint32_t* __restrict__ eidx,
                                uint64_t nev, const uint8_t* __restrict__ kind,
                                const uint64_t* __restrict__ lfrom, uint8_t* __restrict__ keep) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t b = off[r], e = off[r + 1];
    if (b == e) return;
    uint64_t lo = 0, hi = nev;   // first event of row r
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)ekeys[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    for (uint32_t k = b; k < e; k++) {
        bool kp = !single || k == e - 1;
        for (uint64_t q = lo; kp && q < nev && ekeys[q] == (uint32_t)r; q++) {
            const uint32_t pl = eidx[q];
            if (pl + 1 <= seq[k]) continue;   // before the entry was added
            if (kind[pl] == VH_LOG_CLEAR || (from && lfrom[pl] == from[k])) kp = false;
        }
        keep[k] = kp ? 1 : 0;
    }
}

// explode an existing CSR into (key, source, state) entries in the new layout
__global__ void k_base_expand(const uint32_t* __restrict__ off, const void* __restrict__ src,
                              uint32_t src64, uint32_t src_type, const uint8_t* __restrict__ st,
                              uint32_t bytes, RowLayout oldL, RowLayout newL, uint64_t rows,
                              uint32_t* __restrict__ keys, uint64_t* __restrict__ from,
                              uint8_t* __restrict__ stout, uint32_t* err) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    uint32_t t = 0;
    for (uint32_t q = 1; q <= VH_MAXA; q++)
        if (oldL.nrows[q] && r >= oldL.base[q] && r < oldL.base[q] + oldL.nrows[q]) t = q;
    uint64_t nr = r - oldL.base[t] + 1;
    bool ok = t != 0 && nr <= newL.nrows[t];
    if (!ok && off[r + 1] > off[r]) atomicOr(err, VH_DERR_TARGET);
    uint32_t key = ok ? (uint32_t)(newL.base[t] + nr - 1) : 0u;
    for (uint32_t k = off[r]; k < off[r + 1]; k++) {
        keys[k] = key;
        if (from) from[k] = src64 ? ((const uint64_t*)src)[k] : vh_make_id(src_type, ((const uint32_t*)src)[k]);
        for (uint32_t b = 0; b < bytes; b++) stout[(uint64_t)k * bytes + b] = st[(uint64_t)k * bytes + b];
    }
}

// off[r] = first index whose key >= r, for r in [0, rows]: one thread per
// row, binary search over the sorted keys (rows without keys cost no more
// than rows with keys)
__global__ void k_offsets_from_sorted(const uint32_t* __restrict__ keys, uint64_t n, uint64_t rows,
                                      uint32_t* __restrict__ off) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > rows) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    off[r] = (uint32_t)lo;
}

// row-wise keep flags: SingleEdge keeps the last entry; deaths drop rows of
// targets and entries of tracked sources that died in this step
struct DiedNow {
    const uint8_t* d[VH_MAXA + 1];
};
__global__ void k_filter_rows(const uint32_t* __restrict__ off, const void* __restrict__ src,
                              uint32_t src64, uint32_t src_type, RowLayout L, uint64_t rows,
                              uint32_t keep_last, uint32_t track_src, DiedNow dn,
                              uint8_t* __restrict__ keep) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    uint32_t b = off[r], e = off[r + 1];
    if (b == e) return;
    uint32_t t = 0;
    for (uint32_t q = 1; q <= VH_MAXA; q++)
        if (L.nrows[q] && r >= L.base[q] && r < L.base[q] + L.nrows[q]) t = q;
    uint64_t nr = r - L.base[t] + 1;
    bool tdead = dn.d[t] && dn.d[t][nr - 1];
    for (uint32_t k = b; k < e; k++) {
        bool kp = !tdead && (!keep_last || k == e - 1);
        if (kp && track_src) {
            uint64_t id = src64 ? ((const uint64_t*)src)[k]
                                : vh_make_id(src_type, ((const uint32_t*)src)[k]);
            uint32_t ft = vh_type_of(id);
            if (ft <= VH_MAXA && dn.d[ft] && dn.d[ft][vh_nr_of(id) - 1]) kp = false;
        }
        keep[k] = kp ? 1 : 0;
    }
}

// compaction of CSR entries by keep flags; pos = exclusive scan of keep
__global__ void k_compact_entries(const uint8_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                                  uint64_t n, const void* __restrict__ src, void* __restrict__ srco,
                                  uint32_t srcw, const uint8_t* __restrict__ st,
                                  uint8_t* __restrict__ sto, uint32_t bytes) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    uint64_t p = pos[i];
    if (srcw == 8) ((uint64_t*)srco)[p] = ((const uint64_t*)src)[i];
    else if (srcw == 4) ((uint32_t*)srco)[p] = ((const uint32_t*)src)[i];
    for (uint32_t b = 0; b < bytes; b++) sto[p * bytes + b] = st[i * bytes + b];
}
__global__ void k_remap_offsets(const uint32_t* __restrict__ off, const uint32_t* __restrict__ pos,
                                uint64_t rows, uint32_t* __restrict__ offo) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r <= rows) offo[r] = pos[off[r]];
}
__global__ void k_u64_to_nr(const uint64_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)vh_nr_of(in[i]);
}
__global__ void k_u8_to_u32(const uint8_t* __restrict__ in, uint32_t* __restrict__ out, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

}  // namespace

static RowLayout layout_of(const EdgeStore& E) {
    RowLayout L;
    for (int t = 0; t <= VH_MAXA; t++) {
        L.base[t] = E.rowbase[t];
        L.nrows[t] = E.nrows[t];
    }
    return L;
}

// row layout of an edge type for the current slot counts
static void compute_layout(Sim& s, const EdgeStore& E, RowLayout& L, uint64_t& rows) {
    rows = 0;
    for (int t = 0; t <= VH_MAXA; t++) { L.base[t] = 0; L.nrows[t] = 0; }
    for (int t = 1; t <= s.m->na; t++) {
        bool tgt = has(E.hints, VH_HINT_SINGLETYPE) ? t == E.target : true;
        uint64_t n = tgt ? s.a[t].nextid - 1 : 0;
        L.base[t] = rows;
        L.nrows[t] = n;
        rows += n;
    }
}

int vh::readback(Sim& s, int n) {
    HIPCHECK(hipMemcpyAsync(s.h_stats, s.d_stats.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    return VH_OK;
}

// Build the CSR of E from (optionally) its current entries followed by `nlog`
// log entries (to, from, states) in add_edge! order: a stable sort by target
// row keeps, per target, the existing edges first and the new ones in call
// order (src/EdgeMethods.jl:477-523); :SingleEdge keeps the last one. With
// `d_kind`, log entries can be remove_edges! calls, applied to what precedes
// them in call order (src/EdgeMethods.jl:527-599).
static int build_csr(Sim& s, EdgeStore& E, bool keep_base, const uint64_t* d_to,
                     const uint64_t* d_from, const uint8_t* d_st, uint64_t nlog,
                     const uint8_t* d_kind = nullptr) {
    hipStream_t st = s.stream;
    Scratch& sc = s.scratch;
    RowLayout oldL = layout_of(E), newL;
    uint64_t rows;
    compute_layout(s, E, newL, rows);
    if (rows >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "more than 2^32 target rows");
    const bool ignorefrom = has(E.hints, VH_HINT_IGNOREFROM);
    const uint32_t bytes = E.bytes;
    uint64_t nbase = (keep_base && E.rows) ? E.nedges : 0;
    uint64_t n = nbase + nlog;
    if (n >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "more than 2^32 edges of one type");

    uint32_t* keys = (uint32_t*)sc.take(n * 4 + 4);
    uint32_t* vals = (uint32_t*)sc.take(n * 4 + 4);
    uint32_t* kalt = (uint32_t*)sc.take(n * 4 + 4);
    uint32_t* valt = (uint32_t*)sc.take(n * 4 + 4);
    uint64_t* from = ignorefrom ? nullptr : (uint64_t*)sc.take(n * 8 + 8);
    uint8_t* sts = bytes ? (uint8_t*)sc.take(n * bytes + 8) : nullptr;
    uint32_t* flag = (uint32_t*)sc.take(4);
    if (!keys || !vals || !kalt || !valt || (!ignorefrom && !from) || (bytes && !sts) || !flag)
        return set_error(VH_ERR_OOM, "scratch allocation failed");
    uint32_t* err = s.d_err.as<uint32_t>();
    if (nbase)
        hipLaunchKernelGGL(k_base_expand, dim3(grid_for(E.rows, 256)), dim3(256), 0, st,
                           E.off.as<uint32_t>(), E.src.p, (uint32_t)E.src64, E.src_type,
                           E.st.as<uint8_t>(), bytes, oldL, newL, E.rows, keys, from, sts, err);
    uint32_t* ekeys = nullptr;   // removal events (remove_edges!)
    uint32_t* eidx = nullptr;
    uint32_t* seq = nullptr;
    if (nlog) {
        hipLaunchKernelGGL(k_log_keys, dim3(grid_for(nlog, 256)), dim3(256), 0, st, d_to, nlog, newL,
                           keys + nbase, err);
        if (d_kind) {
            ekeys = (uint32_t*)sc.take(nlog * 4 + 4);
            eidx = (uint32_t*)sc.take(nlog * 4 + 4);
            uint32_t* ka = (uint32_t*)sc.take(nlog * 4 + 4);
            uint32_t* va = (uint32_t*)sc.take(nlog * 4 + 4);
            seq = (uint32_t*)sc.take(n * 4 + 4);
            if (!ekeys || !eidx || !ka || !va || !seq) return set_error(VH_ERR_OOM, "scratch");
            hipLaunchKernelGGL(k_log_removals, dim3(grid_for(nlog, 256)), dim3(256), 0, st, d_kind, nlog,
                               keys + nbase, ekeys);
            if (radix_sort_pairs(&ekeys, &eidx, ka, va, nlog, 32, true, sc, st))
                return set_error(VH_ERR_DEVICE, "radix sort failed");
            hipLaunchKernelGGL(k_entry_seq, dim3(grid_for(n, 256)), dim3(256), 0, st, nbase, n, seq);
        }
        if (from) HIPCHECK(hipMemcpyAsync(from + nbase, d_from, nlog * 8, hipMemcpyDeviceToDevice, st));
        if (bytes)
            HIPCHECK(hipMemcpyAsync(sts + nbase * bytes, d_st, nlog * bytes, hipMemcpyDeviceToDevice, st));
    }
    // already in row order?  (rebuilds whose appends target their callers)
    bool sorted = nlog == 0;
    if (!sorted) {
        check_sorted(keys, n, flag, st);
        HIPCHECK(hipMemcpyAsync(s.d_stats.p, flag, 4, hipMemcpyDeviceToDevice, st));
        RC(readback(s, 1));
        sorted = (uint32_t)s.h_stats[0] == 1u;
    }
    uint64_t* from_s = from;
    uint8_t* st_s = sts;
    uint32_t* skeys = keys;
    uint32_t* seq_s = seq;
    if (!sorted) {
        uint32_t bits = 1;
        while (bits < 32 && (rows >> bits) != 0) bits++;
        if (ekeys) bits = 32;   // removal entries sort last
        uint32_t* sv = vals;
        if (radix_sort_pairs(&skeys, &sv, kalt, valt, n, bits, true, sc, st))
            return set_error(VH_ERR_DEVICE, "radix sort failed");
        if (from) {
            from_s = (uint64_t*)sc.take(n * 8 + 8);
            if (!from_s) return set_error(VH_ERR_OOM, "scratch");
            gather_records(from_s, from, sv, n, 8, st);
        }
        if (bytes) {
            st_s = (uint8_t*)sc.take(n * bytes + 8);
            if (!st_s) return set_error(VH_ERR_OOM, "scratch");
            gather_records(st_s, sts, sv, n, bytes, st);
        }
        if (seq) {
            seq_s = (uint32_t*)sc.take(n * 4 + 4);
            if (!seq_s) return set_error(VH_ERR_OOM, "scratch");
            gather_records(seq_s, seq, sv, n, 4, st);
        }
    }
    // offsets
    uint32_t* off = (uint32_t*)sc.take((rows + 1) * 4);
    if (!off) return set_error(VH_ERR_OOM, "scratch");
    hipLaunchKernelGGL(k_offsets_from_sorted, dim3(grid_for(rows + 1, 256)), dim3(256), 0, st, skeys, n,
                       rows, off);
    uint64_t nfinal = n;
    if (ekeys && n) {
        // remove_edges!: drop entries removed after they were added (and,
        // with :SingleEdge, all but the last)
        uint8_t* keep = (uint8_t*)sc.take(n + 8);
        uint32_t* pos = (uint32_t*)sc.take((n + 1) * 4);
        uint32_t* k32 = (uint32_t*)sc.take((n + 1) * 4);
        uint32_t* off2 = (uint32_t*)sc.take((rows + 1) * 4);
        uint64_t* from2 = from_s ? (uint64_t*)sc.take(n * 8 + 8) : nullptr;
        uint8_t* st2 = bytes ? (uint8_t*)sc.take(n * bytes + 8) : nullptr;
        if (!keep || !pos || !k32 || !off2 || (from_s && !from2) || (bytes && !st2))
            return set_error(VH_ERR_OOM, "scratch");
        fill_u8(keep, 0, n, st);   // (removal entries, past the last row, are dropped)
        hipLaunchKernelGGL(k_remove_filter, dim3(grid_for(rows, 256)), dim3(256), 0, st, off, from_s, seq_s,
                           rows, has(E.hints, VH_HINT_SINGLEEDGE) ? 1u : 0u, ekeys, eidx, nlog, d_kind, d_from,
                           keep);
        hipLaunchKernelGGL(k_u8_to_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, k32, n);
        if (scan_excl_u32(k32, pos, n, sc, st)) return set_error(VH_ERR_DEVICE, "scan");
        hipLaunchKernelGGL(k_compact_entries, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, pos, n,
                           (const void*)from_s, (void*)from2, from_s ? 8u : 0u, st_s, st2, bytes);
        hipLaunchKernelGGL(k_remap_offsets, dim3(grid_for(rows + 1, 256)), dim3(256), 0, st, off, pos,
                           rows, off2);
        off = off2;
        from_s = from2;
        st_s = st2;
        HIPCHECK(hipMemcpyAsync(s.d_stats.p, off + rows, 4, hipMemcpyDeviceToDevice, st));
        RC(readback(s, 1));
        nfinal = (uint32_t)s.h_stats[0];
    } else if (has(E.hints, VH_HINT_SINGLEEDGE) && n) {
        // :SingleEdge keeps the last added edge per target
        uint8_t* keep = (uint8_t*)sc.take(n + 8);
        uint32_t* pos = (uint32_t*)sc.take((n + 1) * 4);
        uint32_t* k32 = (uint32_t*)sc.take((n + 1) * 4);
        uint32_t* off2 = (uint32_t*)sc.take((rows + 1) * 4);
        uint64_t* from2 = from_s ? (uint64_t*)sc.take(n * 8 + 8) : nullptr;
        uint8_t* st2 = bytes ? (uint8_t*)sc.take(n * bytes + 8) : nullptr;
        if (!keep || !pos || !k32 || !off2) return set_error(VH_ERR_OOM, "scratch");
        DiedNow dn;
        memset(&dn, 0, sizeof dn);
        hipLaunchKernelGGL(k_filter_rows, dim3(grid_for(rows, 256)), dim3(256), 0, st, off,
                           (const void*)from_s, 1u, 0u, newL, rows, 1u, 0u, dn, keep);
        hipLaunchKernelGGL(k_u8_to_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, k32, n);
        if (scan_excl_u32(k32, pos, n, sc, st)) return set_error(VH_ERR_DEVICE, "scan");
        hipLaunchKernelGGL(k_compact_entries, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, pos, n,
                           (const void*)from_s, (void*)from2, from_s ? 8u : 0u, st_s, st2, bytes);
        hipLaunchKernelGGL(k_remap_offsets, dim3(grid_for(rows + 1, 256)), dim3(256), 0, st, off, pos,
                           rows, off2);
        off = off2;
        from_s = from2;
        st_s = st2;
        HIPCHECK(hipMemcpyAsync(s.d_stats.p, off + rows, 4, hipMemcpyDeviceToDevice, st));
        RC(readback(s, 1));
        nfinal = (uint32_t)s.h_stats[0];
    }
    // publish into the store
    RC(E.off.ensure((rows + 1) * 4, st));
    HIPCHECK(hipMemcpyAsync(E.off.p, off, (rows + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (!ignorefrom) {
        if (E.src64) {
            RC(E.src.ensure(nfinal * 8 + 8, st));
            HIPCHECK(hipMemcpyAsync(E.src.p, from_s, nfinal * 8, hipMemcpyDeviceToDevice, st));
        } else {
            RC(E.src.ensure(nfinal * 4 + 8, st));
            if (nfinal)
                hipLaunchKernelGGL(k_u64_to_nr, dim3(grid_for(nfinal, 256)), dim3(256), 0, st, from_s,
                                   E.src.as<uint32_t>(), nfinal);
        }
    }
    if (bytes) {
        RC(E.st.ensure(nfinal * bytes + 8, st));
        HIPCHECK(hipMemcpyAsync(E.st.p, st_s, nfinal * bytes, hipMemcpyDeviceToDevice, st));
    }
    for (int t = 0; t <= VH_MAXA; t++) {
        E.rowbase[t] = newL.base[t];
        E.nrows[t] = newL.nrows[t];
    }
    E.rows = rows;
    E.nedges = nfinal;
    E.version++;
    return VH_OK;
}

// drop rows of targets and entries of tracked sources that died in this step
static int filter_deaths(Sim& s, EdgeStore& E, const DiedNow& dn) {
    if (!E.rows || !E.nedges) return VH_OK;
    hipStream_t st = s.stream;
    Scratch& sc = s.scratch;
    uint64_t n = E.nedges, rows = E.rows;
    bool track = !has(E.hints, VH_HINT_IGNOREFROM);
    RowLayout L = layout_of(E);
    uint8_t* keep = (uint8_t*)sc.take(n + 8);
    uint32_t* k32 = (uint32_t*)sc.take((n + 1) * 4);
    uint32_t* pos = (uint32_t*)sc.take((n + 1) * 4);
    uint32_t* off2 = (uint32_t*)sc.take((rows + 1) * 4);
    uint32_t srcw = track ? (E.src64 ? 8u : 4u) : 0u;
    void* src2 = track ? sc.take(n * srcw + 8) : nullptr;
    uint8_t* st2 = E.bytes ? (uint8_t*)sc.take(n * E.bytes + 8) : nullptr;
    if (!keep || !k32 || !pos || !off2 || (track && !src2) || (E.bytes && !st2))
        return set_error(VH_ERR_OOM, "scratch");
    fill_u8(keep, 1, n, st);
    hipLaunchKernelGGL(k_filter_rows, dim3(grid_for(rows, 256)), dim3(256), 0, st,
                       E.off.as<uint32_t>(), E.src.p, (uint32_t)E.src64, E.src_type, L, rows, 0u,
                       track ? 1u : 0u, dn, keep);
    hipLaunchKernelGGL(k_u8_to_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, k32, n);
    if (scan_excl_u32(k32, pos, n, sc, st)) return set_error(VH_ERR_DEVICE, "scan");
    hipLaunchKernelGGL(k_compact_entries, dim3(grid_for(n, 256)), dim3(256), 0, st, keep, pos, n,
                       E.src.p, src2, srcw, E.st.as<uint8_t>(), st2, E.bytes);
    hipLaunchKernelGGL(k_remap_offsets, dim3(grid_for(rows + 1, 256)), dim3(256), 0, st,
                       E.off.as<uint32_t>(), pos, rows, off2);
    HIPCHECK(hipMemcpyAsync(s.d_stats.p, off2 + rows, 4, hipMemcpyDeviceToDevice, st));
    RC(readback(s, 1));
    uint64_t nn = (uint32_t)s.h_stats[0];
    HIPCHECK(hipMemcpyAsync(E.off.p, off2, (rows + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (track && nn) HIPCHECK(hipMemcpyAsync(E.src.p, src2, nn * srcw, hipMemcpyDeviceToDevice, st));
    if (E.bytes && nn) HIPCHECK(hipMemcpyAsync(E.st.p, st2, nn * E.bytes, hipMemcpyDeviceToDevice, st));
    E.nedges = nn;
    E.version++;
    return VH_OK;
}

// ================================================================ agents
int vh::ensure_slots(Sim& s, int t, uint64_t slots) {
    AgentStore& A = s.a[t];
    if (slots <= A.cap && A.state[0].p) return VH_OK;
    uint64_t nc = std::max<uint64_t>(slots, A.cap + A.cap / 4);
    if (nc < 64) nc = 64;
    size_t keep = A.nslots * A.bytes;
    size_t keepw = (std::max(A.nslots, A.nextid - 1)) * A.bytes;
    RC(A.state[0].ensure(nc * A.bytes + 8, s.stream, A.rd == 0 ? keep : keepw));
    if (!A.indep) RC(A.state[1].ensure(nc * A.bytes + 8, s.stream, A.rd == 1 ? keep : keepw));
    if (A.mortal) {
        RC(A.died.ensure(nc + 8, s.stream, A.nextid - 1));
        RC(A.died_now.ensure(nc + 8, s.stream, A.nextid - 1));
        RC(A.reuse.ensure(nc * 4 + 8, s.stream, A.nreuse * 4));
    }
    A.cap = nc;
    return VH_OK;
}

// Bulk init-phase edges: the symmetrised Barabasi-Albert graph that
// vh_graph_ba + add_edge! of both directions of every pair (a -> b, then
// b -> a, pair by pair) would give, Float32 weights by that position,
// generated on the device at finish_init! (add_graph! of a generated graph,
// src/GraphsSupport.jl:34-57).
extern "C" int vh_add_edges_ba(vh_sim_t h, uint8_t et, uint32_t m, uint64_t graph_seed, uint64_t weight_seed) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    if (s.initialized)
        return set_error(VH_ERR_ASSERT, "You can call add_edge! only in the initialization phase");
    EdgeStore& E = s.e[et];
    if (E.target < 1) return set_error(VH_ERR_ARG, "device Barabasi-Albert: the edge type needs a target type");
    if (has(E.hints, VH_HINT_IGNOREFROM) || has(E.hints, VH_HINT_SINGLEEDGE))
        return set_error(VH_ERR_ARG, "device Barabasi-Albert: the edge type must keep sources and allow many edges");
    if (E.bytes != 0 && E.bytes != 4)
        return set_error(VH_ERR_ARG, "device Barabasi-Albert: edge state must be empty or one Float32 weight");
    if (!E.h_to.empty() || E.gen || E.ba_m || E.stencil)
        return set_error(VH_ERR_ARG, "device Barabasi-Albert: edges already added");
    if (m < 1 || m > 255) return set_error(VH_ERR_ARG, "device Barabasi-Albert: 1 <= m <= 255");
    E.ba_m = m;
    E.gen_seed = graph_seed;
    E.gen_wseed = weight_seed;
    return VH_OK;
}

// Bulk init-phase edges: the G(n, p) graph that vh_graph_gnp + add_edge!
// in target order would give, generated on the device at finish_init!.
extern "C" int vh_add_edges_gnp(vh_sim_t h, uint8_t et, double mean_deg, uint64_t graph_seed,
                                uint64_t weight_seed) {
    if (!h) return set_error(VH_ERR_ARG, "null simulation");
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    if (s.initialized)
        return set_error(VH_ERR_ASSERT, "You can call add_edge! only in the initialization phase");
    EdgeStore& E = s.e[et];
    if (s.m->na != 1 || E.target != 1)
        return set_error(VH_ERR_ARG, "device G(n, p): one agent type, edges targeting it");
    if (has(E.hints, VH_HINT_IGNOREFROM) || has(E.hints, VH_HINT_SINGLEEDGE))
        return set_error(VH_ERR_ARG, "device G(n, p): the edge type must keep sources and allow many edges");
    if (E.bytes != 0 && E.bytes != 4)
        return set_error(VH_ERR_ARG, "device G(n, p): edge state must be empty or one Float32 weight");
    if (!E.h_to.empty() || E.gen) return set_error(VH_ERR_ARG, "device G(n, p): edges already added");
    if (!(mean_deg >= 0.0)) return set_error(VH_ERR_ARG, "device G(n, p): mean degree");
    E.gen = true;
    E.gen_mean_deg = mean_deg;
    E.gen_seed = graph_seed;
    E.gen_wseed = weight_seed;
    return VH_OK;
}

// connect_raster_neighbors! (src/Raster.jl:134-164) as one request: the
// stencil edges of a raster whose cells are consecutive agents of one type,
// generated on the device at finish_init! (the CSR that adding them one by
// one gives). Offsets as `_stencil_core` (src/Raster.jl:80-96): the cube
// [-d, d]^n, d = floor(distance), first component fastest, without the
// origin, Euclidean / Manhattan filtered by <= distance.
extern "C" int vh_add_edges_stencil(vh_sim_t h, uint8_t et, uint32_t cell_type, uint64_t first_nr,
                                    uint32_t ndims, const int64_t* dims, double distance, int metric,
                                    int periodic, const uint8_t* edge_state) {
    if (!h || !dims) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (et >= s.m->ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge type");
    if (s.initialized)
        return set_error(VH_ERR_ASSERT, "connect_raster_neighbors! before finish_init! only");
    EdgeStore& E = s.e[et];
    if (!E.h_to.empty() || E.gen || E.stencil)
        return set_error(VH_ERR_ARG, "raster stencil: the edge type already has edges");
    if (has(E.hints, VH_HINT_IGNOREFROM) || has(E.hints, VH_HINT_SINGLEEDGE) || E.bytes > 16)
        return set_error(VH_ERR_ARG, "raster stencil: sources kept, many edges, state <= 16 bytes");
    if (cell_type < 1 || (int)cell_type > s.m->na) return set_error(VH_ERR_UNREGISTERED, "cell type");
    if (has(E.hints, VH_HINT_SINGLETYPE) && E.target != (int32_t)cell_type)
        return set_error(VH_ERR_ASSERT, "raster stencil: the edge type targets another agent type");
    if (ndims < 1 || ndims > 4) return set_error(VH_ERR_ARG, "raster stencil: 1 to 4 dimensions");
    StencilSpec sp;
    memset(&sp, 0, sizeof sp);
    uint64_t ncells = 1;
    for (uint32_t d = 0; d < 4; d++) {
        sp.dims[d] = d < ndims ? dims[d] : 1;
        if (sp.dims[d] < 1) return set_error(VH_ERR_ARG, "raster stencil: empty dimension");
        ncells *= (uint64_t)sp.dims[d];
    }
    if (first_nr < 1 || first_nr - 1 + ncells > s.a[cell_type].nextid - 1)
        return set_error(VH_ERR_ARG, "raster stencil: the cells are not all added");
    if (metric < 0 || metric > 2 || !(distance >= 0.0) || distance > 127.0)
        return set_error(VH_ERR_ARG, "raster stencil: metric / distance");
    const int64_t d = (int64_t)floor(distance);
    int64_t side = 2 * d + 1, total = 1;
    for (uint32_t k = 0; k < ndims; k++) total *= side;
    uint32_t nst = 0;
    for (int64_t idx = 0; idx < total; idx++) {
        int64_t o[4] = {0, 0, 0, 0}, q = idx, sq = 0, sa = 0;
        bool origin = true;
        for (uint32_t k = 0; k < ndims; k++) {
            o[k] = q % side - d;
            q /= side;
            origin = origin && o[k] == 0;
            sq += o[k] * o[k];
            sa += o[k] < 0 ? -o[k] : o[k];
        }
        if (origin) continue;
        if (metric == 1 && sqrt((double)sq) > distance) continue;
        if (metric == 2 && (double)sa > distance) continue;
        if (nst >= (uint32_t)VH_STENCIL_MAX)
            return set_error(VH_ERR_CAPACITY, "raster stencil: more than %d offsets", VH_STENCIL_MAX);
        for (uint32_t k = 0; k < 4; k++) sp.off[nst][k] = (int8_t)o[k];
        nst++;
    }
    sp.nst = nst;
    sp.ndims = ndims;
    sp.first_nr = first_nr;
    sp.cell_type = cell_type;
    sp.periodic = periodic ? 1u : 0u;
    if (edge_state && E.bytes) memcpy(sp.state, edge_state, E.bytes);
    E.sspec = sp;
    E.stencil = true;
    return VH_OK;
}

static int build_stencil(Sim& s, EdgeStore& E) {
    RowLayout L;
    uint64_t rows;
    compute_layout(s, E, L, rows);
    if (rows >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "more than 2^32 target rows");
    const StencilSpec& sp = E.sspec;
    RC(stencil_csr(s, E, L.base[sp.cell_type] + sp.first_nr - 1, rows));
    for (int t = 0; t <= VH_MAXA; t++) {
        E.rowbase[t] = L.base[t];
        E.nrows[t] = L.nrows[t];
    }
    E.rows = rows;
    E.version++;
    E.stencil = false;
    E.stencil_version = E.version;
    return VH_OK;
}

// the CSR of a device-generated edge type: rows = the agents of type 1
static int build_gnp_csr(Sim& s, EdgeStore& E) {
    hipStream_t st = s.stream;
    RowLayout L;
    uint64_t rows;
    compute_layout(s, E, L, rows);
    uint64_t ne = 0;
    if (rows) {
        DevBuf cdf;
        int rc = gnp_device(rows, E.gen_mean_deg, E.gen_seed, E.off, E.src, cdf, s.scratch, st, &ne);
        if (!rc && E.bytes) {
            rc = E.st.ensure(ne * 4 + 8, st);
            if (!rc) fill_weights(E.st.as<float>(), ne, E.gen_wseed, st);
        }
        HIPCHECK(hipStreamSynchronize(st));
        cdf.release();
        RC(rc);
    } else {
        RC(E.off.ensure(8, st));
        fill_u32(E.off.as<uint32_t>(), 0, 1, st);
    }
    for (int t = 0; t <= VH_MAXA; t++) {
        E.rowbase[t] = L.base[t];
        E.nrows[t] = L.nrows[t];
    }
    E.rows = rows;
    E.nedges = ne;
    E.src64 = false;
    E.src_type = 1;
    E.version++;
    E.gen = false;
    return VH_OK;
}

// edges 2k: a_k -> b_k and 2k+1: b_k -> a_k of the Barabasi-Albert pairs, as
// AgentIDs of type t (numbers through `map` when given: map[nr - 1])
__global__ void k_ba_log(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t K, uint32_t t,
                         const uint32_t* __restrict__ map, uint64_t* __restrict__ to, uint64_t* __restrict__ from) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 2 * K) return;
    uint32_t x = a[j >> 1], y = b[j >> 1];
    if (map) {
        x = map[x - 1];
        y = map[y - 1];
    }
    const uint32_t f = (j & 1) ? y : x, g = (j & 1) ? x : y;
    to[j] = vh_make_id(t, g);
    from[j] = vh_make_id(t, f);
}
// the same edges as 1-based number pairs (ea = from, eb = to) at `ea`, `eb`
__global__ void k_ba_ends(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint64_t K,
                          uint32_t* __restrict__ ea, uint32_t* __restrict__ eb) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= 2 * K) return;
    const uint32_t x = a[j >> 1], y = b[j >> 1];
    ea[j] = (j & 1) ? y : x;
    eb[j] = (j & 1) ? x : y;
}

static int ba_agents(Sim& s, EdgeStore& E, uint64_t* n, uint64_t* K) {
    *n = s.a[E.target].nextid - 1;
    if (*n < E.ba_m) return set_error(VH_ERR_ARG, "device Barabasi-Albert: fewer agents than m");
    *K = ba_pairs(*n, E.ba_m);
    return VH_OK;
}

// right after a scratch reset: room for the pairs, the log, the weights, the
// rounds' state and build_csr's sort of the 2K edges
static int ba_reserve(Sim& s, EdgeStore& E) {
    uint64_t n, K;
    RC(ba_agents(s, E, &n, &K));
    if (s.scratch.reserve(128 * K + 8 * n + ((size_t)64 << 20), s.stream))
        return set_error(VH_ERR_OOM, "scratch for the Barabasi-Albert graph");
    return VH_OK;
}

static int build_ba_csr(Sim& s, EdgeStore& E) {
    hipStream_t st = s.stream;
    Scratch& sc = s.scratch;
    uint64_t n, K;
    RC(ba_agents(s, E, &n, &K));
    const bool have = E.ba_pairs_ok;   // generated by the numbering
    uint32_t* a = have ? E.ba_pairs.as<uint32_t>() : (uint32_t*)sc.take(K * 4 + 4);
    uint32_t* b = have ? a + K : (uint32_t*)sc.take(K * 4 + 4);
    uint64_t* to = (uint64_t*)sc.take(2 * K * 8 + 8);
    uint64_t* fr = (uint64_t*)sc.take(2 * K * 8 + 8);
    float* w = E.bytes ? (float*)sc.take(2 * K * 4 + 8) : nullptr;
    if (!a || !b || !to || !fr || (E.bytes && !w)) return set_error(VH_ERR_OOM, "scratch");
    if (!have) RC(ba_device(n, E.ba_m, E.gen_seed, a, b, sc, st));
    const uint32_t* map =
        (s.relabel_type == E.target && s.relabel_map.p) ? s.relabel_map.as<uint32_t>() : nullptr;
    if (K)
        hipLaunchKernelGGL(k_ba_log, dim3(grid_for(2 * K, 256)), dim3(256), 0, st, a, b, K, (uint32_t)E.target,
                           map, to, fr);
    if (w) fill_weights(w, 2 * K, E.gen_wseed, st);
    RC(build_csr(s, E, false, to, fr, (const uint8_t*)w, 2 * K));
    HIPCHECK(hipStreamSynchronize(st));
    E.ba_pairs.release();
    E.ba_pairs_ok = false;
    E.ba_m = 0;
    return VH_OK;
}

// finish_init! with the agents of type t renumbered for locality, as the
// reference renumbers them when it distributes them (src/Simulation.jl:414-453,
// `return_idmapping`): vh_bfs_order over the init edges among agents of
// type t. States travel with their agents; every edge keeps its place in the
// insertion order, so each target keeps its in-edges in the same order.
// new_nr[k] (k = old number - 1) receives the new number. AgentIDs held
// inside agent or edge states are not renumbered.
extern "C" int vh_finish_init_relabel(vh_sim_t h, uint8_t t, uint64_t* new_nr) {
    if (!h || !new_nr) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (s.initialized)
        return set_error(VH_ERR_ASSERT, "You can not call finish_init! twice for the same simulation");
    if (t < 1 || t > s.m->na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent type");
    AgentStore& A = s.a[t];
    const uint64_t n = A.nextid - 1;
    if (n >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "too many agents");
    // the endpoints of every init edge among agents of type t, in edge-type
    // and insertion order (device-generated Barabasi-Albert edges in their
    // edge type's place), numbered breadth-first on the device (bfs_device,
    // the order of vh_bfs_order)
    hipStream_t st = s.stream;
    std::vector<std::vector<uint32_t>> ha(s.m->ne), hb(s.m->ne);
    std::vector<uint64_t> cnt(s.m->ne, 0), bak(s.m->ne, 0);
    uint64_t M = 0;
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        if (has(E.hints, VH_HINT_IGNOREFROM)) continue;
        if (E.ba_m && E.target == (int)t) {
            uint64_t nn, K;
            RC(ba_agents(s, E, &nn, &K));
            bak[et] = K;
            cnt[et] = 2 * K;
        } else {
            for (size_t k = 0; k < E.h_to.size(); k++) {
                uint64_t to = E.h_to[k], fr = E.h_from[k];
                if (vh_type_of(to) != t || vh_type_of(fr) != t) continue;
                ha[et].push_back((uint32_t)vh_nr_of(fr));
                hb[et].push_back((uint32_t)vh_nr_of(to));
            }
            cnt[et] = ha[et].size();
        }
        M += cnt[et];
    }
    Scratch& sc = s.scratch;
    sc.reset(st);
    uint64_t bak_max = 0;
    for (uint64_t k : bak) bak_max = k > bak_max ? k : bak_max;
    if (sc.reserve(8 * M + 8 * bak_max + 5 * n + bfs_scratch_bytes(n, M) + ((size_t)1 << 20), st))
        return set_error(VH_ERR_OOM, "scratch for the breadth-first numbering");
    uint32_t* ea = (uint32_t*)sc.take(M * 4 + 4);
    uint32_t* eb = (uint32_t*)sc.take(M * 4 + 4);
    if (!ea || !eb) return set_error(VH_ERR_OOM, "scratch");
    uint64_t o = 0;
    for (int et = 0; et < s.m->ne; et++) {
        if (!cnt[et]) continue;
        if (bak[et]) {
            // the pairs are kept for finish_init!'s CSR build
            const uint64_t K = bak[et];
            EdgeStore& E = s.e[et];
            RC(E.ba_pairs.ensure(2 * K * 4 + 8, st));
            uint32_t* pa = E.ba_pairs.as<uint32_t>();
            uint32_t* pb = pa + K;
            RC(ba_device(n, E.ba_m, E.gen_seed, pa, pb, sc, st));
            E.ba_pairs_ok = true;
            hipLaunchKernelGGL(k_ba_ends, dim3(grid_for(2 * K, 256)), dim3(256), 0, st, pa, pb, K, ea + o, eb + o);
        } else {
            HIPCHECK(hipMemcpyAsync(ea + o, ha[et].data(), cnt[et] * 4, hipMemcpyHostToDevice, st));
            HIPCHECK(hipMemcpyAsync(eb + o, hb[et].data(), cnt[et] * 4, hipMemcpyHostToDevice, st));
        }
        o += cnt[et];
    }
    RC(s.relabel_map.ensure(n * 4 + 4, st));
    RC(bfs_device(n, M, ea, eb, s.relabel_map.as<uint32_t>(), sc, st));
    s.relabel_type = t;
    {
        std::vector<uint32_t> nr32(n);
        if (n) HIPCHECK(hipMemcpyAsync(nr32.data(), s.relabel_map.p, n * 4, hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        for (uint64_t k = 0; k < n; k++) new_nr[k] = nr32[k];
    }
    sc.reset(st);
    ha.clear();
    hb.clear();
    // states to their new slots
    if (A.bytes && n) {
        std::vector<uint8_t> ns(n * A.bytes);
        for (uint64_t o = 0; o < n; o++)
            memcpy(ns.data() + (new_nr[o] - 1) * A.bytes, A.h_state.data() + o * A.bytes, A.bytes);
        A.h_state.swap(ns);
    }
    // edge endpoints of type t renumbered, edges in place
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        for (auto* v : {&E.h_to, &E.h_from})
            for (auto& id : *v)
                if (id && vh_type_of(id) == t) id = vh_make_id(t, new_nr[vh_nr_of(id) - 1]);
    }
    const int rc = vh_finish_init(h);
    s.relabel_map.release();
    s.relabel_type = 0;
    return rc;
}

// finish_init! src/Simulation.jl:403-476 for one process: publish the
// init-phase agents and build every edge type's CSR.
extern "C" int vh_finish_init(vh_sim_t h) {
    Sim& s = h->s;
    if (s.initialized)
        return set_error(VH_ERR_ASSERT, "You can not call finish_init! twice for the same simulation");
    hipStream_t st = s.stream;
    for (int t = 1; t <= s.m->na; t++) {
        AgentStore& A = s.a[t];
        uint64_t n = A.nextid - 1;
        RC(ensure_slots(s, t, n));
        if (n && A.bytes)
            HIPCHECK(hipMemcpyAsync(A.state[A.rd].p, A.h_state.data(), n * A.bytes,
                                    hipMemcpyHostToDevice, st));
        if (A.mortal) {
            fill_u8(A.died.as<uint8_t>(), 0, n, st);
            fill_u8(A.died_now.as<uint8_t>(), 0, n, st);
        }
        A.nslots = n;
        HIPCHECK(hipStreamSynchronize(st));   // h_state is freed next
        std::vector<uint8_t>().swap(A.h_state);
    }
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        uint64_t n = E.h_to.size();
        s.scratch.reset(st);
        if (E.ba_m) RC(ba_reserve(s, E));
        uint64_t* dto = (uint64_t*)s.scratch.take(n * 8 + 8);
        uint64_t* dfrom = has(E.hints, VH_HINT_IGNOREFROM) ? nullptr : (uint64_t*)s.scratch.take(n * 8 + 8);
        uint8_t* dst = E.bytes ? (uint8_t*)s.scratch.take(n * E.bytes + 8) : nullptr;
        if (!dto || (!dfrom && !has(E.hints, VH_HINT_IGNOREFROM)) || (E.bytes && !dst))
            return set_error(VH_ERR_OOM, "scratch");
        if (n) {
            HIPCHECK(hipMemcpyAsync(dto, E.h_to.data(), n * 8, hipMemcpyHostToDevice, st));
            if (dfrom) HIPCHECK(hipMemcpyAsync(dfrom, E.h_from.data(), n * 8, hipMemcpyHostToDevice, st));
            if (dst) HIPCHECK(hipMemcpyAsync(dst, E.h_st.data(), n * E.bytes, hipMemcpyHostToDevice, st));
        }
        if (s.m->na != 1) E.src64 = true;
        if (E.gen)
            RC(build_gnp_csr(s, E));
        else if (E.ba_m)
            RC(build_ba_csr(s, E));
        else if (E.stencil)
            RC(build_stencil(s, E));
        else
            RC(build_csr(s, E, false, dto, dfrom, dst, n));
        HIPCHECK(hipStreamSynchronize(st));
        std::vector<uint64_t>().swap(E.h_to);
        std::vector<uint64_t>().swap(E.h_from);
        std::vector<uint8_t>().swap(E.h_st);
        std::unordered_map<uint64_t, size_t>().swap(E.h_first);
    }
    // degree bins of the neighbour-gathering bodies, so that their first
    // apply! needs no host round trip (rebuilt when the edges change)
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        if (E.rows && E.target >= 1 && !E.src64 && !has(E.hints, VH_HINT_IGNOREFROM))
            RC(build_heavy_bins(s, E, E.target));
    }
    s.scratch.reset(st);
    HIPCHECK(hipStreamSynchronize(st));
    // the arena grown for the init-time sorts (device graphs, numbering) is
    // not kept for the steps, which size it for themselves
    if (s.scratch.bytes_held() > ((size_t)1 << 30)) s.scratch.release();
    uint32_t e = 0;
    HIPCHECK(hipMemcpy(&e, s.d_err.p, 4, hipMemcpyDeviceToHost));
    if (e) return set_error(VH_ERR_ASSERT, "invalid edge target during finish_init! (flags %u)", e);
    s.initialized = true;
    s.num_transitions = 1;
    return VH_OK;
}

// ================================================================ apply!
namespace {

__global__ void k_births_mark(uint8_t* __restrict__ died, const uint32_t* __restrict__ reuse,
                              uint64_t nreuse, uint64_t nextid, uint64_t nb) {
    uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nb) return;
    uint64_t slot = g < nreuse ? reuse[nreuse - 1 - g] : nextid + (g - nreuse);
    died[slot - 1] = 0;
}

// deaths of this step in ascending slot order onto the reuse stack
// (readreuseable = [readreuseable; writereuseable], src/AgentMethods.jl:430)
__global__ void k_push_deaths(const uint8_t* __restrict__ died_now, const uint32_t* __restrict__ pos,
                              uint64_t n, uint32_t* __restrict__ reuse, uint64_t base,
                              uint8_t* __restrict__ died) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !died_now[i]) return;
    reuse[base + pos[i]] = (uint32_t)(i + 1);
    died[i] = 1;
}


// contact rebuild: rows are the callers (edges target their caller); copy
// the kept entries of the fixed-capacity log (sources not dead this step)
__global__ void k_copy_nondead(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

enum Kind { K_PLAIN = 0, K_TWOPASS = 1, K_SIR = 2, K_OPINION = 3, K_CONTACT = 4 };
static int kind_of(uint32_t tr) {
    switch (tr) {
    case VH_TR_NOOP: case VH_TR_SUM_NEIGHBORS: case VH_TR_DIE_IF: case VH_TR_COUNT_EDGES:
    case VH_TR_DIFFUSE:
        return K_PLAIN;
    case VH_TR_ADD_SELF_EDGE: case VH_TR_READD_EDGES: case VH_TR_NEIGHBOR_FOO_EDGES: case VH_TR_SPAWN:
    case VH_TR_MOVE_TO: case VH_TR_REMOVE_EDGES:
        return K_TWOPASS;
    case VH_TR_SIR: return K_SIR;
    case VH_TR_OPINION_F64: case VH_TR_OPINION_F32: return K_OPINION;
    case VH_TR_CONTACT: return K_CONTACT;
    }
    uint32_t f;
    if (user_body_flags(tr, &f) == 0) return (f & (VH_BODY_ADDS_EDGES | VH_BODY_ADDS_AGENTS)) ? K_TWOPASS : K_PLAIN;
    return -1;
}

static void fill_ctx(Sim& s, const vh_apply_desc* d, DCtx& c) {
    memset(&c, 0, sizeof c);
    for (int t = 1; t <= s.m->na; t++) {
        AgentStore& A = s.a[t];
        DAgent& D = c.a[t];
        D.rstate = A.rstate();
        D.wstate = A.wstate();
        D.died = A.mortal ? A.died.as<uint8_t>() : nullptr;
        D.died_now = A.mortal ? A.died_now.as<uint8_t>() : nullptr;
        D.nslots = A.nslots;
        D.bytes = A.bytes;
        D.mortal = A.mortal;
        D.indep = A.indep;
        D.readable = in_mask(d->read_agents, t);
        D.writeable = A.writeable;
        D.reuse = A.reuse.as<uint32_t>();
        D.nreuse = A.nreuse;
        D.nextid = A.nextid;
    }
    for (int et = 0; et < s.m->ne; et++) {
        EdgeStore& E = s.e[et];
        DEdge& D = c.e[et];
        D.off = E.rows ? E.off.as<uint32_t>() : nullptr;
        D.src = E.src.p;
        D.st = E.st.as<uint8_t>();
        for (int t = 0; t <= VH_MAXA; t++) {
            D.rowbase[t] = E.rowbase[t];
            D.nrows[t] = E.nrows[t];
        }
        D.bytes = E.bytes;
        D.hints = E.hints;
        D.src64 = E.src64;
        D.src_type = E.src_type;
        D.readable = in_mask(d->read_edges, et);
        D.writeable = E.writeable;
        D.add_existing = in_mask(d->add_existing_edges, et);
    }
    for (int i = 0; i < 4; i++) {
        c.role_a[i] = d->role_agent[i];
        c.role_e[i] = d->role_edge[i];
    }
    c.err = s.d_err.as<uint32_t>();
    c.ee[0].et = c.ee[1].et = -1;
    c.ea[0].at = -1;
    if (d->params_bytes) memcpy(c.params, d->params, std::min<uint32_t>(d->params_bytes, VH_PARAMS_BYTES));
}

static int upload_ctx(Sim& s, const DCtx& c) {
    if (s.ctx_valid && memcmp(&c, &s.h_ctx, sizeof c) == 0) return VH_OK;
    s.h_ctx = c;
    HIPCHECK(hipMemcpyAsync(s.d_ctx.p, &s.h_ctx, sizeof(DCtx), hipMemcpyHostToDevice, s.stream));
    // the source is pageable host memory: make sure the copy has consumed it
    // before h_ctx can change again
    HIPCHECK(hipStreamSynchronize(s.stream));
    s.ctx_valid = true;
    return VH_OK;
}

static int check_errors(Sim& s) {
    HIPCHECK(hipStreamSynchronize(s.stream));
    uint32_t e = 0;
    HIPCHECK(hipMemcpy(&e, s.d_err.p, 4, hipMemcpyDeviceToHost));
    if (!e) return VH_OK;
    fill_u32(s.d_err.as<uint32_t>(), 0, 1, s.stream);
    HIPCHECK(hipStreamSynchronize(s.stream));
    if (e & VH_DERR_CAP) return set_error(VH_ERR_CAPACITY, "per-caller emission capacity exceeded");
    if (e & VH_DERR_READ_EDGE)
        return set_error(VH_ERR_ASSERT, "You try to access an edge of a type that is not in the "
                                        "`read` argument of the apply! function.");
    if (e & VH_DERR_READ_AGENT)
        return set_error(VH_ERR_ASSERT, "an agent type must be in the `read` argument of the "
                                        "transition function.");
    if (e & VH_DERR_DIED)
        return set_error(VH_ERR_ASSERT, "agentstate was requested for an agent that has been removed");
    if (e & VH_DERR_IMMORTAL)
        return set_error(VH_ERR_ASSERT, "You can not return `nothing` for immortal agents");
    if (e & VH_DERR_WRITE)
        return set_error(VH_ERR_ASSERT, "a type must be in the `write` argument of the transition function.");
    if (e & VH_DERR_HINT)
        return set_error(VH_ERR_ASSERT, "the accessor is not defined for the hint combination");
    if (e & VH_DERR_TARGET)
        return set_error(VH_ERR_ASSERT, "invalid edge target (type or id)");
    if (e & VH_DERR_ADDEX)
        return set_error(VH_ERR_ASSERT, "Edges can not be removed, as the edge type is not in the "
                                        "`add_existing` argument of the transition function.");
    return set_error(VH_ERR_ASSERT, "device assertion flags %u", e);
}

// diffuse whose cells' in-edges are still the device stencil built at
// finish_init!, on an immortal, readable, written (not :Independent) cell
// type called alone: the stencil kernel gives what k_generic<DIFFUSE> gives
static bool diffuse_on_stencil(Sim& s, const vh_apply_desc* d) {
    if (d->transition != VH_TR_DIFFUSE || d->n_call != 1 || d->with_edge >= 0) return false;
    if (!d->params || d->params_bytes < sizeof(vh_raster_params)) return false;
    const int t = d->call[0], et = d->role_edge[0];
    if (et < 0 || et >= s.m->ne) return false;
    const AgentStore& A = s.a[t];
    const EdgeStore& E = s.e[et];
    const vh_raster_params* rp = (const vh_raster_params*)d->params;
    return !A.mortal && A.writeable && !A.indep && A.bytes && rp->field_offset < A.bytes &&
           in_mask(d->read_agents, t) && in_mask(d->read_edges, et) && E.rows &&
           E.stencil_version == E.version && E.target == t && (int)E.sspec.cell_type == t;
}

static int apply_one(Sim& s, const vh_apply_desc* d) {
    Model& m = *s.m;
    hipStream_t st = s.stream;
    if (!s.initialized) return set_error(VH_ERR_ASSERT, "You must call finish_init! before apply!");
    if (d->with_edge >= 0 && (d->with_edge >= m.ne || has(m.e[d->with_edge].hints, VH_HINT_SINGLETYPE)))
        return set_error(VH_ERR_ASSERT, "The `with_edge` keyword can only be used for edgetypes "
                                        "without the :SingleType hint.");
    int kind = kind_of(d->transition);
    if (kind < 0) return set_error(VH_ERR_UNREGISTERED, "unregistered transition %u", d->transition);
    if (d->n_call > VH_MAX_CALL) return set_error(VH_ERR_ARG, "too many call types");
    if (d->params_bytes > VH_PARAMS_BYTES) return set_error(VH_ERR_ARG, "params too large");
    uint32_t callmask = 0;
    for (uint32_t i = 0; i < d->n_call; i++) {
        if (d->call[i] < 1 || d->call[i] > m.na) return set_error(VH_ERR_UNREGISTERED, "unregistered call type");
        callmask |= 1u << d->call[i];
    }
    if (callmask & d->add_existing_agents)
        return set_error(VH_ERR_ASSERT, "a call type can not be element of `add_existing`");
    if ((d->add_existing_agents & ~d->write_agents) || (d->add_existing_edges & ~d->write_edges))
        return set_error(VH_ERR_ASSERT, "a type is in `add_existing` but not in `write`");
    for (int i = 0; i < 4; i++) {
        if (d->role_edge[i] >= m.ne) return set_error(VH_ERR_UNREGISTERED, "unregistered edge role");
        if (d->role_agent[i] > m.na) return set_error(VH_ERR_UNREGISTERED, "unregistered agent role");
    }
    const uint64_t step = s.num_transitions;
    s.scratch.reset(st);

    // ---- prepare_write! (src/AgentMethods.jl:271-298, src/EdgeMethods.jl:639-663)
    for (int t = 1; t <= m.na; t++) {
        AgentStore& A = s.a[t];
        A.writeable = in_mask(d->write_agents, t);
        if (!A.writeable) continue;
        bool ae = in_mask(callmask, t) || in_mask(d->add_existing_agents, t);
        if (!A.mortal && !ae)
            return set_error(VH_ERR_ASSERT, "You can add the type to the `write` argument only when it "
                                            "is also included in the `add_existing` argument, as it "
                                            "has the hint :Immortal.");
        if (!ae) {
            A.nslots = 0;
            A.nextid = 1;
            A.nreuse = 0;
        } else if (!in_mask(callmask, t) && !A.indep && A.nslots && A.bytes) {
            // retained agents of a written, uncalled type: write = read
            HIPCHECK(hipMemcpyAsync(A.wstate(), A.rstate(), A.nslots * A.bytes,
                                    hipMemcpyDeviceToDevice, st));
        }
        if (A.mortal && A.nslots) fill_u8(A.died_now.as<uint8_t>(), 0, A.nslots, st);
    }
    for (int et = 0; et < m.ne; et++) s.e[et].writeable = in_mask(d->write_edges, et);

    DCtx c;
    fill_ctx(s, d, c);
    uint64_t ncallers = 0;
    for (uint32_t i = 0; i < d->n_call; i++) {
        c.cbase[i] = ncallers;
        ncallers += s.a[d->call[i]].nslots;
    }

    bool deaths_possible = false;
    for (int t = 1; t <= m.na; t++)
        if (s.a[t].writeable && s.a[t].mortal && s.a[t].nslots) deaths_possible = true;

    // logs of the two-pass bodies
    uint64_t ne_tot[2] = {0, 0}, nb_tot = 0;
    uint64_t* e_to[2] = {nullptr, nullptr};
    uint8_t* e_kind = nullptr;   // remove_edges! bodies: kind of each log entry
    uint64_t* e_from[2] = {nullptr, nullptr};
    uint8_t* e_st[2] = {nullptr, nullptr};
    int contact_t = -1;
    uint64_t contact_deaths = 0;

    if (kind == K_PLAIN && diffuse_on_stencil(s, d)) {
        const vh_raster_params* rp = (const vh_raster_params*)d->params;
        RC(launch_diffuse_stencil(s, d->call[0], d->role_edge[0], rp->field_offset));
    } else if (kind == K_PLAIN) {
        RC(upload_ctx(s, c));
        for (uint32_t i = 0; i < d->n_call; i++)
            RC(launch_generic(d->transition, s.d_ctx.as<DCtx>(), (int)i, d->call[i],
                              s.a[d->call[i]].nslots, 1, d->with_edge, st));
    } else if (kind == K_TWOPASS) {
        // pass 0 counts the add_edge!/add_agent! calls of every caller, an
        // exclusive scan turns the counts into log positions in call order,
        // pass 1 runs the body again (it is deterministic) and writes there.
        // Edges go to role_edge[0] (role_edge[1] for NEIGHBOR_FOO_EDGES),
        // births to role_agent[0] (SPAWN).
        int erole = d->transition == VH_TR_NEIGHBOR_FOO_EDGES ? d->role_edge[1] : d->role_edge[0];
        // remove_edges! bodies log removals beside the add_edge!s; the
        // SendId body of that test adds to its second edge role; a
        // registered body adds to edge role 0 when it says so, and may
        // remove there
        uint32_t uflags = 0;
        const bool user = user_body_flags(d->transition, &uflags) == 0;
        if (user && !(uflags & VH_BODY_ADDS_EDGES)) erole = -1;
        const bool removals = d->transition == VH_TR_REMOVE_EDGES || (user && (uflags & VH_BODY_ADDS_EDGES));
        if (d->transition == VH_TR_REMOVE_EDGES && d->params && d->params_bytes >= sizeof(vh_test_params) &&
            ((const vh_test_params*)d->params)->value == VH_RM_SEND_ID)
            erole = d->role_edge[1];
        bool births = (d->transition == VH_TR_SPAWN || (user && (uflags & VH_BODY_ADDS_AGENTS))) &&
                      d->role_agent[0] > 0;
        uint32_t* ecnt = nullptr;
        uint64_t* eoff = nullptr;
        uint32_t* bcnt = nullptr;
        uint64_t* boff = nullptr;
        if (erole >= 0) {
            ecnt = (uint32_t*)s.scratch.take(ncallers * 4 + 8);
            eoff = (uint64_t*)s.scratch.take((ncallers + 1) * 8);
            if (!ecnt || !eoff) return set_error(VH_ERR_OOM, "scratch");
            fill_u32(ecnt, 0, ncallers, st);
        }
        if (births) {
            bcnt = (uint32_t*)s.scratch.take(ncallers * 4 + 8);
            boff = (uint64_t*)s.scratch.take((ncallers + 1) * 8);
            if (!bcnt || !boff) return set_error(VH_ERR_OOM, "scratch");
            fill_u32(bcnt, 0, ncallers, st);
        }
        auto set_emit = [&](DCtx& cc) {
            uint64_t nc = 0;
            for (uint32_t i = 0; i < d->n_call; i++) {
                cc.cbase[i] = nc;
                nc += s.a[d->call[i]].nslots;
            }
            if (erole >= 0)
                cc.ee[0] = DEmitEdge{ecnt, eoff, e_to[0], e_from[0], e_st[0], erole, m.e[erole].bytes, e_kind};
            if (births) cc.ea[0] = DEmitAgent{bcnt, boff, d->role_agent[0], 0};
        };
        set_emit(c);
        RC(upload_ctx(s, c));
        for (uint32_t i = 0; i < d->n_call; i++)
            RC(launch_generic(d->transition, s.d_ctx.as<DCtx>(), (int)i, d->call[i],
                              s.a[d->call[i]].nslots, 0, d->with_edge, st));
        if (ecnt) {
            if (scan_excl_u64(ecnt, eoff, ncallers, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
            HIPCHECK(hipMemcpyAsync((uint64_t*)s.d_stats.p + 0, eoff + ncallers, 8,
                                    hipMemcpyDeviceToDevice, st));
        }
        if (bcnt) {
            if (scan_excl_u64(bcnt, boff, ncallers, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
            HIPCHECK(hipMemcpyAsync((uint64_t*)s.d_stats.p + 1, boff + ncallers, 8,
                                    hipMemcpyDeviceToDevice, st));
        }
        RC(readback(s, 2));
        ne_tot[0] = ecnt ? s.h_stats[0] : 0;
        nb_tot = bcnt ? s.h_stats[1] : 0;
        RC(check_errors(s));
        if (ecnt) {
            uint32_t b = m.e[erole].bytes;
            e_to[0] = (uint64_t*)s.scratch.take(ne_tot[0] * 8 + 8);
            e_from[0] = (uint64_t*)s.scratch.take(ne_tot[0] * 8 + 8);
            e_st[0] = b ? (uint8_t*)s.scratch.take(ne_tot[0] * b + 8) : nullptr;
            if (!e_to[0] || !e_from[0] || (b && !e_st[0])) return set_error(VH_ERR_OOM, "scratch");
            if (removals) {
                e_kind = (uint8_t*)s.scratch.take(ne_tot[0] + 8);
                if (!e_kind) return set_error(VH_ERR_OOM, "scratch");
            }
        }
        if (nb_tot) {
            AgentStore& A = s.a[d->role_agent[0]];
            if (!A.writeable)
                return set_error(VH_ERR_ASSERT, "the agent type must be in the `write` argument");
            uint64_t fresh = nb_tot > A.nreuse ? nb_tot - A.nreuse : 0;
            RC(ensure_slots(s, d->role_agent[0], A.nextid - 1 + fresh));
            if (A.mortal && A.nextid - 1 + fresh > A.nslots)
                fill_u8(A.died_now.as<uint8_t>() + A.nslots, 0, A.nextid - 1 + fresh - A.nslots, st);
        }
        fill_ctx(s, d, c);   // buffers may have moved
        set_emit(c);
        RC(upload_ctx(s, c));
        for (uint32_t i = 0; i < d->n_call; i++)
            RC(launch_generic(d->transition, s.d_ctx.as<DCtx>(), (int)i, d->call[i],
                              s.a[d->call[i]].nslots, 1, d->with_edge, st));
        if (nb_tot) {
            AgentStore& A = s.a[d->role_agent[0]];
            uint64_t consumed = std::min<uint64_t>(nb_tot, A.nreuse);
            if (A.mortal)
                hipLaunchKernelGGL(k_births_mark, dim3(grid_for(nb_tot, 256)), dim3(256), 0, st,
                                   A.died.as<uint8_t>(), A.reuse.as<uint32_t>(), A.nreuse, A.nextid,
                                   nb_tot);
            A.nreuse -= consumed;
            A.nextid += nb_tot - consumed;
        }
    } else if (kind == K_SIR) {
        if (d->n_call != 1 || d->role_edge[0] < 0 || d->params_bytes < sizeof(vh_sir_params))
            return set_error(VH_ERR_ARG, "SIR needs one call type, an edge role and vh_sir_params");
        vh_sir_params p;
        memcpy(&p, d->params, sizeof p);
        RC(launch_sir(s, d->call[0], d->role_edge[0], p, step));
    } else if (kind == K_OPINION) {
        if (d->n_call != 1 || d->role_edge[0] < 0 || d->params_bytes < sizeof(vh_opinion_params))
            return set_error(VH_ERR_ARG, "opinion needs one call type, an edge role and vh_opinion_params");
        vh_opinion_params p;
        memcpy(&p, d->params, sizeof p);
        RC(launch_opinion(s, d->call[0], d->role_edge[0], p, d->transition == VH_TR_OPINION_F32, step));
    } else {  // K_CONTACT
        if (d->n_call != 1 || d->role_edge[0] < 0 || d->params_bytes < sizeof(vh_contact_params))
            return set_error(VH_ERR_ARG, "contact needs one call type, an edge role and vh_contact_params");
        vh_contact_params p;
        memcpy(&p, d->params, sizeof p);
        contact_t = d->call[0];
        AgentStore& A = s.a[contact_t];
        if (!A.mortal || A.bytes != 8 || !A.writeable || !in_mask(d->write_edges, d->role_edge[0]) ||
            in_mask(d->add_existing_edges, d->role_edge[0]))
            return set_error(VH_ERR_ARG, "contact: mortal 8-byte agent type in write, edge type "
                                         "rebuilt (in write, not add_existing)");
        RC(contact_apply(s, contact_t, d->role_edge[0], p, step, &contact_deaths));
    }

    // ---- finish_write! for agents: deaths (the contact body did its own
    // births and deaths above)
    DiedNow dn;
    memset(&dn, 0, sizeof dn);
    bool any_death = false;
    if (contact_t > 0 && contact_deaths) {
        any_death = true;
        dn.d[contact_t] = s.a[contact_t].died_now.as<uint8_t>();
    }
    if (deaths_possible) {
        uint32_t* dpos[VH_MAXA + 1] = {nullptr};
        int slot = 0;
        int dslot[VH_MAXA + 1];
        for (int t = 1; t <= m.na; t++) {
            AgentStore& A = s.a[t];
            dslot[t] = -1;
            if (!(A.writeable && A.mortal && A.nslots) || t == contact_t) continue;
            uint32_t* d32 = (uint32_t*)s.scratch.take(A.nslots * 4 + 8);
            dpos[t] = (uint32_t*)s.scratch.take((A.nslots + 1) * 4);
            if (!d32 || !dpos[t]) return set_error(VH_ERR_OOM, "scratch");
            hipLaunchKernelGGL(k_u8_to_u32, dim3(grid_for(A.nslots, 256)), dim3(256), 0, st,
                               A.died_now.as<uint8_t>(), d32, A.nslots);
            if (scan_excl_u32(d32, dpos[t], A.nslots, s.scratch, st)) return set_error(VH_ERR_DEVICE, "scan");
            HIPCHECK(hipMemcpyAsync((uint32_t*)s.d_stats.p + 2 * slot, dpos[t] + A.nslots, 4,
                                    hipMemcpyDeviceToDevice, st));
            dslot[t] = slot++;
        }
        if (slot) RC(readback(s, slot));
        const uint32_t* h32 = (const uint32_t*)s.h_stats;
        for (int t = 1; t <= m.na; t++) {
            AgentStore& A = s.a[t];
            uint64_t nd = dslot[t] >= 0 ? h32[2 * dslot[t]] : 0;
            if (!nd) continue;
            any_death = true;
            RC(A.reuse.ensure((A.nreuse + nd) * 4 + 8, st, A.nreuse * 4));
            hipLaunchKernelGGL(k_push_deaths, dim3(grid_for(A.nslots, 256)), dim3(256), 0, st,
                               A.died_now.as<uint8_t>(), dpos[t], A.nslots, A.reuse.as<uint32_t>(),
                               A.nreuse, A.died.as<uint8_t>());
            A.nreuse += nd;
            dn.d[t] = A.died_now.as<uint8_t>();
        }
    }

    // ---- finish_write! for edges: rebuild written types
    for (int et = 0; et < m.ne; et++) {
        EdgeStore& E = s.e[et];
        if (!E.writeable) continue;
        bool keep_base = in_mask(d->add_existing_edges, et);
        if (kind == K_CONTACT && et == d->role_edge[0]) continue;   // rebuilt by contact_apply
        const uint64_t* lto = nullptr;
        const uint64_t* lfrom = nullptr;
        const uint8_t* lst = nullptr;
        uint64_t nlog = 0;
        const uint8_t* lkind = nullptr;
        if (kind == K_TWOPASS && c.ee[0].cnt && c.ee[0].et == et) {
            lto = e_to[0]; lfrom = e_from[0]; lst = e_st[0]; nlog = ne_tot[0]; lkind = e_kind;
        }
        if (!keep_base || nlog || E.rows == 0)
            RC(build_csr(s, E, keep_base, lto, lfrom, lst, nlog, lkind));
    }
    // deaths remove edges of every edge type (src/AgentMethods.jl:314-345)
    if (any_death)
        for (int et = 0; et < m.ne; et++)
            if (!(kind == K_CONTACT && et == d->role_edge[0])) {
                const uint64_t before = s.e[et].nedges;
                RC(filter_deaths(s, s.e[et], dn));
                if (s.e[et].nedges != before) s.e[et].last_change = step;   // network_changed
            }

    // ---- publish (double-buffer swap instead of the reference's memcpy)
    for (int t = 1; t <= m.na; t++) {
        AgentStore& A = s.a[t];
        if (A.writeable) A.nslots = std::max(A.nslots, A.nextid - 1);
        if (A.writeable && !A.indep) A.rd ^= 1;
        if (A.writeable) A.last_change = step;
        A.writeable = false;
    }
    for (int et = 0; et < m.ne; et++) {
        if (s.e[et].writeable) s.e[et].last_change = step;
        s.e[et].writeable = false;
    }
    s.num_transitions++;
    if (g_asserts) RC(check_errors(s));
    return VH_OK;
}

// ============================================================ params / globals
static uint32_t dt_size(int dt) {
    switch (dt) {
    case VH_DT_I64: case VH_DT_F64: case VH_DT_U64: return 8;
    case VH_DT_F32: case VH_DT_U32: case VH_DT_I32: return 4;
    case VH_DT_U16: return 2;
    case VH_DT_BOOL: case VH_DT_U8: return 1;
    default: return 0;
    }
}
static bool dt_float(int dt) { return dt == VH_DT_F64 || dt == VH_DT_F32; }

// the stored value (raw 8 bytes of type `from`) converted to `to` and
// written to `out`; false when Julia's convert would throw (a non-integral
// or out-of-range value for an integer field)
static bool convert_value(int from, uint64_t raw, int to, void* out) {
    double f = 0.0;
    int64_t i = 0;
    uint64_t u = 0;
    bool isf = dt_float(from), neg = false;
    switch (from) {
    case VH_DT_F64: memcpy(&f, &raw, 8); break;
    case VH_DT_F32: { float g; memcpy(&g, &raw, 4); f = g; break; }
    case VH_DT_I64: memcpy(&i, &raw, 8); neg = i < 0; u = (uint64_t)i; break;
    case VH_DT_I32: { int32_t v; memcpy(&v, &raw, 4); i = v; neg = v < 0; u = (uint64_t)i; break; }
    case VH_DT_U64: memcpy(&u, &raw, 8); i = (int64_t)u; break;
    case VH_DT_U32: { uint32_t v; memcpy(&v, &raw, 4); u = v; i = v; break; }
    case VH_DT_U16: { uint16_t v; memcpy(&v, &raw, 2); u = v; i = v; break; }
    case VH_DT_U8: case VH_DT_BOOL: { uint8_t v; memcpy(&v, &raw, 1); u = v; i = v; break; }
    default: return false;
    }
    if (dt_float(to)) {
        const double v = isf ? f : (neg ? (double)i : (double)u);
        if (to == VH_DT_F64) memcpy(out, &v, 8);
        else { const float g = (float)v; memcpy(out, &g, 4); }
        return true;
    }
    // integer target: the value as a sign and magnitude
    if (isf) {
        if (!(f == floor(f)) || f < -9.2233720368547758e18 || f >= 1.8446744073709552e19) return false;
        neg = f < 0;
        if (neg) i = (int64_t)f; else u = (uint64_t)f;
        if (!neg) i = (int64_t)u;
    }
    auto fits = [&](int64_t lo, uint64_t hi) { return neg ? (lo < 0 && i >= lo) : u <= hi; };
    switch (to) {
    case VH_DT_I64: if (!fits(INT64_MIN, INT64_MAX)) return false; { int64_t v = neg ? i : (int64_t)u; memcpy(out, &v, 8); } return true;
    case VH_DT_U64: if (!fits(0, UINT64_MAX)) return false; memcpy(out, &u, 8); return true;
    case VH_DT_I32: if (!fits(INT32_MIN, INT32_MAX)) return false; { int32_t v = (int32_t)(neg ? i : (int64_t)u); memcpy(out, &v, 4); } return true;
    case VH_DT_U32: if (!fits(0, UINT32_MAX)) return false; { uint32_t v = (uint32_t)u; memcpy(out, &v, 4); } return true;
    case VH_DT_U16: if (!fits(0, UINT16_MAX)) return false; { uint16_t v = (uint16_t)u; memcpy(out, &v, 2); } return true;
    case VH_DT_U8: if (!fits(0, UINT8_MAX)) return false; { uint8_t v = (uint8_t)u; memcpy(out, &v, 1); } return true;
    case VH_DT_BOOL: if (!fits(0, 1)) return false; { uint8_t v = (uint8_t)u; memcpy(out, &v, 1); } return true;
    default: return false;
    }
}

extern "C" int vh_sim_set_value(vh_sim_t h, int kind, uint32_t slot, int dtype, const void* value) {
    if (!h || !value) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if (kind != VH_VALUE_PARAM && kind != VH_VALUE_GLOBAL) return set_error(VH_ERR_ARG, "kind must be 0 or 1");
    if (slot >= VH_MAX_VALUES) return set_error(VH_ERR_ARG, "slot %u >= %d", slot, VH_MAX_VALUES);
    const uint32_t sz = dt_size(dtype);
    if (!sz) return set_error(VH_ERR_ARG, "unknown dtype %d", dtype);
    if (kind == VH_VALUE_PARAM && s.initialized)
        return set_error(VH_ERR_ASSERT,
                         "You can call `set_param!` only before `finish_init!`. For values that change "
                         "while the simulation is running, register them as globals and use "
                         "`set_global!` instead.");
    uint64_t raw = 0;
    memcpy(&raw, value, sz);
    s.vals[kind][slot] = raw;
    s.vdt[kind][slot] = (uint8_t)dtype;
    s.vset[kind] |= 1u << slot;
    return VH_OK;
}

extern "C" int vh_sim_get_value(vh_sim_t h, int kind, uint32_t slot, int* dtype, void* value) {
    if (!h || !value) return set_error(VH_ERR_ARG, "null argument");
    Sim& s = h->s;
    if ((kind != VH_VALUE_PARAM && kind != VH_VALUE_GLOBAL) || slot >= VH_MAX_VALUES ||
        !((s.vset[kind] >> slot) & 1u))
        return set_error(VH_ERR_ARG, "value slot %u of kind %d is not set", slot, kind);
    if (dtype) *dtype = s.vdt[kind][slot];
    memcpy(value, &s.vals[kind][slot], 8);
    return VH_OK;
}

// the desc with its bound blob fields filled in from the simulation's values
// (`blob` holds the copy the returned desc points to)
static int resolve_binds(const Sim& s, const vh_apply_desc* d, vh_apply_desc& out, uint8_t* blob) {
    out = *d;
    if (!d->n_bind) return VH_OK;
    if (!d->bind) return set_error(VH_ERR_ARG, "n_bind > 0 without bind");
    if (d->params_bytes > VH_PARAMS_BYTES || (d->params_bytes && !d->params))
        return set_error(VH_ERR_ARG, "params too large");
    memset(blob, 0, VH_PARAMS_BYTES);
    if (d->params_bytes) memcpy(blob, d->params, d->params_bytes);
    for (uint32_t k = 0; k < d->n_bind; k++) {
        const vh_value_bind& b = d->bind[k];
        const uint32_t sz = dt_size(b.dtype);
        if ((b.kind != VH_VALUE_PARAM && b.kind != VH_VALUE_GLOBAL) || b.slot >= VH_MAX_VALUES || !sz ||
            (uint64_t)b.offset + sz > d->params_bytes)
            return set_error(VH_ERR_ARG, "invalid value binding %u", k);
        if (!((s.vset[b.kind] >> b.slot) & 1u))
            return set_error(VH_ERR_ASSERT, "the %s of slot %u is not set", b.kind ? "global" : "parameter",
                             (unsigned)b.slot);
        if (!convert_value(s.vdt[b.kind][b.slot], s.vals[b.kind][b.slot], b.dtype, blob + b.offset))
            return set_error(VH_ERR_ASSERT, "InexactError: the %s of slot %u does not convert to the field's type",
                             b.kind ? "global" : "parameter", (unsigned)b.slot);
    }
    out.params = blob;
    out.n_bind = 0;
    out.bind = nullptr;
    return VH_OK;
}

extern "C" int vh_apply(vh_sim_t h, const vh_apply_desc* d) {
    if (!h || !d) return set_error(VH_ERR_ARG, "null argument");
    vh_apply_desc r;
    uint8_t blob[VH_PARAMS_BYTES];
    RC(resolve_binds(h->s, d, r, blob));
    RC(apply_one(h->s, &r));
    HIPCHECK(hipEventRecord(h->s.last, h->s.stream));
    return VH_OK;
}
// ============================================================ batched apply!
struct vh_batch {
    hipStream_t stream = nullptr;     // the stream of the last call's simulations
    hipEvent_t done = nullptr;        // end of the last call's launches (staging reuse)
    DevBuf dev;                       // member tables of one call
    void* host = nullptr;             // their pinned staging
    size_t host_cap = 0;
};

extern "C" int vh_batch_create(vh_batch_t* out) {
    if (!out) return set_error(VH_ERR_ARG, "null argument");
    vh_batch* b = new vh_batch();
    hipError_t e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "hipEventCreate");
    }
    *out = b;
    return VH_OK;
}

extern "C" int vh_batch_destroy(vh_batch_t b) {
    if (!b) return VH_OK;
    (void)hipEventSynchronize(b->done);
    b->dev.release();
    if (b->host) (void)hipHostFree(b->host);
    (void)hipEventDestroy(b->done);
    delete b;
    return VH_OK;
}

// apply! of `desc` to every simulation, step by step: simulation i takes
// part in the first steps[i] steps. The one-pass opinion bodies of all
// simulations of a step run as one launch per state type; everything else of
// an apply! (validation, bookkeeping, other bodies) is each simulation's own.
extern "C" int vh_batch_apply(vh_batch_t b, const vh_sim_t* sims, uint32_t n, const vh_apply_desc* desc,
                              const uint32_t* steps) {
    if (!b || !desc || (n && (!sims || !steps))) return set_error(VH_ERR_ARG, "null argument");
    if (!n) return VH_OK;
    hipStream_t st = sims[0]->s.stream;
    uint32_t smax = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!sims[i]) return set_error(VH_ERR_ARG, "null simulation");
        if (sims[i]->s.stream != st) return set_error(VH_ERR_ARG, "batched simulations must share one stream");
        smax = std::max(smax, steps[i]);
    }
    // the previous call's staging is free once its copies ran
    HIPCHECK(hipEventSynchronize(b->done));
    const size_t need = (size_t)smax * 2 * (((size_t)n * sizeof(OpMember) + 255) & ~(size_t)255);
    if (need > b->host_cap) {
        if (b->host) (void)hipHostFree(b->host);
        b->host = nullptr;
        b->host_cap = 0;
        HIPCHECK(hipHostMalloc(&b->host, need, hipHostMallocDefault));
        b->host_cap = need;
    }
    RC(b->dev.ensure(need + 256, st));
    size_t off = 0;
    OpinionDefer col;
    col.stream = st;
    int rc = VH_OK;
    for (uint32_t k = 0; k < smax && rc == VH_OK; k++) {
        g_opinion_defer = &col;
        for (uint32_t i = 0; i < n && rc == VH_OK; i++) {
            if (steps[i] <= k) continue;
            vh_apply_desc r;
            uint8_t blob[VH_PARAMS_BYTES];
            rc = resolve_binds(sims[i]->s, desc, r, blob);
            if (rc == VH_OK) rc = apply_one(sims[i]->s, &r);
        }
        g_opinion_defer = nullptr;
        if (rc == VH_OK) rc = opinion_defer_launch(col, st, b->dev, &b->host, &b->host_cap, &off);
    }
    g_opinion_defer = nullptr;
    RC(rc);
    HIPCHECK(hipEventRecord(b->done, st));
    for (uint32_t i = 0; i < n; i++) HIPCHECK(hipEventRecord(sims[i]->s.last, st));
    return VH_OK;
}

extern "C" int vh_apply_n(vh_sim_t h, const vh_apply_desc* d, uint32_t steps) {
    if (!h || !d) return set_error(VH_ERR_ARG, "null argument");
    vh_apply_desc r;
    uint8_t blob[VH_PARAMS_BYTES];
    RC(resolve_binds(h->s, d, r, blob));
    for (uint32_t i = 0; i < steps; i++) RC(apply_one(h->s, &r));
    HIPCHECK(hipEventRecord(h->s.last, h->s.stream));
    return VH_OK;
}
