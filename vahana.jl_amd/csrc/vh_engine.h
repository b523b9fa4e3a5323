// vh_engine.h — host-side state of models and simulations (internal).
#pragma once
#include <stdint.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "vh_common.h"
#include "vh_prims.h"

namespace vh {

// device buffer with grow-only capacity
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    // ensure capacity for `bytes`; keeps the first `keep` bytes when growing
    int ensure(size_t bytes, hipStream_t st, size_t keep = 0);
    void release();
    template <class T> T* as() const { return (T*)p; }
};

struct Model {
    int na = 0, ne = 0;
    struct AT { uint32_t bytes, hints; } a[VH_MAXA + 1];
    struct ET { uint32_t bytes, hints; int32_t target; int64_t size; } e[VH_MAXE];
};

struct AgentStore {
    uint32_t bytes = 0;
    bool mortal = true, indep = false;
    std::vector<uint8_t> h_state;   // init-phase states (slot order)
    uint64_t nextid = 1;            // next fresh slot (1-based)
    uint64_t nslots = 0;            // length(readstate)
    uint64_t cap = 0;               // device slot capacity
    DevBuf state[2];
    int rd = 0;                     // index of the read buffer
    DevBuf died, died_now;          // uint8 per slot
    DevBuf reuse;                   // LIFO stack of free slots (uint32, bottom first)
    uint64_t nreuse = 0;
    bool writeable = false;
    uint64_t last_change = 0;       // transition of the last write (src/AgentMethods.jl:433)
    const uint8_t* rstate() const { return state[rd].as<uint8_t>(); }
    uint8_t* wstate() const { return indep ? state[rd].as<uint8_t>() : state[rd ^ 1].as<uint8_t>(); }
};

// a raster stencil (connect_raster_neighbors!): cells are the agents
// first_nr .. first_nr + prod(dims) - 1 of cell_type in column-major order
constexpr int VH_STENCIL_MAX = 128;
struct StencilSpec {
    int64_t dims[4];
    uint64_t first_nr;
    uint32_t cell_type, ndims, nst, periodic;
    int8_t off[VH_STENCIL_MAX][4];   // offsets in stencil order
    uint8_t state[16];
};

struct EdgeStore {
    uint32_t bytes = 0, hints = 0;
    int32_t target = 0;
    std::vector<uint64_t> h_from, h_to;   // init-phase edges in add_edge! order
    std::vector<uint8_t> h_st;
    // :SingleEdge init phase: index in h_to of the first edge to each target
    // (kept across add_edge! calls, cleared at finish_init!)
    std::unordered_map<uint64_t, size_t> h_first;
    uint64_t rowbase[VH_MAXA + 1] = {0}, nrows[VH_MAXA + 1] = {0};
    uint64_t rows = 0;             // total rows of the CSR (0: never built)
    uint64_t nedges = 0;
    DevBuf off, src, st;           // CSR: uint32 offsets[rows+1], sources, states
    bool src64 = true;
    uint32_t src_type = 0;
    uint64_t version = 0;          // bumped on every rebuild
    // degree bins of the opinion kernels, valid for `bins_version`
    DevBuf heavy;                  // uint32 rows with degree > 64
    uint64_t nheavy = 0, bins_version = ~0ull;
    // gather plan of the opinion kernels (build_gather_plan), valid for
    // `plan_version` with windows of plan_wsz sources, plan_n agents and
    // states of plan_tb bytes
    DevBuf psrc, pval;             // per plan entry: local source index (uint16); gathered states
    DevBuf wstart;                 // first plan entry of each window
    DevBuf ablk, lbase;            // per block: first agent; first light-list entry
    DevBuf tpk, twt;               // light lists: slot | run start, weight
    DevBuf rbase, prun, pfirst;    // runs: first run per block; plan position, list index in the block
    DevBuf cbase;                  // run starts before each 64-entry chunk of a block's list (by wave)
    DevBuf wq;                     // pass-2 queue: next block, finished launch blocks
    DevBuf huge;                   // rows of in-degree above the plan's staging limit
    DevBuf hbase, hty, hslot;      // their chunk lists: first entry per row; plan position, slot
    DevBuf hinfo;                  // per heavy row: row, CSR start, degree, hbase (one 16-byte load)
    bool plan_ok = false;          // the lists fit their sort keys (else the one-pass kernel)
    uint64_t plan_version = ~0ull, plan_n = 0;
    uint32_t plan_wsz = 0, plan_tb = 0, plan_nb = 0, plan_w = 0;
    uint64_t plan_nhuge = 0;
    // out-edges of the SIR exposure pass (build_out_csr), valid for
    // `out_version`; `expo`: one bit per agent, set when it has an infected
    // in-neighbour in this step; `expo_cnt`: infected counts by step parity
    DevBuf out_off, out_dst, expo, expo_cnt;
    uint64_t out_version = ~0ull;
    // every device buffer of the store (release, accounting)
    template <class F>
    void each_buf(F f) {
        DevBuf* all[] = {&off,  &src,   &st,    &heavy, &psrc, &pval,    &wstart,  &ablk, &lbase,    &tpk,
                         &twt,  &rbase, &prun,  &pfirst, &cbase, &wq,     &huge,    &hbase, &hty,     &hslot,
                         &hinfo,
                         &out_off, &out_dst, &expo, &expo_cnt, &ba_pairs};
        for (DevBuf* b : all) f(*b);
    }
    bool writeable = false;
    // transition of the last write or death removal (src/EdgeMethods.jl:682,
    // src/AgentMethods.jl:347-353)
    uint64_t last_change = 0;
    // in-edges requested as a raster stencil during init (vh_add_edges_stencil)
    bool stencil = false;
    StencilSpec sspec;
    uint64_t stencil_version = ~0ull;   // the CSR is the stencil of sspec while version equals it
    // in-edges requested as a device G(n, p) during init (vh_add_edges_gnp)
    bool gen = false;
    double gen_mean_deg = 0.0;
    uint64_t gen_seed = 0, gen_wseed = 0;
    // in-edges requested as a device Barabasi-Albert graph, both directions
    // of every pair (vh_add_edges_ba; seeds in gen_seed / gen_wseed)
    uint32_t ba_m = 0;             // 0: none
    DevBuf ba_pairs;               // its pairs when the numbering generated them (a[K], then b[K])
    bool ba_pairs_ok = false;
};

struct Sim {
    // the simulation's own copy of its model's types: a model handle may be
    // destroyed while simulations of it live
    Model model{};
    Model* m = nullptr;
    int device = 0;                // the device current at vh_sim_create
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t io = nullptr;      // reads, when set (vh_set_io_stream)
    hipEvent_t last = nullptr;     // end of the last apply! on `stream`
    AgentStore a[VH_MAXA + 1];
    EdgeStore e[VH_MAXE];
    bool initialized = false;
    uint64_t num_transitions = 0;
    DCtx h_ctx;                    // last uploaded context
    DevBuf d_ctx;
    bool ctx_valid = false;
    DevBuf d_err;                  // uint32 error word
    DevBuf d_stats;                // small readback area
    uint64_t* h_stats = nullptr;   // pinned mirror of d_stats
    Scratch scratch;
    // param / get_global values (vh_sim_set_value): raw 8 bytes and vh_dtype
    uint64_t vals[2][VH_MAX_VALUES] = {};
    uint8_t vdt[2][VH_MAX_VALUES] = {};
    uint32_t vset[2] = {0, 0};     // slots set, per kind
    // vh_finish_init_relabel: the new numbers of the agents of relabel_type
    // (uint32, 1-based, by old number - 1), applied to device-generated edges
    DevBuf relabel_map;
    int relabel_type = 0;
};

// model-body launchers (vh_bodies.hip)
int launch_generic(int transition, const DCtx* d_ctx, int ci, int call_t, uint64_t nslots, int pass,
                   int with_edge, hipStream_t st);
int launch_sir(Sim& s, int t, int et, const vh_sir_params& p, uint64_t step);
int launch_opinion(Sim& s, int t, int et, const vh_opinion_params& p, bool f32, uint64_t step);
// one-pass opinion launches of many simulations collected into one fused
// launch per state type (vh_batch_apply): while `g_opinion_defer` is set,
// launch_opinion records a simulation's launch instead of enqueuing it when
// the simulation runs on the collector's stream and needs no plan
struct OpMember {
    const uint32_t* off;
    const uint32_t* src;
    const void* x;
    const float* w;
    void* xo;
    const uint32_t* heavy;
    uint64_t n, nrows, rowbase, nheavy;
    double eps;
    uint32_t blk0, nh;   // first block of the member in the launch; its heavy-wave blocks
};
struct OpinionDefer {
    hipStream_t stream = nullptr;
    std::vector<OpMember> m[2];   // [0] Float64, [1] Float32 states
};
extern thread_local OpinionDefer* g_opinion_defer;
// enqueue the fused launches of the collected members on `st`, member tables
// staged through `host` (pinned) into `dev` (both grown as needed)
int opinion_defer_launch(OpinionDefer& c, hipStream_t st, DevBuf& dev, void** host, size_t* host_cap,
                         size_t* host_off);
// diffuse over an unchanged device stencil: neighbours from the stencil
// arithmetic instead of the CSR (same neighbour set; the body is an OR)
int launch_diffuse_stencil(Sim& s, int t, int et, uint32_t field);
// contact model: per-step flag bitmaps, one bit per slot (64 per word)
struct ContactBits {
    uint64_t* sp;     // dead in the read state or dying in this step (padding set)
    uint64_t* dead;   // dead in the read state (padding set)
    uint64_t* inf;    // infected in the read state
    uint32_t* dcnt;   // deaths per word
    uint64_t nwords;
};
// one apply! of the contact model incl. its births, deaths and contact CSR
int contact_apply(Sim& s, int t, int et, const vh_contact_params& p, uint64_t step, uint64_t* ndeaths);
// device G(n, p) CSR in target order (vh_ensemble.hip): off[n + 1], sources as
// 1-based numbers; one host synchronisation
int gnp_device(uint64_t n, double mean_deg, uint64_t seed, DevBuf& off, DevBuf& src, DevBuf& cdf,
               Scratch& sc, hipStream_t st, uint64_t* nedges);
// CSR of a raster stencil (vh_bodies.hip): off[rows + 1] with the cell rows
// at rowbase, sources as AgentIDs (src64) or numbers, uniform states
int stencil_csr(Sim& s, EdgeStore& E, uint64_t rowbase, uint64_t rows);
// bodies registered at run time (vh_userbody.hip): flags (-1: not one) and
// the launch of its agent-loop kernel
int user_body_flags(uint32_t tr, uint32_t* flags);
int launch_user_body(uint32_t tr, const DCtx* d_ctx, int ci, int call_t, uint64_t nslots, int pass,
                     int with_edge, hipStream_t st);
// Barabasi-Albert pairs on the device (vh_graphgen.hip), equal to
// vh_graph_ba: a[k], b[k] for k < ba_pairs(n, m)
uint64_t ba_pairs(uint64_t n, uint32_t m);
int ba_device(uint64_t n, uint32_t m, uint64_t seed, uint32_t* a, uint32_t* b, Scratch& sc, hipStream_t st);
// breadth-first numbering on the device, equal to vh_bfs_order: newnr[v - 1]
// for the undirected edges (ea[j], eb[j]) (1-based)
int bfs_device(uint64_t n, uint64_t M, const uint32_t* ea, const uint32_t* eb, uint32_t* newnr, Scratch& sc,
               hipStream_t st);
// scratch bytes bfs_device takes (reserve them beforehand: it takes many pieces)
size_t bfs_scratch_bytes(uint64_t n, uint64_t M);
// Float32 weights in (0, 1) by CSR position (vh_ensemble.hip)
void fill_weights(float* w, uint64_t n, uint64_t wseed, hipStream_t st);
// host graph helper (vh_graph.cpp)
extern "C" int vh_bfs_order(uint64_t n, uint64_t m, const uint32_t* a, const uint32_t* b,
                            uint64_t* new_nr);
// engine services the bodies use
int readback(Sim& s, int n);                           // d_stats -> h_stats, synchronous
int ensure_slots(Sim& s, int t, uint64_t slots);
// degree bins for the opinion kernels
int build_heavy_bins(Sim& s, EdgeStore& E, int t);

// process-wide kernel options (vh_set_options)
extern vh_options g_opts;

int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* where);

#define HIPCHECK(x)                                        \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) return vh::hip_fail(e_, #x); \
    } while (0)
#define RC(x)              \
    do {                   \
        int r_ = (x);      \
        if (r_) return r_; \
    } while (0)

}  // namespace vh

struct vh_model {
    vh::Model m;
};
struct vh_sim {
    vh::Sim s;
};
