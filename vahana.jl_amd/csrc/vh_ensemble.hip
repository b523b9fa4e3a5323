// vh_ensemble.hip — many SIR simulations advanced by one segmented launch
// per step (ensembles / parameter sweeps, and the step batches of the serving
// loop).  Every member keeps its own device buffers; a per-step member table
// (member descriptors + a block map) lets members join and leave between
// steps.  Each member's result is bit-identical to the same simulation run
// alone through apply!: the body, its counter-hash keys (seed, AgentID, step,
// edge position) and the CSR order are the same.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "vh_engine.h"

using namespace vh;

namespace {

constexpr uint64_t KEY_DEG = 11ull << 40;
constexpr uint64_t KEY_SRC = 12ull << 40;

// inverse-CDF table of Binomial(n - 1, p); the last entry is forced to 2.0
std::vector<double> binom_cdf(uint64_t n, double p) {
    std::vector<double> cdf;
    if (n < 2 || p <= 0.0) {
        cdf.push_back(2.0);
        return cdf;
    }
    double N = (double)(n - 1);
    double lp = log(p), lq = log1p(-p), acc = 0.0;
    for (uint64_t k = 0; k <= n - 1; k++) {
        double lpmf = lgamma(N + 1.0) - lgamma((double)k + 1.0) - lgamma(N - (double)k + 1.0) +
                      (double)k * lp + (N - (double)k) * lq;
        acc += exp(lpmf);
        cdf.push_back(acc);
        if (k > (uint64_t)(N * p) && 1.0 - acc < 1e-18) break;
        if (cdf.size() >= 4096) break;
    }
    cdf.back() = 2.0;
    return cdf;
}

__host__ __device__ inline uint32_t gnp_degree(const double* cdf, uint32_t ncdf, uint64_t seed,
                                               uint64_t t) {
    double u = vh_u01(vh_hash(seed, t, 0, KEY_DEG));
    uint32_t d = 0;
    while (d + 1 < ncdf && !(u < cdf[d])) d++;
    return d;
}
// the d distinct sources of target t (1-based), in draw order
__host__ __device__ inline void gnp_sources(uint64_t n, uint64_t seed, uint64_t t, uint32_t d,
                                            uint32_t* out) {
    uint64_t draw = 0;
    for (uint32_t j = 0; j < d;) {
        uint64_t h = vh_hash(seed, t, ++draw, KEY_SRC);
        uint64_t s = 1 + (((h >> 32) * (n - 1)) >> 32);
        if (s >= t) s++;
        bool dup = false;
        for (uint32_t q = 0; q < j; q++) dup |= out[q] == (uint32_t)s;
        if (dup) continue;
        out[j++] = (uint32_t)s;
    }
}

__global__ void k_gnp_degree(const double* cdf, uint32_t ncdf, uint64_t seed, uint64_t n,
                             uint32_t* deg) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) deg[i] = gnp_degree(cdf, ncdf, seed, i + 1);
}
__global__ void k_gnp_fill(uint64_t n, uint64_t seed, const uint32_t* off, uint32_t* src) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) gnp_sources(n, seed, i + 1, off[i + 1] - off[i], src + off[i]);
}
// initial state: I with probability n_init / n (hash keyed like the solo input)
__global__ void k_sir_init(uint32_t* st, uint64_t n, uint64_t seed, double frac, uint32_t type) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t id = vh_make_id(type, i + 1);
    st[i] = vh_u01(vh_hash(seed, id, 0, VH_KEY_INIT)) < frac ? (uint32_t)VH_I : (uint32_t)VH_S;
}

// in-edge rows as transpose input: key = source index, value = row
__global__ void k_member_edge_rows(const uint32_t* __restrict__ off, const uint32_t* __restrict__ src,
                                   uint64_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (uint32_t k = off[r]; k < off[r + 1]; k++) {
        keys[k] = src[k] - 1u;
        vals[k] = (uint32_t)r;
    }
}
// out_off[v] = first index whose key >= v (keys sorted), v in [0, n]
__global__ void k_member_offsets(const uint32_t* __restrict__ keys, uint64_t ne, uint64_t n,
                                 uint32_t* __restrict__ out_off) {
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    uint64_t lo = 0, hi = ne;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    out_off[v] = (uint32_t)lo;
}

constexpr int ENS_CLASSES = 3;   // member size classes of the tile choice

struct MemberDesc {
    const uint32_t* off;
    const uint32_t* src;
    const uint32_t* rs;
    uint32_t* ws;
    uint64_t n;
    uint64_t step;
    uint64_t seed;
    double p;
    uint32_t rec, type;
    const uint32_t* out_off;   // out-edges (exposure pass)
    const uint32_t* out_dst;
    uint32_t* bits;            // two exposure bitmaps of nw words (launch parity)
    uint64_t nw;
    uint32_t tile;             // agents per logical block (the launch block size times 1, 2 or 4)
    uint32_t pad;
};

// Exposure (as in the solo SIR step): per launch, k_ens_decide turns the
// member's infected count of the previous launch (per-block counts stored by
// k_sir_ensemble, all present) into dec[member]: bit 0 on while fewer than
// 1/32 of its agents are infected; bit 1 "full" when the previous launch was
// off or the launch number is a multiple of EXPO_FULL. With dec on,
// k_ens_expose sets the bit of every out-neighbour of an infected agent in
// the launch's bitmap (full), or only of the agents infected in the previous
// launch (days 0) on top of the previous launch's bitmap, and k_sir_ensemble
// lets a susceptible agent without its bit skip the draws (no infected
// in-neighbour: no infection). Every infected agent was either infected since
// the last full launch or marked in it, so the bitmap covers all of them; bits
// of agents that recovered since stay until the next full launch, and a stale
// bit or count can only make an agent draw, never skip wrongly.
// k_sir_ensemble prepares the other bitmap's words of its agents for the
// next launch: cleared before a full launch, else a copy of this launch's.
constexpr uint64_t EXPO_FULL = 8;   // a full exposure launch at least every 8 launches
__global__ __launch_bounds__(256) void k_ens_decide(const MemberDesc* __restrict__ md,
                                                    const uint32_t* __restrict__ mblk,
                                                    const uint32_t* __restrict__ blkcnt, uint32_t nblocks,
                                                    uint32_t parity, uint64_t lidx, uint32_t* __restrict__ dec) {
    __shared__ uint64_t s_tot[4];
    __shared__ uint32_t s_bad[4];
    const uint32_t mi = blockIdx.x, tid = threadIdx.x;
    const uint32_t b0 = mblk[2 * mi], nb = mblk[2 * mi + 1];
    uint64_t tot = 0;
    uint32_t bad = 0;
    for (uint32_t q = tid; q < nb; q += 256) {
        const uint32_t c = blkcnt[(uint64_t)parity * nblocks + b0 + q];
        bad |= c == 0xFFFFFFFFu ? 1u : 0u;
        tot += c == 0xFFFFFFFFu ? 0u : (c & 0xFFFFu);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        tot += __shfl_xor(tot, o, 64);
        bad |= __shfl_xor(bad, o, 64);
    }
    if ((tid & 63u) == 0) {
        s_tot[tid >> 6] = tot;
        s_bad[tid >> 6] = bad;
    }
    __syncthreads();
    if (tid == 0) {
        const uint64_t t = s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3];
        const uint32_t bd = s_bad[0] | s_bad[1] | s_bad[2] | s_bad[3];
        const uint64_t n = md[mi].n;
        const bool on = n && !bd && t * 32 <= n;
        const bool full = !(dec[mi] & 1u) || lidx % EXPO_FULL == 0;
        dec[mi] = on ? (full ? 3u : 1u) : 0u;
    }
}

__device__ __forceinline__ uint32_t ens_block(uint32_t nblocks, uint32_t remap) {
    uint32_t b = blockIdx.x;
    if (remap) {
        // XCD-aware order: consecutive blocks of one member land on one XCD
        // (blocks are dispatched round-robin over the 8 XCDs)
        uint32_t per = (nblocks + 7) / 8;
        uint32_t x = b % 8, q = b / 8;
        b = x * per + q;
    }
    return b;   // >= nblocks: no block
}

// grid-stride over the logical blocks (of their member's `tile` agents):
// members whose pass is off, and blocks without an agent to mark (by the
// block's counts of the previous launch: infected, and infected in it), cost
// a read or two per logical block, not a block launch
__global__ __launch_bounds__(256) void k_ens_expose(const MemberDesc* __restrict__ md,
                                                    const uint32_t* __restrict__ bmember,
                                                    const uint32_t* __restrict__ bstart, uint32_t nblocks,
                                                    const uint32_t* __restrict__ dec, uint32_t parity,
                                                    const uint32_t* __restrict__ blkcnt) {
    for (uint32_t lb = blockIdx.x; lb < nblocks; lb += gridDim.x) {
        const uint32_t mi = bmember[lb];
        const uint32_t dm = dec[mi];
        if (!(dm & 1u)) continue;   // uniform per block
        const bool full = (dm & 2u) != 0u;
        // (on: every count of the member is present)
        const uint32_t bc = blkcnt[(uint64_t)parity * nblocks + lb];
        if ((full ? (bc & 0xFFFFu) : (bc >> 16)) == 0u) continue;
        const MemberDesc& m = md[mi];
        uint32_t* bits = m.bits + parity * m.nw;
        const uint64_t a0 = bstart[lb];
        for (uint32_t o = threadIdx.x; o < m.tile; o += 256) {
            const uint64_t i = a0 + o;
            if (i >= m.n) continue;
            const uint32_t st = m.rs[i];
            // infected, and (unless full) infected in the previous launch
            if ((st & 0xFFu) != VH_I || (!full && (st >> 16) != 0u)) continue;
            for (uint32_t k = m.out_off[i]; k < m.out_off[i + 1]; k++) {
                const uint32_t v = m.out_dst[k];
                atomicOr(bits + (v >> 5), 1u << (v & 31u));
            }
        }
    }
}

// one launch for every member: logical block b belongs to member
// bmember[b] and covers agents [bstart[b], bstart[b] + tile) of it, thread t
// the agents bstart[b] + t + k * blockDim, k < tile / blockDim (dec null: no
// exposure; else per-block infected counts of the written state go to
// blkcnt[(parity ^ 1) * nblocks + b])
__global__ void k_sir_ensemble(const MemberDesc* __restrict__ md, const uint32_t* __restrict__ bmember,
                               const uint32_t* __restrict__ bstart, uint32_t nblocks, uint32_t remap,
                               const uint32_t* __restrict__ dec, uint32_t parity, uint64_t lidx,
                               uint32_t* __restrict__ blkcnt) {
    __shared__ uint32_t s_inf[16], s_new[16];
    const uint32_t b = ens_block(nblocks, remap);
    if (b >= nblocks) return;   // whole block
    const uint32_t mi = bmember[b];
    const MemberDesc& m = md[mi];
    const uint64_t a0 = bstart[b];
    if (a0 >= m.n) return;      // whole block (a member past its horizon has n = 0)
    const uint32_t dm = dec ? dec[mi] : 0u;
    const bool ex = (dm & 1u) != 0u;
    const uint32_t apt = m.tile / blockDim.x;
    uint32_t c_inf = 0, c_new = 0;   // this wave's infected (and newly infected) agents
    for (uint32_t k = 0; k < apt; k++) {
        const uint64_t i = a0 + threadIdx.x + (uint64_t)k * blockDim.x;
        const bool in = i < m.n;
        uint32_t s = in ? m.rs[i] : 0u;
        const uint32_t h = s & 0xFFu;
        const bool skip = in && h == VH_S && ex && !((m.bits[parity * m.nw + (i >> 5)] >> (i & 31u)) & 1u);
        if (in && h == VH_S && !skip) {
            const uint64_t pre = vh_hash_pre(m.seed, vh_make_id(m.type, i + 1), m.step);
            const uint64_t pt = vh_u01_below(m.p);
            const uint32_t bb = m.off[i], e = m.off[i + 1];
            // the draw first, the source's state only below p (as in k_sir)
            if (vh_any_infecting(pre, pt, bb, e,
                                 [&](uint32_t q) { return (m.rs[m.src[q] - 1] & 0xFFu) == VH_I; }))
                s = (s & 0xFF00u) | VH_I;
        } else if (in && h == VH_I) {
            const uint32_t days = ((s >> 16) + 1u) & 0xFFFFu;
            s = (s & 0xFF00u) | (days << 16) | (days >= m.rec ? VH_R : VH_I);
        }
        if (in) m.ws[i] = s;
        if (dec) {
            c_inf += (uint32_t)__popcll(__ballot(in && (s & 0xFFu) == VH_I));
            c_new += (uint32_t)__popcll(__ballot(in && (s & 0xFFu) == VH_I && (s >> 16) == 0u));
        }
    }
    if (dec) {
        // this block's words of the other bitmap, for the next launch:
        // cleared before a full launch (this one off, or the next launch
        // number a multiple of EXPO_FULL), else this launch's words
        const uint64_t w0 = a0 >> 5;
        const bool next_full = !ex || (lidx + 1) % EXPO_FULL == 0;
        for (uint32_t q = threadIdx.x; q < (m.tile >> 5); q += blockDim.x)
            if (w0 + q < m.nw)
                m.bits[(parity ^ 1u) * m.nw + w0 + q] = next_full ? 0u : m.bits[parity * m.nw + w0 + q];
        // infected agents of the written state, and those infected in this
        // launch (days 0): count | new count << 16
        if ((threadIdx.x & 63u) == 0) {
            s_inf[threadIdx.x >> 6] = c_inf;
            s_new[threadIdx.x >> 6] = c_new;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t c = 0, cn = 0;
            for (uint32_t q = 0; q < (blockDim.x >> 6); q++) {
                c += s_inf[q];
                cn += s_new[q];
            }
            blkcnt[(uint64_t)(parity ^ 1u) * nblocks + b] = c | cn << 16;
        }
    }
}

}  // namespace

namespace {
__global__ void k_fill_weights(float* __restrict__ w, uint64_t n, uint64_t wseed) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) w[k] = (float)((double)((vh_hash(wseed, k, 0, VH_KEY_WEIGHT) >> 41) * 2 + 1) * 0x1.0p-24);
}
}  // namespace

namespace vh {
// Directed G(n, p), p = mean_deg / (n - 1), as a CSR in target order: the
// in-degree of target t is Binomial(n - 1, p) by inverse CDF, its sources are
// distinct uniform draws != t in draw order (vh_graph_gnp is the host twin).
// `cdf` and `sc` are working buffers; one host synchronisation (edge count).
int gnp_device(uint64_t n, double mean_deg, uint64_t seed, DevBuf& off, DevBuf& src, DevBuf& cdf_buf,
               Scratch& sc, hipStream_t st, uint64_t* nedges) {
    if (n == 0 || n >= 0xFFFFFFFFull) return set_error(VH_ERR_ARG, "G(n, p): bad n");
    double pe = n > 1 ? mean_deg / (double)(n - 1) : 0.0;
    std::vector<double> cdf = binom_cdf(n, pe);
    RC(cdf_buf.ensure(cdf.size() * 8, st));
    RC(off.ensure((n + 1) * 4, st));
    uint32_t* deg = (uint32_t*)sc.take(n * 4 + 8);
    if (!deg) return set_error(VH_ERR_OOM, "scratch");
    HIPCHECK(hipMemcpyAsync(cdf_buf.p, cdf.data(), cdf.size() * 8, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gnp_degree, dim3(grid_for(n, 256)), dim3(256), 0, st, cdf_buf.as<double>(),
                       (uint32_t)cdf.size(), seed, n, deg);
    if (scan_excl_u32(deg, off.as<uint32_t>(), n, sc, st)) return set_error(VH_ERR_DEVICE, "scan");
    uint32_t ne = 0;
    HIPCHECK(hipMemcpyAsync(&ne, off.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));   // also keeps `cdf` alive until the copy is done
    RC(src.ensure((uint64_t)ne * 4 + 8, st));
    hipLaunchKernelGGL(k_gnp_fill, dim3(grid_for(n, 256)), dim3(256), 0, st, n, seed, off.as<uint32_t>(),
                       src.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    *nedges = ne;
    return VH_OK;
}
// w[k] = (2 (h >> 41) + 1) 2^-24, h = vh_hash(wseed, k, 0, VH_KEY_WEIGHT): the
// values vh_fill_open01f(wseed, 0, VH_KEY_WEIGHT, n, w) gives on the host
void fill_weights(float* w, uint64_t n, uint64_t wseed, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_weights, dim3(grid_for(n, 256)), dim3(256), 0, st, w, n, wseed);
}
}  // namespace vh

// ====================================================================== host side
// completion of one vh_ensemble_step call, shared by the members it advanced
struct StepDone {
    hipEvent_t ev = nullptr;
    ~StepDone() {
        if (ev) (void)hipEventDestroy(ev);
    }
};

struct Member {
    uint32_t id;
    uint64_t n, nedges, step;
    uint64_t horizon = 0;   // last step to apply (0: unbounded)
    uint64_t seed;
    double p;
    uint32_t rec;
    DevBuf off, src, st[2];
    DevBuf out_off, out_dst, bits;   // out-edges and the two exposure bitmaps
    uint64_t nw = 0;                 // words per bitmap
    bool out_ready = false;          // built (on the first step with the exposure pass)
    int cur = 0;
    std::shared_ptr<StepDone> done;   // the last launch that wrote this member
    void release() {
        off.release(); src.release(); st[0].release(); st[1].release();
        out_off.release(); out_dst.release(); bits.release();
    }
};

struct vh_ensemble {
    std::map<uint32_t, Member*> members;
    uint32_t next_id = 1;
    hipStream_t stream = nullptr;   // the step launches
    hipStream_t aux = nullptr;      // member builds and reads, beside running steps
    bool own_streams = true;
    hipGraphExec_t exec = nullptr;  // last graph launched, destroyed after a sync
    std::shared_ptr<StepDone> last; // completion of the last step call
    // pinned staging of the block table and the descriptor tables: the
    // uploads stay asynchronous (reused after the next call's stream sync)
    void* h_table = nullptr;
    size_t h_table_cap = 0;
    void* h_descs = nullptr;
    size_t h_descs_cap = 0;
    uint32_t block = 256, remap = 1;
    uint32_t apt[ENS_CLASSES] = {1, 1, 1};   // agents per thread of each size class
    DevBuf d_desc, d_bmember, d_bstart;
    DevBuf d_mblk, d_blkcnt, d_dec;   // exposure: member block ranges, per-block counts, decisions
    Scratch oscr;                     // scratch of the members' out-edge builds (step stream)
    uint64_t launches = 0;            // step launches so far (bitmap parity)
    bool expose_last = false;         // the exposure pass ran in the last step call
    std::vector<MemberDesc> h_desc;
    std::vector<uint32_t> h_bmember, h_bstart;
    uint32_t nblocks = 0;
    bool table_valid = false;
    uint32_t table_block = 0;
    uint32_t table_apt[ENS_CLASSES] = {0, 0, 0};
    Scratch scratch;
    DevBuf d_cdf;
};

static int pinned(void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap) return VH_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t n = std::max(bytes, (size_t)4096);
    HIPCHECK(hipHostMalloc(p, n, hipHostMallocDefault));
    *cap = n;
    return VH_OK;
}

// size class of a member: [0, 10^4), [10^4, 10^5), [10^5, ...) agents
static int size_class(uint64_t n) { return n < 10000 ? 0 : n < 100000 ? 1 : 2; }
static uint32_t member_tile(const vh_ensemble* e, uint64_t n) { return e->block * e->apt[size_class(n)]; }
static bool table_current(const vh_ensemble* e) {
    if (!e->table_valid || e->table_block != e->block) return false;
    for (int c = 0; c < ENS_CLASSES; c++)
        if (e->table_apt[c] != e->apt[c]) return false;
    return true;
}

static int build_table(vh_ensemble* e) {
    e->h_bmember.clear();
    e->h_bstart.clear();
    e->h_desc.clear();
    uint32_t mi = 0;
    std::vector<uint32_t> mblk;
    for (auto& kv : e->members) {
        Member* m = kv.second;
        mblk.push_back((uint32_t)e->h_bmember.size());
        const uint32_t tile = member_tile(e, m->n);
        for (uint64_t a = 0; a < m->n; a += tile) {
            e->h_bmember.push_back(mi);
            e->h_bstart.push_back((uint32_t)a);
        }
        mblk.push_back((uint32_t)e->h_bmember.size() - mblk.back());
        mi++;
    }
    e->nblocks = (uint32_t)e->h_bmember.size();
    // exposure: member block ranges; block counts unknown until a launch
    // with this table has written them
    RC(e->d_mblk.ensure(mblk.size() * 4 + 8, e->stream));
    RC(e->d_blkcnt.ensure((size_t)e->nblocks * 8 + 8, e->stream));
    RC(e->d_dec.ensure(e->members.size() * 4 + 8, e->stream));
    if (!mblk.empty())
        HIPCHECK(hipMemcpy(e->d_mblk.p, mblk.data(), mblk.size() * 4, hipMemcpyHostToDevice));
    fill_u32(e->d_blkcnt.as<uint32_t>(), 0xFFFFFFFFu, (uint64_t)e->nblocks * 2, e->stream);
    RC(e->d_bmember.ensure(e->h_bmember.size() * 4 + 8, e->stream));
    RC(e->d_bstart.ensure(e->h_bstart.size() * 4 + 8, e->stream));
    RC(e->d_desc.ensure((e->members.size() + 1) * sizeof(MemberDesc), e->stream));
    if (e->nblocks) {
        // e->stream is idle here (vh_ensemble_step synchronised it)
        RC(pinned(&e->h_table, &e->h_table_cap, (size_t)e->nblocks * 8));
        uint32_t* ht = (uint32_t*)e->h_table;
        memcpy(ht, e->h_bmember.data(), e->nblocks * 4);
        memcpy(ht + e->nblocks, e->h_bstart.data(), e->nblocks * 4);
        HIPCHECK(hipMemcpyAsync(e->d_bmember.p, ht, e->nblocks * 4, hipMemcpyHostToDevice, e->stream));
        HIPCHECK(hipMemcpyAsync(e->d_bstart.p, ht + e->nblocks, e->nblocks * 4, hipMemcpyHostToDevice,
                                e->stream));
    }
    e->table_valid = true;
    e->table_block = e->block;
    for (int c = 0; c < ENS_CLASSES; c++) e->table_apt[c] = e->apt[c];
    return VH_OK;
}

// member descriptors for one step (read/write buffers and step numbers)
static int upload_desc(vh_ensemble* e) {
    e->h_desc.clear();
    for (auto& kv : e->members) {
        Member* m = kv.second;
        MemberDesc d;
        d.off = m->off.as<uint32_t>();
        d.src = m->src.as<uint32_t>();
        d.rs = m->st[m->cur].as<uint32_t>();
        d.ws = m->st[m->cur ^ 1].as<uint32_t>();
        d.n = m->n;
        d.step = m->step;
        d.seed = m->seed;
        d.p = m->p;
        d.rec = m->rec;
        d.type = 1;
        d.out_off = m->out_off.as<uint32_t>();
        d.out_dst = m->out_dst.as<uint32_t>();
        d.bits = m->bits.as<uint32_t>();
        d.nw = m->nw;
        d.tile = member_tile(e, m->n);
        d.pad = 0;
        e->h_desc.push_back(d);
    }
    return VH_OK;
}

extern "C" int vh_ensemble_create(vh_ensemble_t* out) {
    vh_ensemble* e = new vh_ensemble();
    hipError_t r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (r == hipSuccess) {
        r = hipStreamCreateWithFlags(&e->aux, hipStreamNonBlocking);
        if (r != hipSuccess) (void)hipStreamDestroy(e->stream);
    }
    if (r != hipSuccess) {
        delete e;
        return hip_fail(r, "hipStreamCreate");
    }
    *out = e;
    return VH_OK;
}

extern "C" int vh_ensemble_destroy(vh_ensemble_t e) {
    if (!e) return VH_OK;
    (void)hipStreamSynchronize(e->stream);
    (void)hipStreamSynchronize(e->aux);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    for (auto& kv : e->members) {
        Member* m = kv.second;
        m->release();
        delete m;
    }
    e->d_desc.release(); e->d_bmember.release(); e->d_bstart.release(); e->d_cdf.release();
    e->d_mblk.release(); e->d_blkcnt.release(); e->d_dec.release();
    e->scratch.release();
    e->oscr.release();
    if (e->h_table) (void)hipHostFree(e->h_table);
    if (e->h_descs) (void)hipHostFree(e->h_descs);
    if (e->own_streams) {
        (void)hipStreamDestroy(e->aux);
        (void)hipStreamDestroy(e->stream);
    }
    delete e;
    return VH_OK;
}

// the member's out-edges (the transpose of its in-edge CSR, by a stable radix
// sort on the source) and its two exposure bitmaps, cleared
static int member_out_edges(Member* m, uint64_t ne, Scratch& sc, hipStream_t st) {
    const uint64_t n = m->n;
    uint32_t* keys = (uint32_t*)sc.take(ne * 4 + 4);
    uint32_t* vals = (uint32_t*)sc.take(ne * 4 + 4);
    uint32_t* ka = (uint32_t*)sc.take(ne * 4 + 4);
    uint32_t* va = (uint32_t*)sc.take(ne * 4 + 4);
    if (!keys || !vals || !ka || !va) return set_error(VH_ERR_OOM, "scratch");
    RC(m->out_off.ensure((n + 1) * 4 + 8, st));
    RC(m->out_dst.ensure(ne * 4 + 8, st));
    m->nw = (n + 31) / 32;
    RC(m->bits.ensure(2 * m->nw * 4 + 8, st));
    fill_u32(m->bits.as<uint32_t>(), 0u, 2 * m->nw, st);
    if (ne) {
        hipLaunchKernelGGL(k_member_edge_rows, dim3(grid_for(n, 256)), dim3(256), 0, st, m->off.as<uint32_t>(),
                           m->src.as<uint32_t>(), n, keys, vals);
        uint32_t bits = 1;
        while (bits < 32 && (n >> bits) != 0) bits++;
        if (radix_sort_pairs(&keys, &vals, ka, va, ne, bits, false, sc, st))
            return set_error(VH_ERR_DEVICE, "radix sort failed");
        HIPCHECK(hipMemcpyAsync(m->out_dst.p, vals, ne * 4, hipMemcpyDeviceToDevice, st));
    }
    hipLaunchKernelGGL(k_member_offsets, dim3(grid_for(n + 1, 256)), dim3(256), 0, st, keys, ne, n,
                       m->out_off.as<uint32_t>());
    HIPCHECK(hipGetLastError());
    m->out_ready = true;
    return VH_OK;
}

// a SIR member on a device-generated G(n, p) (p = mean_deg / (n - 1)); the
// same edge list comes from vh_graph_gnp on the host
extern "C" int vh_ensemble_add_sir(vh_ensemble_t e, const vh_sir_member* sm, uint32_t* member_id) {
    if (!e || !sm || sm->n == 0 || sm->n >= 0xFFFFFFFFull) return set_error(VH_ERR_ARG, "bad member");
    Member* m = new Member();
    m->id = e->next_id++;
    m->n = sm->n;
    m->seed = sm->seed;
    m->p = sm->p_inf;
    m->rec = sm->recovery;
    m->step = 1;
    hipStream_t st = e->aux;   // builds run beside the step launches
    int rc = m->st[0].ensure(sm->n * 4, st);
    if (!rc) rc = m->st[1].ensure(sm->n * 4, st);
    uint64_t ne = 0;
    e->scratch.reset(st);
    if (!rc) rc = gnp_device(sm->n, sm->mean_deg, sm->graph_seed, m->off, m->src, e->d_cdf, e->scratch, st, &ne);
    if (rc) { m->release(); delete m; return rc; }
    m->nedges = ne;
    double frac = (double)sm->n_init / (double)sm->n;
    hipLaunchKernelGGL(k_sir_init, dim3(grid_for(sm->n, 256)), dim3(256), 0, st, m->st[0].as<uint32_t>(),
                       sm->n, sm->seed, frac, 1u);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    e->members[m->id] = m;
    e->table_valid = false;
    *member_id = m->id;
    return VH_OK;
}

extern "C" int vh_ensemble_remove(vh_ensemble_t e, uint32_t id) {
    auto it = e->members.find(id);
    if (it == e->members.end()) return set_error(VH_ERR_ARG, "unknown member %u", id);
    Member* m = it->second;
    // launches after the member's last step only read its (zero) size from
    // their descriptor tables, so its buffers are free once that step is done
    if (m->done) HIPCHECK(hipEventSynchronize(m->done->ev));
    m->release();
    delete m;
    e->members.erase(it);
    e->table_valid = false;
    return VH_OK;
}

extern "C" int vh_ensemble_set_horizon(vh_ensemble_t e, uint32_t id, uint64_t last_step) {
    auto it = e->members.find(id);
    if (it == e->members.end()) return set_error(VH_ERR_ARG, "unknown member %u", id);
    it->second->horizon = last_step;
    return VH_OK;
}

extern "C" int vh_ensemble_set_tiles(vh_ensemble_t e, const uint32_t* apt) {
    if (!e || !apt) return set_error(VH_ERR_ARG, "null argument");
    for (int c = 0; c < ENS_CLASSES; c++)
        if (apt[c] != 1 && apt[c] != 2 && apt[c] != 4)
            return set_error(VH_ERR_ARG, "agents per thread must be 1, 2 or 4");
    for (int c = 0; c < ENS_CLASSES; c++) e->apt[c] = apt[c];
    return VH_OK;
}

extern "C" int vh_ensemble_set_config(vh_ensemble_t e, uint32_t block, uint32_t remap) {
    if (block != 64 && block != 128 && block != 256 && block != 512 && block != 1024)
        return set_error(VH_ERR_ARG, "block must be 64..1024 (power of two)");
    e->block = block;
    e->remap = remap ? 1 : 0;
    e->table_valid = false;
    return VH_OK;
}

// advance every member `steps` steps; with use_graph the steps are captured
// once into a hipGraph and launched as one unit
extern "C" int vh_ensemble_step(vh_ensemble_t e, uint32_t steps, int use_graph) {
    if (e->members.empty() || !steps) return VH_OK;
    // the previous call's launches read d_desc and the staging buffers, and
    // its graph must outlive them: wait for them (only them: the stream may
    // be shared with other work)
    if (e->last) HIPCHECK(hipEventSynchronize(e->last->ev));
    if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    if (!table_current(e)) RC(build_table(e));
    size_t nm = e->members.size();
    // exposure pass unless the options turn it off, for ensembles of at
    // least sir_expose_min agents (default 65536: below, its two extra
    // launches per step cost more than the draws it saves); the members'
    // out-edges are built on the first such step, on the step stream
    uint64_t total = 0;
    for (auto& kv : e->members) total += kv.second->n;
    const bool expose = g_opts.sir_expose != 0 && (int64_t)total >= g_opts.sir_expose_min;
    if (expose) {
        for (auto& kv : e->members) {
            Member* m = kv.second;
            if (m->out_ready) continue;
            e->oscr.reset(e->stream);   // (waits for the previous build when it spilled)
            RC(member_out_edges(m, m->nedges, e->oscr, e->stream));
        }
        // the bitmaps were not kept while the pass was off: every member's
        // next exposure launch is a full one
        if (!e->expose_last) fill_u32(e->d_dec.as<uint32_t>(), 0u, nm, e->stream);
    }
    e->expose_last = expose;
    // one descriptor table per step (buffers alternate, step numbers grow)
    RC(e->d_desc.ensure((size_t)steps * nm * sizeof(MemberDesc), e->stream));
    auto done = std::make_shared<StepDone>();
    HIPCHECK(hipEventCreateWithFlags(&done->ev, hipEventDisableTiming));
    std::vector<MemberDesc> all;
    all.reserve((size_t)steps * nm);
    for (uint32_t s = 0; s < steps; s++) {
        upload_desc(e);
        size_t k = 0;
        for (auto& kv : e->members) {
            Member* m = kv.second;
            if (m->horizon && m->step > m->horizon) {
                e->h_desc[k].n = 0;          // finished: its blocks exit at once
            } else {
                m->cur ^= 1;
                m->step++;
                m->done = done;
            }
            k++;
        }
        all.insert(all.end(), e->h_desc.begin(), e->h_desc.end());
    }
    RC(pinned(&e->h_descs, &e->h_descs_cap, all.size() * sizeof(MemberDesc)));
    memcpy(e->h_descs, all.data(), all.size() * sizeof(MemberDesc));
    HIPCHECK(hipMemcpyAsync(e->d_desc.p, e->h_descs, all.size() * sizeof(MemberDesc), hipMemcpyHostToDevice,
                            e->stream));
    // with the XCD-aware order the grid is padded to a multiple of 8 so that
    // every (xcd, slot) pair, and with it every logical block, exists
    uint32_t grid = e->remap ? 8u * ((e->nblocks + 7u) / 8u) : e->nblocks;
    dim3 g(grid), b(e->block);
    const uint64_t launch0 = e->launches;
    auto launch = [&](uint32_t s) {
        const MemberDesc* md = e->d_desc.as<MemberDesc>() + (size_t)s * nm;
        const uint32_t parity = (uint32_t)((launch0 + s) & 1u);
        if (expose) {
            hipLaunchKernelGGL(k_ens_decide, dim3((uint32_t)nm), dim3(256), 0, e->stream, md,
                               e->d_mblk.as<uint32_t>(), e->d_blkcnt.as<uint32_t>(), e->nblocks, parity,
                               (uint64_t)(launch0 + s), e->d_dec.as<uint32_t>());
            hipLaunchKernelGGL(k_ens_expose, dim3(std::min<uint32_t>(e->nblocks, 8192u)), dim3(256), 0,
                               e->stream, md, e->d_bmember.as<uint32_t>(), e->d_bstart.as<uint32_t>(),
                               e->nblocks, (const uint32_t*)e->d_dec.as<uint32_t>(), parity,
                               (const uint32_t*)e->d_blkcnt.as<uint32_t>());
        }
        hipLaunchKernelGGL(k_sir_ensemble, g, b, 0, e->stream, md, e->d_bmember.as<uint32_t>(),
                           e->d_bstart.as<uint32_t>(), e->nblocks, e->remap,
                           expose ? (const uint32_t*)e->d_dec.as<uint32_t>() : nullptr, parity,
                           (uint64_t)(launch0 + s), e->d_blkcnt.as<uint32_t>());
    };
    e->launches += steps;
    if (use_graph && steps > 1) {
        hipGraph_t graph;
        HIPCHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        for (uint32_t s = 0; s < steps; s++) launch(s);
        HIPCHECK(hipStreamEndCapture(e->stream, &graph));
        hipError_t r = hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (r != hipSuccess) return hip_fail(r, "hipGraphInstantiate");
        HIPCHECK(hipGraphLaunch(e->exec, e->stream));   // kept until the next sync
    } else {
        for (uint32_t s = 0; s < steps; s++) launch(s);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipEventRecord(done->ev, e->stream));
    e->last = done;
    return VH_OK;
}

extern "C" int vh_ensemble_set_streams(vh_ensemble_t e, void* step_stream, void* aux_stream) {
    if (!e) return set_error(VH_ERR_ARG, "null ensemble");
    if ((step_stream == nullptr) != (aux_stream == nullptr))
        return set_error(VH_ERR_ARG, "set both streams or neither");
    HIPCHECK(hipStreamSynchronize(e->stream));
    HIPCHECK(hipStreamSynchronize(e->aux));
    if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    if (e->own_streams) {
        (void)hipStreamDestroy(e->aux);
        (void)hipStreamDestroy(e->stream);
    }
    if (step_stream) {
        e->stream = (hipStream_t)step_stream;
        e->aux = (hipStream_t)aux_stream;
        e->own_streams = false;
    } else {
        HIPCHECK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        HIPCHECK(hipStreamCreateWithFlags(&e->aux, hipStreamNonBlocking));
        e->own_streams = true;
    }
    return VH_OK;
}

extern "C" int vh_ensemble_sync(vh_ensemble_t e) {
    HIPCHECK(hipStreamSynchronize(e->stream));
    if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
    }
    return VH_OK;
}

extern "C" int vh_ensemble_read(vh_ensemble_t e, uint32_t id, void* states, uint64_t cap, uint64_t* n,
                                uint64_t* step) {
    auto it = e->members.find(id);
    if (it == e->members.end()) return set_error(VH_ERR_ARG, "unknown member %u", id);
    Member* m = it->second;
    *n = m->n;
    if (step) *step = m->step;
    if (cap < m->n) return set_error(VH_ERR_CAPACITY, "capacity");
    // only the launches that wrote this member must be complete: a member
    // past its horizon is read while later steps of the others run
    if (m->done) HIPCHECK(hipEventSynchronize(m->done->ev));
    HIPCHECK(hipMemcpyAsync(states, m->st[m->cur].p, m->n * 4, hipMemcpyDeviceToHost, e->aux));
    HIPCHECK(hipStreamSynchronize(e->aux));
    return VH_OK;
}

extern "C" int vh_ensemble_info(vh_ensemble_t e, uint64_t* members, uint64_t* agents, uint64_t* edges,
                                uint64_t* bytes) {
    uint64_t a = 0, ed = 0, by = 0;
    for (auto& kv : e->members) {
        Member* m = kv.second;
        a += m->n;
        ed += m->nedges;
        by += m->off.cap + m->src.cap + m->st[0].cap + m->st[1].cap;
    }
    *members = e->members.size();
    *agents = a;
    *edges = ed;
    *bytes = by;
    return VH_OK;
}

// autotuner over result-invariant launch knobs (block size, XCD-aware block
// order): times `trial_steps` steps per candidate on the members' current
// state and keeps the fastest.  Trial steps advance a snapshot, not the
// members: their buffers are restored afterwards.
// The autotuner: every trial runs `trial_steps` launches (after one launch
// that builds the block table) from a snapshot of the members' states and
// step numbers, restored afterwards. First the launch block size (128..1024)
// and block order (XCD-aware or not) with one agent per thread, then, per
// size class present, 2 or 4 agents per thread (larger logical blocks) where
// that is faster. Restoring invalidates the block table, so the first launch
// after it runs without the exposure pass and the next one rebuilds its
// bitmaps (no stale exposure bits).
extern "C" int vh_ensemble_autotune(vh_ensemble_t e, uint32_t trial_steps, uint32_t* best_block,
                                    uint32_t* best_remap, uint32_t* class_apt, float* best_ms) {
    if (e->members.empty()) return set_error(VH_ERR_ARG, "empty ensemble");
    if (trial_steps == 0) trial_steps = 3;
    // snapshot of the current state of every member
    std::vector<DevBuf> snap(e->members.size());
    size_t k = 0;
    std::vector<std::pair<int, uint64_t>> saved;
    bool present[ENS_CLASSES] = {false, false, false};
    for (auto& kv : e->members) {
        Member* m = kv.second;
        RC(snap[k].ensure(m->n * 4 + 4, e->stream));
        HIPCHECK(hipMemcpyAsync(snap[k].p, m->st[m->cur].p, m->n * 4, hipMemcpyDeviceToDevice, e->stream));
        saved.push_back({m->cur, m->step});
        present[size_class(m->n)] = true;
        k++;
    }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    auto trial = [&](uint32_t blk, uint32_t rm, const uint32_t* apt, float* ms) -> int {
        RC(vh_ensemble_set_config(e, blk, rm));
        RC(vh_ensemble_set_tiles(e, apt));
        RC(vh_ensemble_step(e, 1, 0));   // warm-up, table build
        HIPCHECK(hipEventRecord(a, e->stream));
        RC(vh_ensemble_step(e, trial_steps, 0));
        HIPCHECK(hipEventRecord(b, e->stream));
        HIPCHECK(hipEventSynchronize(b));
        HIPCHECK(hipEventElapsedTime(ms, a, b));
        *ms /= trial_steps;
        return VH_OK;
    };
    float best = 1e30f;
    uint32_t bb = e->block, br = e->remap;
    uint32_t bapt[ENS_CLASSES] = {1, 1, 1};
    const uint32_t blocks[] = {128, 256, 512, 1024};
    for (uint32_t blk : blocks)
        for (uint32_t rm = 0; rm < 2; rm++) {
            float ms = 0;
            RC(trial(blk, rm, bapt, &ms));
            if (ms < best) {
                best = ms;
                bb = blk;
                br = rm;
            }
        }
    for (int c = 0; c < ENS_CLASSES; c++) {
        if (!present[c]) continue;
        for (uint32_t v : {2u, 4u}) {
            uint32_t apt[ENS_CLASSES] = {bapt[0], bapt[1], bapt[2]};
            apt[c] = v;
            float ms = 0;
            RC(trial(bb, br, apt, &ms));
            if (ms < best) {
                best = ms;
                bapt[c] = v;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // restore
    k = 0;
    for (auto& kv : e->members) {
        Member* m = kv.second;
        m->cur = saved[k].first;
        m->step = saved[k].second;
        HIPCHECK(hipMemcpyAsync(m->st[m->cur].p, snap[k].p, m->n * 4, hipMemcpyDeviceToDevice, e->stream));
        k++;
    }
    HIPCHECK(hipStreamSynchronize(e->stream));
    for (auto& sb : snap) sb.release();
    RC(vh_ensemble_set_config(e, bb, br));   // (invalidates the block table)
    RC(vh_ensemble_set_tiles(e, bapt));
    if (best_block) *best_block = bb;
    if (best_remap) *best_remap = br;
    if (class_apt)
        for (int c = 0; c < ENS_CLASSES; c++) class_apt[c] = bapt[c];
    if (best_ms) *best_ms = best;
    return VH_OK;
}

// host restatement of the device generator (same table, same hash draws)
extern "C" int vh_graph_gnp(uint64_t n, double mean_deg, uint64_t seed, uint32_t* from, uint32_t* to,
                            uint64_t cap, uint64_t* m) {
    if (n >= 0xFFFFFFFFull) return VH_ERR_ARG;
    double pe = n > 1 ? mean_deg / (double)(n - 1) : 0.0;
    std::vector<double> cdf = binom_cdf(n, pe);
    uint64_t k = 0;
    std::vector<uint32_t> buf;
    for (uint64_t t = 1; t <= n; t++) {
        uint32_t d = gnp_degree(cdf.data(), (uint32_t)cdf.size(), seed, t);
        buf.resize(d);
        gnp_sources(n, seed, t, d, buf.data());
        for (uint32_t j = 0; j < d; j++, k++)
            if (k < cap) {
                from[k] = buf[j];
                to[k] = (uint32_t)t;
            }
    }
    *m = k;
    return k > cap ? VH_ERR_CAPACITY : VH_OK;
}
