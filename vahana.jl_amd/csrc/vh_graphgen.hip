// vh_graphgen.hip — graph construction on the device for finish_init!:
// Barabasi-Albert preferential attachment (add_graph! of a generated graph,
// src/GraphsSupport.jl:34-57) and the breadth-first agent numbering that
// vh_finish_init_relabel applies. Both give exactly the results of the host
// versions in vh_graph.cpp (vh_graph_ba, vh_bfs_order), which the checker's
// inputs use.
#include <string.h>

#include <vector>

#include "vh_engine.h"

using namespace vh;

namespace {

constexpr uint64_t KEY_BA = 10ull << 40;   // vh_graph.cpp's stream key of the BA draws
constexpr uint32_t NONE = 0xFFFFFFFFu;

// ------------------------------------------------------------- Barabasi-Albert
// vh_graph_ba builds the pairs in creation order: the clique on 1..m (pair k:
// a = j, b = i for i < j), then node v = m+1..n with pairs K_v + q = (v,
// chosen_q), K_v = C + (v - m - 1) m, C = m (m - 1) / 2. Its endpoint list
// holds, for pair k, b then a within the clique and a then b after it, and
// node v draws u = ends[((h >> 32) L_v) >> 32], L_v = 2 K_v, h =
// hash(seed, v, draw, KEY_BA), until it has m distinct u. An endpoint at an
// even index past the clique is known (the node of its pair); at an odd index
// it is another node's choice. The device resolves every node's draws in
// rounds: a node stops at a choice not made yet and continues in the next
// round from the same draw. The earliest unfinished node only refers to
// finished nodes, so every round finishes at least that node.
__device__ __forceinline__ void clique_pair(uint64_t k, uint32_t m, uint32_t& i, uint32_t& j) {
    uint32_t r = 1;
    uint64_t len = m - 1;
    while (k >= len) {
        k -= len;
        r++;
        len--;
    }
    i = r;
    j = r + 1 + (uint32_t)k;
}

__global__ void k_ba_init(uint64_t npairs, uint32_t m, uint64_t C, uint32_t* __restrict__ a,
                          uint32_t* __restrict__ b) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    if (k < C) {
        uint32_t i, j;
        clique_pair(k, m, i, j);
        a[k] = j;
        b[k] = i;
    } else {
        a[k] = (uint32_t)(m + 1 + (k - C) / m);
        b[k] = 0u;
    }
}

__global__ void k_ba_round(uint64_t n, uint32_t m, uint64_t C, uint64_t seed, uint32_t* __restrict__ b,
                           uint32_t* __restrict__ draws, uint8_t* __restrict__ got, uint32_t* __restrict__ open,
                           uint32_t budget) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;   // node m + 1 + x
    if (x >= n - m) return;
    uint32_t g = got[x];
    if (g >= m) return;
    const uint64_t v = m + 1 + x;
    const uint64_t base = C + x * m;
    const uint64_t L = 2 * base;
    uint32_t d = draws[x];
    for (uint32_t it = 0; it < budget && g < m; it++) {
        uint32_t u;
        if (L == 0) {
            u = g + 1;   // m == 1: the second node attaches to the first
        } else {
            const uint64_t h = vh_hash(seed, v, d, KEY_BA);
            const uint64_t idx = ((h >> 32) * L) >> 32;
            const uint64_t k = idx >> 1;
            if (k < C) {
                uint32_t i, j;
                clique_pair(k, m, i, j);
                u = (idx & 1) ? j : i;
            } else if (!(idx & 1)) {
                u = (uint32_t)(m + 1 + (k - C) / m);
            } else {
                u = __hip_atomic_load(b + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (u == 0) break;   // that node has not made this choice yet
            }
            d++;
        }
        bool dup = false;
        for (uint32_t q = 0; q < g; q++) dup |= b[base + q] == u;
        if (!dup) {
            __hip_atomic_store(b + base + g, u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            g++;
        }
    }
    draws[x] = d;
    got[x] = (uint8_t)(g < 255 ? g : 255);
    if (g < m) atomicAdd(open, 1u);
}

// ------------------------------------------------------------- BFS numbering
// vh_bfs_order: adjacency of v = the other endpoint of every edge incident to
// v, in edge order (an edge (a, b) adds b to a's list, then a to b's); the
// first root is the vertex of highest degree (the lowest among equals), each
// later one the lowest unvisited vertex; a vertex gets the next position when
// it is first reached. Level by level: the frontier's adjacency entries are
// enumerated in the sequential visit order (frontier position, then list
// position), every unvisited neighbour keeps its smallest enumeration index
// (atomicMin), and the winners, compacted in enumeration order, form the next
// level in exactly the sequential order.
__global__ void k_inc_keys(const uint32_t* __restrict__ ea, const uint32_t* __restrict__ eb, uint64_t M,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ deg) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t x = ea[j], y = eb[j];
    keys[2 * j] = x;
    keys[2 * j + 1] = y;
    atomicAdd(deg + x, 1u);
    atomicAdd(deg + y, 1u);
}

// nbr[p] = the other endpoint of incidence vals[p] (even: the edge's b side)
__global__ void k_inc_nbr(const uint32_t* __restrict__ vals, uint64_t n2, const uint32_t* __restrict__ ea,
                          const uint32_t* __restrict__ eb, uint32_t* __restrict__ nbr) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n2) return;
    const uint32_t q = vals[p];
    nbr[p] = (q & 1) ? ea[q >> 1] : eb[q >> 1];
}

__global__ void k_bfs_root(const uint32_t* __restrict__ deg, uint64_t n, unsigned long long* best) {
    const uint64_t v = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    atomicMax(best, ((unsigned long long)deg[v] << 32) | (unsigned long long)(NONE - (uint32_t)v));
}

__global__ void k_bfs_place(uint32_t* __restrict__ pos, uint32_t* __restrict__ order, uint32_t v, uint32_t p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        pos[v] = p;
        order[p] = v;
    }
}

// fdeg[p] = degree of frontier vertex p, fdeg[F] = 0
__global__ void k_front_deg(const uint32_t* __restrict__ order, uint64_t lo, uint64_t F,
                            const uint32_t* __restrict__ deg, uint32_t* __restrict__ fdeg) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < F) fdeg[p] = deg[order[lo + p]];
    else if (p == F) fdeg[p] = 0u;
}

__device__ __forceinline__ uint32_t front_nbr(uint64_t e, const uint32_t* __restrict__ foff, uint64_t F,
                                              const uint32_t* __restrict__ order, uint64_t lo,
                                              const uint32_t* __restrict__ start,
                                              const uint32_t* __restrict__ nbr) {
    // the last frontier vertex whose entries start at or before e
    uint64_t l = 0, h = F;   // foff[l] <= e < foff[h]
    while (h - l > 1) {
        const uint64_t mid = (l + h) >> 1;
        if (foff[mid] <= e) l = mid;
        else h = mid;
    }
    const uint32_t u = order[lo + l];
    return nbr[start[u] + (uint32_t)(e - foff[l])];
}

__global__ void k_front_min(uint64_t T, const uint32_t* __restrict__ foff, uint64_t F,
                            const uint32_t* __restrict__ order, uint64_t lo, const uint32_t* __restrict__ start,
                            const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ pos,
                            uint32_t* __restrict__ keymin) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T) return;
    const uint32_t w = front_nbr(e, foff, F, order, lo, start, nbr);
    if (pos[w] == NONE) atomicMin(keymin + w, (uint32_t)e);
}

__global__ void k_front_flag(uint64_t T, const uint32_t* __restrict__ foff, uint64_t F,
                             const uint32_t* __restrict__ order, uint64_t lo, const uint32_t* __restrict__ start,
                             const uint32_t* __restrict__ nbr, const uint32_t* __restrict__ pos,
                             const uint32_t* __restrict__ keymin, uint32_t* __restrict__ flag) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e > T) return;
    if (e == T) {
        flag[e] = 0u;
        return;
    }
    const uint32_t w = front_nbr(e, foff, F, order, lo, start, nbr);
    flag[e] = (pos[w] == NONE && keymin[w] == (uint32_t)e) ? 1u : 0u;
}

__global__ void k_front_assign(uint64_t T, const uint32_t* __restrict__ foff, uint64_t F,
                               uint32_t* __restrict__ order, uint64_t lo, uint64_t hi,
                               const uint32_t* __restrict__ start, const uint32_t* __restrict__ nbr,
                               uint32_t* __restrict__ pos, uint32_t* __restrict__ keymin,
                               const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fpos) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= T || !flag[e]) return;
    const uint32_t w = front_nbr(e, foff, F, order, lo, start, nbr);
    const uint32_t p = (uint32_t)(hi + fpos[e]);
    order[p] = w;
    pos[w] = p;
    keymin[w] = NONE;
}

// lowest unvisited vertex >= from with at least one incident edge
__global__ void k_next_root(const uint32_t* __restrict__ pos, const uint32_t* __restrict__ deg, uint64_t from,
                            uint64_t n, uint32_t* __restrict__ out) {
    const uint64_t v = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > n) return;
    if (pos[v] == NONE && deg[v] > 0) atomicMin(out, (uint32_t)v);
}

// unvisited vertices of [from, to) (all without edges): flag, then positions
__global__ void k_single_flag(const uint32_t* __restrict__ pos, uint64_t from, uint64_t cnt,
                              uint32_t* __restrict__ flag) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < cnt) flag[x] = pos[from + x] == NONE ? 1u : 0u;
    else if (x == cnt) flag[x] = 0u;
}
__global__ void k_single_assign(uint32_t* __restrict__ pos, uint32_t* __restrict__ order, uint64_t from,
                                uint64_t cnt, uint64_t hi, const uint32_t* __restrict__ flag,
                                const uint32_t* __restrict__ fpos) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= cnt || !flag[x]) return;
    const uint32_t p = (uint32_t)(hi + fpos[x]);
    pos[from + x] = p;
    order[p] = (uint32_t)(from + x);
}

__global__ void k_newnr(const uint32_t* __restrict__ pos, uint64_t n, uint32_t* __restrict__ newnr) {
    const uint64_t v = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v <= n) newnr[v - 1] = pos[v] + 1;
}

int scan_or_fail(const uint32_t* in, uint32_t* out, uint64_t n, Scratch& sc, hipStream_t st) {
    const int r = scan_excl_u32(in, out, n, sc, st);
    if (r == -3) return set_error(VH_ERR_OOM, "scratch (scan)");
    if (r) return set_error(VH_ERR_DEVICE, "scan launch failed");
    return VH_OK;
}

template <class T>
int read1(const T* d, T* h, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(h, d, sizeof(T), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    return VH_OK;
}

}  // namespace

namespace vh {

size_t bfs_scratch_bytes(uint64_t n, uint64_t M) {
    // per-agent arrays, the sorted incidences, the level buffers, and room
    // for the scans and the sort's histograms
    return 24 * (n + 2) + 64 * (M + 2) + ((size_t)64 << 20);
}

uint64_t ba_pairs(uint64_t n, uint32_t m) {
    if (m < 1 || n < m) return 0;
    return (uint64_t)m * (m - 1) / 2 + (n - m) * (uint64_t)m;
}

int ba_device(uint64_t n, uint32_t m, uint64_t seed, uint32_t* a, uint32_t* b, Scratch& sc, hipStream_t st) {
    if (m < 1 || m > 255 || n < m || n >= 0xFFFFFFFFull)
        return set_error(VH_ERR_ARG, "Barabasi-Albert: need 1 <= m <= min(n, 255), n < 2^32");
    const uint64_t K = ba_pairs(n, m), C = (uint64_t)m * (m - 1) / 2;
    if (2 * K >= 0xFFFFFFFFull) return set_error(VH_ERR_CAPACITY, "Barabasi-Albert: 2 m n must stay below 2^32");
    if (K) hipLaunchKernelGGL(k_ba_init, dim3(grid_for(K, 256)), dim3(256), 0, st, K, m, C, a, b);
    const uint64_t nodes = n - m;
    if (!nodes) return VH_OK;
    uint32_t* draws = (uint32_t*)sc.take(nodes * 4);
    uint8_t* got = (uint8_t*)sc.take(nodes);
    uint32_t* open = (uint32_t*)sc.take(4);
    if (!draws || !got || !open) return set_error(VH_ERR_OOM, "scratch");
    fill_u32(draws, 0u, nodes, st);
    fill_u8(got, 0, nodes, st);
    for (int round = 0;; round++) {
        // every round finishes at least the earliest open node; about
        // 2 log2(n) rounds are typical
        if (round > 4096) return set_error(VH_ERR_DEVICE, "Barabasi-Albert: too many rounds");
        fill_u32(open, 0u, 1, st);
        hipLaunchKernelGGL(k_ba_round, dim3(grid_for(nodes, 256)), dim3(256), 0, st, n, m, C, seed, b, draws,
                           got, open, 64u);
        uint32_t left = 0;
        RC(read1(open, &left, st));
        if (!left) break;
    }
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

int bfs_device(uint64_t n, uint64_t M, const uint32_t* ea, const uint32_t* eb, uint32_t* newnr, Scratch& sc,
               hipStream_t st) {
    if (n >= 0xFFFFFFFFull || 2 * M >= 0xFFFFFFFFull)
        return set_error(VH_ERR_CAPACITY, "breadth-first numbering: n < 2^32 and 2 m < 2^32");
    if (!n) return VH_OK;
    const uint64_t n2 = 2 * M;
    uint32_t* deg = (uint32_t*)sc.take((n + 2) * 4);
    uint32_t* start = (uint32_t*)sc.take((n + 2) * 4);
    uint32_t* pos = (uint32_t*)sc.take((n + 1) * 4);
    uint32_t* order = (uint32_t*)sc.take(n * 4);
    uint32_t* keymin = (uint32_t*)sc.take((n + 1) * 4);
    uint32_t* keys = (uint32_t*)sc.take(n2 * 4 + 4);
    uint32_t* vals = (uint32_t*)sc.take(n2 * 4 + 4);
    uint32_t* kalt = (uint32_t*)sc.take(n2 * 4 + 4);
    uint32_t* valt = (uint32_t*)sc.take(n2 * 4 + 4);
    // frontier buffers: at most n + 1 frontier entries, at most 2 M + 1
    // adjacency entries per level
    const uint64_t fe = (n2 > n ? n2 : n) + 2;
    uint32_t* fdeg = (uint32_t*)sc.take(fe * 4);
    uint32_t* foff = (uint32_t*)sc.take(fe * 4);
    uint32_t* flag = (uint32_t*)sc.take(fe * 4);
    uint32_t* fpos = (uint32_t*)sc.take(fe * 4);
    unsigned long long* best = (unsigned long long*)sc.take(8);
    uint32_t* small = (uint32_t*)sc.take(8);
    if (!deg || !start || !pos || !order || !keymin || !keys || !vals || !kalt || !valt || !fdeg || !foff ||
        !flag || !fpos || !best || !small)
        return set_error(VH_ERR_OOM, "scratch");
    fill_u32(deg, 0u, n + 2, st);
    fill_u32(pos, NONE, n + 1, st);
    fill_u32(keymin, NONE, n + 1, st);
    uint32_t* nbr = kalt;   // free after the sort
    if (M) {
        hipLaunchKernelGGL(k_inc_keys, dim3(grid_for(M, 256)), dim3(256), 0, st, ea, eb, M, keys, deg);
        uint32_t bits = 1;
        while (bits < 32 && (n >> bits)) bits++;
        if (radix_sort_pairs(&keys, &vals, kalt, valt, n2, bits, true, sc, st))
            return set_error(VH_ERR_DEVICE, "radix sort failed");
        nbr = keys;   // the sorted keys are not needed past this point
        hipLaunchKernelGGL(k_inc_nbr, dim3(grid_for(n2, 256)), dim3(256), 0, st, vals, n2, ea, eb, nbr);
    }
    RC(scan_or_fail(deg, start, n + 2, sc, st));
    // the first root: highest degree, lowest number among equals
    fill_u32((uint32_t*)best, 0u, 2, st);
    hipLaunchKernelGGL(k_bfs_root, dim3(grid_for(n, 256)), dim3(256), 0, st, deg, n, best);
    unsigned long long bh = 0;
    RC(read1(best, &bh, st));
    uint32_t root = NONE - (uint32_t)(bh & 0xFFFFFFFFull);
    uint64_t hi = 0, cursor = 1;
    bool first = true;
    while (hi < n) {
        if (!first) {
            // the lowest unvisited vertex with edges; the unvisited ones
            // below it have none and are their own components, in order
            if (cursor > n) break;
            fill_u32(small, NONE, 1, st);
            hipLaunchKernelGGL(k_next_root, dim3(grid_for(n - cursor + 1, 256)), dim3(256), 0, st, pos, deg,
                               cursor, n, small);
            uint32_t r2 = NONE;
            RC(read1(small, &r2, st));
            const uint64_t lim = r2 == NONE ? n + 1 : r2;
            if (lim > cursor) {
                const uint64_t cnt = lim - cursor;
                hipLaunchKernelGGL(k_single_flag, dim3(grid_for(cnt + 1, 256)), dim3(256), 0, st, pos, cursor,
                                   cnt, flag);
                RC(scan_or_fail(flag, fpos, cnt + 1, sc, st));
                hipLaunchKernelGGL(k_single_assign, dim3(grid_for(cnt, 256)), dim3(256), 0, st, pos, order, cursor,
                                   cnt, hi, flag, fpos);
                uint32_t c = 0;
                RC(read1(fpos + cnt, &c, st));
                hi += c;
            }
            if (r2 == NONE) break;
            root = r2;
            cursor = (uint64_t)r2 + 1;
        }
        first = false;
        hipLaunchKernelGGL(k_bfs_place, dim3(1), dim3(64), 0, st, pos, order, root, (uint32_t)hi);
        uint64_t lo = hi;
        hi += 1;
        while (lo < hi) {
            const uint64_t F = hi - lo;
            hipLaunchKernelGGL(k_front_deg, dim3(grid_for(F + 1, 256)), dim3(256), 0, st, order, lo, F, deg, fdeg);
            RC(scan_or_fail(fdeg, foff, F + 1, sc, st));
            uint32_t T32 = 0;
            RC(read1(foff + F, &T32, st));
            const uint64_t T = T32;
            lo = hi;
            if (!T) break;
            const uint64_t flo = hi - F;
            hipLaunchKernelGGL(k_front_min, dim3(grid_for(T, 256)), dim3(256), 0, st, T, foff, F, order, flo,
                               start, nbr, pos, keymin);
            hipLaunchKernelGGL(k_front_flag, dim3(grid_for(T + 1, 256)), dim3(256), 0, st, T, foff, F, order,
                               flo, start, nbr, pos, keymin, flag);
            RC(scan_or_fail(flag, fpos, T + 1, sc, st));
            hipLaunchKernelGGL(k_front_assign, dim3(grid_for(T, 256)), dim3(256), 0, st, T, foff, F, order, flo,
                               hi, start, nbr, pos, keymin, flag, fpos);
            uint32_t c = 0;
            RC(read1(fpos + T, &c, st));
            hi += c;
        }
    }
    hipLaunchKernelGGL(k_newnr, dim3(grid_for(n, 256)), dim3(256), 0, st, pos, n, newnr);
    HIPCHECK(hipGetLastError());
    return VH_OK;
}

}  // namespace vh
