// vh_graph.cpp — host-side graph sources for add_graph! (src/GraphsSupport.jl:34-57)
// and input fills driven by the counter hash, so every consumer of a seed sees
// the same edge list and initial state.
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/vahana_hip.h"

namespace {
constexpr uint64_t G = 0x9E3779B97F4A7C15ull;
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t hash4(uint64_t s, uint64_t a, uint64_t b, uint64_t c) {
    uint64_t h = mix64(s + G);
    h = mix64(h + a + G);
    h = mix64(h + b + G);
    return mix64(h + c + G);
}
inline double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }
constexpr uint64_t KEY_ER = 9ull << 40;
constexpr uint64_t KEY_BA = 10ull << 40;
}  // namespace

extern "C" uint64_t vh_hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return hash4(seed, a, b, c);
}

// out[i] = odd multiple of 2^-24 in (0, 1) from hash(seed, i, b, c): the
// Float32 edge weights of the opinion model
extern "C" int vh_fill_open01f(uint64_t seed, uint64_t b, uint64_t c, uint64_t n, float* out) {
    for (uint64_t i = 0; i < n; i++) {
        uint64_t h = hash4(seed, i, b, c);
        out[i] = (float)((double)((h >> 41) * 2 + 1) * 0x1.0p-24);
    }
    return VH_OK;
}

extern "C" int vh_fill_agent_u01(uint64_t seed, uint32_t type, uint64_t b, uint64_t c, uint64_t n,
                                 double* out) {
    const uint64_t tid = (uint64_t)type << 56;
    for (uint64_t i = 0; i < n; i++) out[i] = (double)(hash4(seed, tid | (i + 1), b, c) >> 11) * 0x1.0p-53;
    return VH_OK;
}

// Breadth-first numbering of n vertices over m undirected edges (a_k, b_k)
// (1-based): adjacency lists in edge order, the first root the vertex of
// highest degree (the lowest number among equals), each further component
// from its lowest-numbered vertex. new_nr[v - 1] = position of v + 1.
extern "C" int vh_bfs_order(uint64_t n, uint64_t m, const uint32_t* a, const uint32_t* b,
                            uint64_t* new_nr) {
    if (n >= 0xFFFFFFFFull) return VH_ERR_CAPACITY;
    for (uint64_t k = 0; k < m; k++)
        if (a[k] < 1 || a[k] > n || b[k] < 1 || b[k] > n) return VH_ERR_ARG;
    std::vector<uint64_t> start(n + 2, 0);
    for (uint64_t k = 0; k < m; k++) {
        start[a[k]]++;
        start[b[k]]++;
    }
    uint64_t acc = 0;
    for (uint64_t v = 1; v <= n + 1; v++) {
        uint64_t c = start[v];
        start[v] = acc;
        acc += c;
    }
    std::vector<uint32_t> adj(acc);
    {
        std::vector<uint64_t> fill(start.begin(), start.end());
        for (uint64_t k = 0; k < m; k++) {
            adj[fill[a[k]]++] = b[k];
            adj[fill[b[k]]++] = a[k];
        }
    }
    std::vector<uint32_t> order;
    order.reserve(n);
    std::vector<uint8_t> seen(n + 1, 0);
    uint64_t root = 1, best = 0;
    for (uint64_t v = 1; v <= n; v++)
        if (start[v + 1] - start[v] > best) {
            best = start[v + 1] - start[v];
            root = v;
        }
    uint64_t next_unvisited = 1;
    while (order.size() < n) {
        if (seen[root]) {
            while (next_unvisited <= n && seen[next_unvisited]) next_unvisited++;
            root = next_unvisited;
        }
        size_t head = order.size();
        seen[root] = 1;
        order.push_back((uint32_t)root);
        while (head < order.size()) {
            uint32_t v = order[head++];
            for (uint64_t k = start[v]; k < start[v + 1]; k++) {
                uint32_t w = adj[k];
                if (!seen[w]) {
                    seen[w] = 1;
                    order.push_back(w);
                }
            }
        }
    }
    for (uint64_t k = 0; k < n; k++) new_nr[order[k] - 1] = k + 1;
    return VH_OK;
}

// Directed G(n, p) by geometric skipping over the n-1 candidate sources of
// each target (Batagelj & Brandes 2005), targets ascending.
extern "C" int vh_graph_er(uint64_t n, double p, uint64_t seed, uint32_t* from, uint32_t* to,
                           uint64_t cap, uint64_t* m) {
    uint64_t k = 0;
    if (n >= 0xFFFFFFFFull) return VH_ERR_ARG;
    double lq = p < 1.0 ? log1p(-p) : 0.0;
    for (uint64_t t = 1; t <= n && p > 0.0; t++) {
        int64_t x = -1;
        uint64_t draw = 0;
        for (;;) {
            uint64_t skip = 0;
            if (p < 1.0) {
                double u = u01(hash4(seed, t, draw++, KEY_ER));
                double q = floor(log1p(-u) / lq);
                skip = q > 1e18 ? (uint64_t)1e18 : (uint64_t)q;
            }
            x += 1 + (int64_t)skip;
            if (x >= (int64_t)n - 1 || x < 0) break;
            uint64_t s = (uint64_t)x + 1;          // 1-based among candidates
            uint64_t src = s < t ? s : s + 1;      // skip the target itself
            if (k < cap) {
                from[k] = (uint32_t)src;
                to[k] = (uint32_t)t;
            }
            k++;
        }
    }
    *m = k;
    return k > cap ? VH_ERR_CAPACITY : VH_OK;
}

// Barabasi-Albert preferential attachment: a clique on nodes 1..m, then each
// node v = m+1..n links to m distinct earlier nodes drawn from the endpoint
// list (probability proportional to degree).  Undirected pairs (v, u) are
// returned in creation order.
extern "C" int vh_graph_ba(uint64_t n, uint32_t m, uint64_t seed, uint32_t* a, uint32_t* b,
                           uint64_t cap, uint64_t* count) {
    if (m < 1 || n < m || n >= 0xFFFFFFFFull) return VH_ERR_ARG;
    uint64_t need = (uint64_t)m * (m - 1) / 2 + (n - m) * (uint64_t)m;
    *count = need;
    if (cap < need) return VH_ERR_CAPACITY;
    std::vector<uint32_t> ends;
    ends.reserve(2 * need);
    uint64_t k = 0;
    for (uint32_t i = 1; i <= m; i++)
        for (uint32_t j = i + 1; j <= m; j++) {
            a[k] = j;
            b[k] = i;
            k++;
            ends.push_back(i);
            ends.push_back(j);
        }
    std::vector<uint32_t> chosen(m);
    for (uint64_t v = (uint64_t)m + 1; v <= n; v++) {
        uint32_t got = 0;
        uint64_t draw = 0;
        uint64_t L = ends.size();
        while (got < m) {
            uint32_t u;
            if (L == 0) {
                u = (uint32_t)(got + 1);   // m == 1 start: attach to node 1
            } else {
                uint64_t h = hash4(seed, v, draw++, KEY_BA);
                u = ends[((h >> 32) * L) >> 32];
            }
            bool dup = false;
            for (uint32_t q = 0; q < got; q++) dup |= chosen[q] == u;
            if (dup) continue;
            chosen[got++] = u;
        }
        for (uint32_t q = 0; q < m; q++) {
            a[k] = (uint32_t)v;
            b[k] = chosen[q];
            k++;
            ends.push_back((uint32_t)v);
            ends.push_back(chosen[q]);
        }
    }
    return VH_OK;
}
