// vh_prims.hip — device-wide scan, stable LSD radix sort, fills and gathers
// for gfx950 (wave64).
#include <atomic>

#include "vh_prims.h"

#include <algorithm>

namespace vh {

// ------------------------------------------------------------------ arena
// every device allocation of the library; a test hook makes the next n fail
static std::atomic<uint32_t> g_fail_allocs{0};
hipError_t device_alloc(void** p, size_t bytes) {
    uint32_t k = g_fail_allocs.load();
    while (k > 0)
        if (g_fail_allocs.compare_exchange_weak(k, k - 1)) return hipErrorOutOfMemory;
    return hipMalloc(p, bytes);
}
void fail_next_allocs(uint32_t n) { g_fail_allocs.store(n); }

void* Scratch::take(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (used + bytes <= cap) {
        void* r = (char*)main + used;
        used += bytes;
        return r;
    }
    if (xb && xused + bytes <= xcap) {
        void* r = xb + xused;
        xused += bytes;
        return r;
    }
    if (nextra >= 64) return nullptr;
    void* p = nullptr;
    if (device_alloc(&p, bytes) != hipSuccess) return nullptr;
    extra[nextra++] = p;
    spilled += bytes;
    return p;
}
int Scratch::room(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (used + bytes <= cap || (xb && xused + bytes <= xcap)) return 0;
    if (nextra >= 64) return -3;
    void* p = nullptr;
    if (device_alloc(&p, bytes) != hipSuccess) return -3;
    extra[nextra++] = p;
    spilled += bytes;
    xb = (char*)p;
    xcap = bytes;
    xused = 0;
    return 0;
}
void Scratch::reset(hipStream_t st) {
    xb = nullptr;
    xcap = xused = 0;
    if (shrink) {
        shrink = false;
        if (bytes_held() > ((size_t)1 << 30)) {
            hipStreamSynchronize(st);
            release();
            return;
        }
    }
    if (nextra) {
        // fold the spill into one main block sized for the whole round
        size_t need = used + spilled;
        hipStreamSynchronize(st);
        for (int i = 0; i < nextra; i++) hipFree(extra[i]);
        nextra = 0;
        spilled = 0;
        if (main) hipFree(main);
        main = nullptr;
        cap = 0;
        size_t ncap = need + need / 4 + (1 << 20);
        if (device_alloc(&main, ncap) == hipSuccess) cap = ncap;
    }
    used = 0;
}
int Scratch::reserve(size_t bytes, hipStream_t st) {
    if (used || nextra || cap >= bytes) return 0;
    hipStreamSynchronize(st);
    if (main) hipFree(main);
    main = nullptr;
    cap = 0;
    if (device_alloc(&main, bytes) != hipSuccess) return -3;
    cap = bytes;
    return 0;
}
void Scratch::release() {
    xb = nullptr;
    xcap = xused = 0;
    for (int i = 0; i < nextra; i++) hipFree(extra[i]);
    nextra = 0;
    spilled = 0;
    if (main) hipFree(main);
    main = nullptr;
    cap = used = 0;
}

// ------------------------------------------------------------------ scan
constexpr int SB = 256;           // threads per block
constexpr int SI = 8;             // items per thread
constexpr int STILE = SB * SI;    // 2048 items per block

template <class T>
__device__ __forceinline__ T wave_incl_scan(T x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive block scan of one value per thread; *total gets the block sum
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T x = wave_incl_scan(v);
    if (lane == 63) lds[w] = x;
    __syncthreads();
    T pre = 0, tot = 0;
    for (int i = 0; i < nw; i++) {
        T s = lds[i];
        if (i < w) pre += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return pre + x - v;
}

template <class IT, class OT>
__global__ __launch_bounds__(SB) void k_scan_reduce(const IT* __restrict__ in, uint64_t n,
                                                    OT* __restrict__ bsum) {
    __shared__ OT lds[SB / 64];
    uint64_t base = (uint64_t)blockIdx.x * STILE + (uint64_t)threadIdx.x * SI;
    OT s = 0;
#pragma unroll
    for (int i = 0; i < SI; i++)
        if (base + i < n) s += (OT)in[base + i];
    OT tot;
    block_excl_scan<OT>(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// exclusive scan of `in` into out[0..n]; boff[b] is the offset of block b
// (nullptr: single block), out[n] receives the total from the last block
template <class IT, class OT>
__global__ __launch_bounds__(SB) void k_scan_down(const IT* __restrict__ in, OT* __restrict__ out,
                                                  uint64_t n, const OT* __restrict__ boff) {
    __shared__ OT lds[SB / 64];
    uint64_t base = (uint64_t)blockIdx.x * STILE + (uint64_t)threadIdx.x * SI;
    OT v[SI];
    OT s = 0;
#pragma unroll
    for (int i = 0; i < SI; i++) {
        v[i] = (base + i < n) ? (OT)in[base + i] : (OT)0;
        s += v[i];
    }
    OT tot;
    OT pre = block_excl_scan<OT>(s, lds, &tot);
    OT off = boff ? boff[blockIdx.x] : (OT)0;
    OT run = off + pre;
#pragma unroll
    for (int i = 0; i < SI; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + tot;
}

template <class IT, class OT>
static int scan_impl(const IT* in, OT* out, uint64_t n, Scratch& s, hipStream_t st) {
    uint64_t nb = (n + STILE - 1) / STILE;
    if (nb == 0) nb = 1;
    if (nb == 1) {
        hipLaunchKernelGGL((k_scan_down<IT, OT>), dim3(1), dim3(SB), 0, st, in, out, n,
                           (const OT*)nullptr);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    OT* bsum = (OT*)s.take(nb * sizeof(OT));
    OT* boff = (OT*)s.take((nb + 1) * sizeof(OT));
    if (!bsum || !boff) return -3;
    hipLaunchKernelGGL((k_scan_reduce<IT, OT>), dim3((unsigned)nb), dim3(SB), 0, st, in, n, bsum);
    int r = scan_impl<OT, OT>(bsum, boff, nb, s, st);
    if (r) return r;
    hipLaunchKernelGGL((k_scan_down<IT, OT>), dim3((unsigned)nb), dim3(SB), 0, st, in, out, n,
                       (const OT*)boff);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, Scratch& s, hipStream_t st) {
    return scan_impl<uint32_t, uint32_t>(in, out, n, s, st);
}
int scan_excl_u64(const uint32_t* in, uint64_t* out, uint64_t n, Scratch& s, hipStream_t st) {
    return scan_impl<uint32_t, uint64_t>(in, out, n, s, st);
}

// ------------------------------------------------------------------ radix sort
// 8-bit digits, 256 threads x 16 items per tile.  Tile element e of wave w is
// tile_base + w*1024 + i*64 + lane (i = 0..15), so each wave walks its
// contiguous quarter in index order; ranks inside an item row come from a
// ballot match over the 8 digit bits, which keeps the sort stable.
constexpr int RB = 256;
constexpr int RI = 16;
constexpr int RTILE = RB * RI;
constexpr int RW = RB / 64;

__global__ __launch_bounds__(RB) void k_radix_hist(const uint32_t* __restrict__ keys, uint64_t n,
                                                   uint32_t shift, uint32_t ntiles,
                                                   uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RTILE;
#pragma unroll 4
    for (int i = 0; i < RI; i++) {
        uint64_t e = base + (uint64_t)i * RB + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RB) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, uint64_t n,
                                                      uint32_t shift, uint32_t ntiles,
                                                      const uint32_t* __restrict__ goff) {
    __shared__ uint32_t whist[RW][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < RW * 256; d += RB) (&whist[0][0])[d] = 0;
    __syncthreads();
    const uint64_t wbase = (uint64_t)blockIdx.x * RTILE + (uint64_t)w * (64 * RI);
    uint32_t k[RI], v[RI];
#pragma unroll
    for (int i = 0; i < RI; i++) {
        uint64_t e = wbase + (uint64_t)i * 64 + lane;
        if (e < n) {
            k[i] = kin[e];
            v[i] = vin ? vin[e] : (uint32_t)e;
            atomicAdd(&whist[w][(k[i] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    {
        // digit d: running offset over the waves of this tile
        const int d = threadIdx.x;
        uint32_t b = goff[(uint64_t)d * ntiles + blockIdx.x];
#pragma unroll
        for (int q = 0; q < RW; q++) {
            uint32_t c = whist[q][d];
            whist[q][d] = b;
            b += c;
        }
    }
    __syncthreads();
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int i = 0; i < RI; i++) {
        uint64_t e = wbase + (uint64_t)i * 64 + lane;
        const bool valid = e < n;
        uint32_t d = valid ? ((k[i] >> shift) & 255u) : 256u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 9; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t rank = 0, base = 0;
        if (valid) {
            rank = (uint32_t)__popcll(peers & lt);
            base = whist[w][d];
        }
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            kout[base + rank] = k[i];
            vout[base + rank] = v[i];
            if (rank == 0) whist[w][d] = base + (uint32_t)__popcll(peers);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

int radix_sort_pairs(uint32_t** keys, uint32_t** vals, uint32_t* kalt, uint32_t* valt, uint64_t n,
                     uint32_t bits, bool iota_vals, Scratch& s, hipStream_t st) {
    if (n == 0) return 0;
    uint32_t ntiles = (uint32_t)((n + RTILE - 1) / RTILE);
    uint32_t* hist = (uint32_t*)s.take((uint64_t)256 * ntiles * sizeof(uint32_t));
    uint32_t* goff = (uint32_t*)s.take(((uint64_t)256 * ntiles + 1) * sizeof(uint32_t));
    if (!hist || !goff) return -3;
    uint32_t *kin = *keys, *vin = *vals, *kout = kalt, *vout = valt;
    if (bits == 0) bits = 1;
    bool first = true;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RB), 0, st, kin, n, shift, ntiles,
                           hist);
        int r = scan_excl_u32(hist, goff, (uint64_t)256 * ntiles, s, st);
        if (r) return r;
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RB), 0, st, kin,
                           (first && iota_vals) ? (const uint32_t*)nullptr : vin, kout, vout, n,
                           shift, ntiles, goff);
        first = false;
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys = kin;
    *vals = vin;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------ helpers
__global__ void k_fill_u8(uint8_t* p, uint8_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
void fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256)), dim3(256), 0, st, p, v, n);
}
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, v, n);
}
void iota_u32(uint32_t* p, uint64_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n);
}

template <class W>
__global__ void k_gather(W* __restrict__ dst, const W* __restrict__ src,
                         const uint32_t* __restrict__ idx, uint64_t n, uint32_t words) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t j = idx ? idx[i] : i;
    for (uint32_t w = 0; w < words; w++) dst[i * words + w] = src[j * words + w];
}
void gather_records(void* dst, const void* src, const uint32_t* idx, uint64_t n, uint32_t bytes,
                    hipStream_t st) {
    if (!n || !bytes) return;
    unsigned g = grid_for(n, 256);
    if (bytes % 8 == 0)
        hipLaunchKernelGGL(k_gather<uint64_t>, dim3(g), dim3(256), 0, st, (uint64_t*)dst,
                           (const uint64_t*)src, idx, n, bytes / 8);
    else if (bytes % 4 == 0)
        hipLaunchKernelGGL(k_gather<uint32_t>, dim3(g), dim3(256), 0, st, (uint32_t*)dst,
                           (const uint32_t*)src, idx, n, bytes / 4);
    else if (bytes % 2 == 0)
        hipLaunchKernelGGL(k_gather<uint16_t>, dim3(g), dim3(256), 0, st, (uint16_t*)dst,
                           (const uint16_t*)src, idx, n, bytes / 2);
    else
        hipLaunchKernelGGL(k_gather<uint8_t>, dim3(g), dim3(256), 0, st, (uint8_t*)dst,
                           (const uint8_t*)src, idx, n, bytes);
}

__global__ void k_check_sorted(const uint32_t* __restrict__ k, uint64_t n, uint32_t* flag) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i < n && k[i] < k[i - 1]) *flag = 0;
}
void check_sorted(const uint32_t* keys, uint64_t n, uint32_t* flag, hipStream_t st) {
    fill_u32(flag, 1, 1, st);
    if (n > 1)
        hipLaunchKernelGGL(k_check_sorted, dim3(grid_for(n, 256)), dim3(256), 0, st, keys, n,
                           flag);
}

}  // namespace vh
