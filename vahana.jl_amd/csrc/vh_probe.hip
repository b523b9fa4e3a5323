// vh_probe.hip — memory-system probes that bound the transition kernels:
// random 8-byte gathers over a footprint (what a neighbour-state gather costs
// when the graph has no locality) and a streaming read.
#include "vh_engine.h"

using namespace vh;

namespace {
__global__ void k_probe_gather(const double* __restrict__ x, uint64_t nx, uint64_t n, double* out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    for (uint64_t k = i; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        acc += x[(vh_mix64(k) >> 32) * nx >> 32];
    if (acc == 1.2345e300) out[0] = acc;   // never true: keeps the loads
}
// a fraction `hot_permille` of the gathers goes to the first `hot` elements
// (normal loads), the rest is spread over the footprint with non-temporal
// loads when nt != 0 (so they do not displace the hot prefix from L2)
__global__ void k_probe_hotcold(const double* __restrict__ x, uint64_t nx, uint64_t hot,
                                uint32_t hot_permille, uint32_t nt, uint64_t n, double* out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    for (uint64_t k = i; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t h = vh_mix64(k);
        if ((h & 1023u) < hot_permille * 1024u / 1000u) {
            acc += x[(h >> 32) * hot >> 32];
        } else {
            const double* p = x + ((h >> 32) * nx >> 32);
            acc += nt ? __builtin_nontemporal_load(p) : *p;
        }
    }
    if (acc == 1.2345e300) out[0] = acc;
}
// streaming read with 16-byte loads, 4 in flight per thread
typedef uint32_t u32x4p __attribute__((ext_vector_type(4)));
__global__ void k_probe_stream(const u32x4p* __restrict__ a, uint64_t n, double* out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    u32x4p acc = {0, 0, 0, 0};
    uint64_t k = i;
    for (; k + 3 * stride < n; k += 4 * stride) {
        u32x4p v0 = __builtin_nontemporal_load(a + k), v1 = __builtin_nontemporal_load(a + k + stride);
        u32x4p v2 = __builtin_nontemporal_load(a + k + 2 * stride);
        u32x4p v3 = __builtin_nontemporal_load(a + k + 3 * stride);
        acc += v0 ^ v1 ^ v2 ^ v3;
    }
    for (; k < n; k += stride) acc += a[k];
    if (acc.x + acc.y + acc.z + acc.w == 0x12345679u) out[0] = 1.0;
}
}  // namespace

// Device memory allocated with hipDeviceMallocUncached or
// hipDeviceMallocFinegrained for the probes below is kept until the process
// ends (freeing uncached memory that a kernel has read is not safe here);
// one buffer per kind, reused while large enough.
static void* g_special[2] = {nullptr, nullptr};
static size_t g_special_bytes[2] = {0, 0};

static int special_buffer(int kind, size_t bytes, void** out) {
    if (g_special_bytes[kind] < bytes) {
        void* p = nullptr;
        unsigned flags = kind == 0 ? hipDeviceMallocUncached : hipDeviceMallocFinegrained;
        HIPCHECK(hipExtMallocWithFlags(&p, bytes, flags));
        g_special[kind] = p;   // a smaller previous buffer stays allocated
        g_special_bytes[kind] = bytes;
    }
    *out = g_special[kind];
    return VH_OK;
}

// mode 0: n random 8-byte gathers over `footprint` bytes; mode 1: streaming
// read of `footprint` bytes; modes 2/3: 22% of the gathers into a 2 MiB hot
// prefix, the rest over the footprint with normal (2) or non-temporal (3)
// loads; modes 4/5: as mode 0 over uncached (4) or fine-grained (5) device
// memory.  *ms = mean of `reps` launches.
extern "C" int vh_probe_memory(int mode, uint64_t footprint, uint64_t n, uint32_t reps, float* ms) {
    if (mode < 0 || mode > 5) return set_error(VH_ERR_ARG, "probe mode %d", mode);
    void* buf = nullptr;
    double* out = nullptr;
    if (mode == 4 || mode == 5)
        RC(special_buffer(mode - 4, footprint, &buf));
    else
        HIPCHECK(hipMalloc(&buf, footprint));
    HIPCHECK(hipMalloc(&out, 8));
    HIPCHECK(hipMemset(buf, 0, footprint));
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    dim3 g(256 * 64), blk(256);
    auto launch = [&]() {
        if (mode == 0 || mode == 4 || mode == 5)
            hipLaunchKernelGGL(k_probe_gather, g, blk, 0, 0, (const double*)buf, footprint / 8, n, out);
        else if (mode == 2 || mode == 3)
            hipLaunchKernelGGL(k_probe_hotcold, g, blk, 0, 0, (const double*)buf, footprint / 8,
                               (uint64_t)(2u << 20) / 8, 220u, mode == 3 ? 1u : 0u, n, out);
        else
            hipLaunchKernelGGL(k_probe_stream, g, blk, 0, 0, (const u32x4p*)buf, footprint / 16, out);
    };
    launch();
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipEventRecord(a, 0));
    for (uint32_t r = 0; r < reps; r++) launch();
    HIPCHECK(hipEventRecord(b, 0));
    HIPCHECK(hipEventSynchronize(b));
    float t = 0;
    HIPCHECK(hipEventElapsedTime(&t, a, b));
    *ms = t / (reps ? reps : 1);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (mode != 4 && mode != 5) (void)hipFree(buf);
    (void)hipFree(out);
    return VH_OK;
}
