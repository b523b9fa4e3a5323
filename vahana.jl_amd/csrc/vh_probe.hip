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
__global__ void k_probe_stream(const uint32_t* __restrict__ a, uint64_t n, double* out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc = 0;
    for (uint64_t k = i; k < n; k += (uint64_t)gridDim.x * blockDim.x) acc += a[k];
    if (acc == 0x123456789ull) out[0] = (double)acc;
}
}  // namespace

// mode 0: n random 8-byte gathers over `footprint` bytes; mode 1: streaming
// read of `footprint` bytes; modes 2/3: 22% of the gathers into a 2 MiB hot
// prefix, the rest over the footprint with normal (2) or non-temporal (3)
// loads.  *ms = mean of `reps` launches.
extern "C" int vh_probe_memory(int mode, uint64_t footprint, uint64_t n, uint32_t reps, float* ms) {
    void* buf = nullptr;
    double* out = nullptr;
    HIPCHECK(hipMalloc(&buf, footprint));
    HIPCHECK(hipMalloc(&out, 8));
    HIPCHECK(hipMemset(buf, 0, footprint));
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    dim3 g(256 * 64), blk(256);
    auto launch = [&]() {
        if (mode == 0)
            hipLaunchKernelGGL(k_probe_gather, g, blk, 0, 0, (const double*)buf, footprint / 8, n, out);
        else if (mode == 2 || mode == 3)
            hipLaunchKernelGGL(k_probe_hotcold, g, blk, 0, 0, (const double*)buf, footprint / 8,
                               (uint64_t)(2u << 20) / 8, 220u, mode == 3 ? 1u : 0u, n, out);
        else
            hipLaunchKernelGGL(k_probe_stream, g, blk, 0, 0, (const uint32_t*)buf, footprint / 4, out);
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
    (void)hipFree(buf);
    (void)hipFree(out);
    return VH_OK;
}
