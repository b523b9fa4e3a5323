// vh_prims.h — device-wide primitives used by the transition engine:
// exclusive scans, a stable LSD radix sort of (uint32 key, uint32 value)
// pairs, fills and gathers.  All launches are asynchronous on `st`; scratch
// memory comes from the caller's Scratch arena.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vh {

// Grow-only device scratch arena owned by one simulation.  take() hands out
// 256-byte aligned pieces valid until the next reset(); a round that overflows
// the main block spills into extra blocks, which reset() folds into one larger
// main block (after synchronising the stream) so steady-state steps allocate
// nothing.
// hipMalloc for everything the library allocates; fail_next_allocs(n) makes
// the next n calls return hipErrorOutOfMemory (injected failures for tests)
hipError_t device_alloc(void** p, size_t bytes);
void fail_next_allocs(uint32_t n);

struct Scratch {
    void* main = nullptr;
    size_t cap = 0, used = 0, spilled = 0;
    void* extra[64];
    int nextra = 0;
    char* xb = nullptr;                  // spill block handed out piecewise (room())
    size_t xcap = 0, xused = 0;
    bool shrink = false;                 // at the next reset: release when more than 1 GiB is held
    void* take(size_t bytes);            // nullptr on allocation failure
    void reset(hipStream_t st);
    // after reset: grow the main block to at least `bytes` (work that takes
    // many pieces before its next reset); 0 or -3
    int reserve(size_t bytes, hipStream_t st);
    // at any point: room for `bytes` more in one piece (the main block's rest,
    // or one spill block that the following takes share); 0 or -3
    int room(size_t bytes);
    void release();
    size_t bytes_held() const { return cap + spilled; }
};

// out[i] = sum(in[0..i-1]) for i in [0, n], out has n+1 entries
int scan_excl_u32(const uint32_t* in, uint32_t* out, uint64_t n, Scratch& s, hipStream_t st);
int scan_excl_u64(const uint32_t* in, uint64_t* out, uint64_t n, Scratch& s, hipStream_t st);

// stable sort of keys (and the values riding along) by the low `bits` bits;
// keys_alt/vals_alt are ping-pong buffers of n entries; the result is left in
// (*keys, *vals) (pointers may be swapped).  iota_vals: the values are
// 0..n-1 and *vals need not be initialised.
int radix_sort_pairs(uint32_t** keys, uint32_t** vals, uint32_t* keys_alt, uint32_t* vals_alt,
                     uint64_t n, uint32_t bits, bool iota_vals, Scratch& s, hipStream_t st);

void fill_u8(uint8_t* p, uint8_t v, uint64_t n, hipStream_t st);
void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st);
void iota_u32(uint32_t* p, uint64_t n, hipStream_t st);
// dst[i] = src[idx[i]] for records of `bytes` bytes (bytes % 4 == 0 or 1/2)
void gather_records(void* dst, const void* src, const uint32_t* idx, uint64_t n, uint32_t bytes,
                    hipStream_t st);
// 1 if keys[0..n) is non-decreasing, written to *flag (device)
void check_sorted(const uint32_t* keys, uint64_t n, uint32_t* flag, hipStream_t st);

inline unsigned grid_for(uint64_t n, unsigned block) {
    uint64_t g = (n + block - 1) / block;
    return (unsigned)(g ? g : 1);
}

}  // namespace vh
