// vh_common.h — shared device/host definitions of the MI355X transition backend.
//
// The public device API (include/vahana_device.h: device views, accessors,
// codec, RNG) and what only the library's model bodies use.
#pragma once

#include "../../include/vahana_device.h"

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
// stream keys of the model bodies (the `c` word of vh_hash)
#define VH_KEY_RECOVER (1ull << 40)
#define VH_KEY_DEATH   (2ull << 40)
#define VH_KEY_BIRTH   (3ull << 40)
#define VH_KEY_PARTNER (4ull << 40)
#define VH_KEY_INIT    (5ull << 40)
#define VH_KEY_WEIGHT  (6ull << 40)
#define VH_KEY_OPINION (7ull << 40)
#define VH_KEY_SUSC    (8ull << 40)

