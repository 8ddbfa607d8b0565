// butterfly.h -- the NTT's lazy radix-2 butterflies (host and device).
#pragma once
#include "modarith.h"

namespace shc {

// Forward (Cooley-Tukey, x + w y, x - w y): x, y in [0, 8q) -> [0, 8q).
// Inverse (Gentleman-Sande, x + y, (x - y) w): x, y in [0, 4q) -> [0, 4q).
// The twiddle product is shoup_lazier's, in [0, 4q); 8q fits 64 bits for q < 2^61.
// Same congruences as the reference's Harvey butterflies (src/utils/dwthandler.h:88-372),
// with wider lazy ranges.
template <bool INV>
__host__ __device__ __forceinline__ void butterfly(uint64_t &x, uint64_t &y, ulonglong2 w, uint64_t four_q,
                                                   uint64_t neg_q) {
    // v >= 4q ? v - 4q : v for v < 8q: with 4q < 2^63, v - 4q wraps above 2^63 exactly
    // when v < 4q, so the sign of the high word selects
    auto sub_if_ge = [four_q](uint64_t v) {
        const uint64_t t = v - four_q;
        return (int32_t)(t >> 32) < 0 ? v : t;
    };
    if (!INV) {
        const uint64_t u = sub_if_ge(x);
        const uint64_t v = shoup_lazier(y, w.x, w.y, neg_q);
        x = u + v;
        y = u + four_q - v;
    } else {
        const uint64_t s = x + y;
        const uint64_t d = x + four_q - y;
        x = sub_if_ge(s);
        y = shoup_lazier(d, w.x, w.y, neg_q);
    }
}

}  // namespace shc
