// butterfly.h -- the NTT's lazy radix-2 butterflies (host and device).
#pragma once
#include "modarith.h"

namespace shc {

// v >= c ? v - c : v, for c < 2^63 and v < c + 2^63: v - c wraps above 2^63 exactly
// when v < c, so the sign of the high word selects.
__host__ __device__ __forceinline__ uint64_t sub_if_ge(uint64_t v, uint64_t c) {
    const uint64_t t = v - c;
    return (int32_t)(t >> 32) < 0 ? v : t;
}

// The same with neg_c = 2^64 - c: one 64-bit add (v_lshl_add_u64) where v - c takes a
// subtract and a subtract-with-borrow.
__host__ __device__ __forceinline__ uint64_t sub_if_ge_neg(uint64_t v, uint64_t neg_c) {
    const uint64_t t = v + neg_c;
    return (int32_t)(t >> 32) < 0 ? v : t;
}

// Forward (Cooley-Tukey, x + w y, x - w y): u = x, less red_q when red and x >= red_q
// (neg_red = 2^64 - red_q); outputs u + v and u + 4q - v with the twiddle product
// v = shoup_lazier(y) in [0, 4q) (y may be any 64-bit word).  Schedules:
//   standard, q < 2^61: red with red_q = 4q on every layer, [0, 8q) -> [0, 8q);
//   wide, q < 2^60: red with red_q = 8q takes x < 16q to outputs < 12q, and no
//     reduction takes x < 12q to outputs < 16q (16q <= 2^64).
// Inverse (Gentleman-Sande, x + y, (x - y) w): x, y in [0, 4q) -> [0, 4q); neg_red is
// 2^64 - 4q.
// Same congruences as the reference's Harvey butterflies (src/utils/dwthandler.h:88-372),
// with wider lazy ranges.
template <bool INV>
__host__ __device__ __forceinline__ void butterfly(uint64_t &x, uint64_t &y, ulonglong2 w, uint64_t four_q,
                                                   uint64_t neg_q, bool red, uint64_t neg_red) {
    if (!INV) {
        const uint64_t u = red ? sub_if_ge_neg(x, neg_red) : x;
        const uint64_t v = shoup_lazier(y, w.x, w.y, neg_q);
        x = u + v;
        y = u + four_q - v;
    } else {
        const uint64_t s = x + y;
        const uint64_t d = x + four_q - y;
        x = sub_if_ge_neg(s, neg_red);
        y = shoup_lazier(d, w.x, w.y, neg_q);
    }
}

// The standard schedule.
template <bool INV>
__host__ __device__ __forceinline__ void butterfly(uint64_t &x, uint64_t &y, ulonglong2 w, uint64_t four_q,
                                                   uint64_t neg_q) {
    butterfly<INV>(x, y, w, four_q, neg_q, true, 0 - four_q);
}

}  // namespace shc
