// accum.h -- sums of 64 x 64-bit products without carry chains (key-switch inner
// products).
//
// A sum of products t * k with t, k < 2^63 is kept as three 64-bit words
//   value = al + ml * 2^32 + h * 2^64,
// where each product adds
//   al += lo32(tl kl),  m = th kl + tl kh + hi32(tl kl),  ml += lo32(m),
//   h  += th kh + hi32(m)
// (tl, th, kl, kh: 32-bit halves).  m < 2^64 because th, kh < 2^31; al and ml grow by
// less than 2^32 per product; h stays below 2^64 while the value is below 2^128, which
// key switching guarantees: its terms are below 4q * q < 2^122 (q < 2^60) and a level
// has at most 63 digits plus one correction term.  Every
// step is one v_mad_u64_u32 (a 32-bit word joins a 64-bit sum as word * 1 + sum), about
// 9 VALU instructions per product instead of a 128-bit add with carries.  The same
// sums as the reference's 128-bit accumulation of key-switch products
// (src/evaluator.cpp:2520-2590: multiply_uint64 + add_uint128, then barrett_reduce_128).
#pragma once
#include "common.h"
#include "modarith.h"

namespace shc {

struct Acc3 {
    uint64_t al, ml, h;
};

// A uniform 1 the compiler cannot fold away, so that "word * one + sum" stays a
// multiply-add.
__device__ __forceinline__ uint32_t opaque_one() {
    uint32_t one = 1;
    asm volatile("" : "+s"(one));
    return one;
}

__device__ __forceinline__ void mac(Acc3 &a, uint64_t t, uint64_t k, uint32_t one) {
    const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32), kl = (uint32_t)k, kh = (uint32_t)(k >> 32);
    const uint64_t p = (uint64_t)tl * kl;
    a.al = (uint64_t)(uint32_t)p * one + a.al;
    const uint64_t m = (uint64_t)th * kl + ((uint64_t)tl * kh + (p >> 32));
    a.ml = (uint64_t)(uint32_t)m * one + a.ml;
    a.h = (uint64_t)th * kh + ((uint64_t)(uint32_t)(m >> 32) * one + a.h);
}

// value + x for x < 2^64 (one more term, e.g. a correction).
__device__ __forceinline__ void acc_add(Acc3 &a, uint64_t x, uint32_t one) {
    a.al = (uint64_t)(uint32_t)x * one + a.al;
    a.ml = (uint64_t)(uint32_t)(x >> 32) * one + a.ml;
}

__device__ __forceinline__ u128 acc_value(const Acc3 &a) {
    const uint64_t lo = a.al + (a.ml << 32);
    return u128{lo, a.h + (a.ml >> 32) + (lo < a.al ? 1ull : 0ull)};
}

// value mod q (< q).
__device__ __forceinline__ uint64_t acc_reduce(const Acc3 &a, const DevPrime &Q) {
    return barrett128(acc_value(a), Q.q, Q.ratio_lo, Q.ratio_hi);
}

// Replaces the value by value mod q (for sums longer than the 128-bit bound allows).
__device__ __forceinline__ void acc_fold(Acc3 &a, const DevPrime &Q) {
    const uint64_t r = acc_reduce(a, Q);
    a = Acc3{(uint32_t)r, r >> 32, 0};
}

}  // namespace shc
