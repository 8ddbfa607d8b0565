// accum.h -- sums of 64 x 64-bit products without carry chains (key-switch inner
// products).
//
// A sum of products t * k with t, k < 2^63 is kept as three 64-bit words
//   value = al + ml * 2^32 + h * 2^64,
// where each product adds
//   al += lo32(tl kl),  ml += hi32(tl kl) + lo32(m),  h += th kh + hi32(m),
//   m = th kl + tl kh
// (tl, th, kl, kh: 32-bit halves).  m < 2^64 because th, kh < 2^31; al grows by less
// than 2^32 and ml by less than 2^33 per product; h stays below 2^64 while the value is
// below 2^128, which key switching guarantees: its terms are below 8q * q < 2^123
// (q < 2^60) and a level has at most 31 digits at that size (63 below 4q * q) plus one
// correction term.  Every step is one v_mad_u64_u32 (a 32-bit word joins a 64-bit sum
// as word * 1 + sum; 64-bit products take one each): 8 VALU instructions per product,
// with no carry chains and no 64-bit shifts.  The device code states the eight
// multiply-adds in inline assembly, since the compiler otherwise re-associates the
// 32-bit pieces into 64-bit adds and register moves; they go in three groups (the two
// products of tl, then the ones that need those products' halves, then the rest)
// because the compiler puts a wait state between consecutive inline-assembly
// statements that share a register.  The same
// sums as the reference's 128-bit accumulation of key-switch products
// (src/evaluator.cpp:2520-2590: multiply_uint64 + add_uint128, then barrett_reduce_128).
#pragma once
#include "common.h"
#include "modarith.h"

namespace shc {

struct Acc3 {
    uint64_t al, ml, h;
};

// value + t * k.  KU: k is uniform across the wave (an SGPR operand, e.g. a conversion
// table entry; one per instruction, the constant-bus limit).
template <bool KU = false>
__host__ __device__ __forceinline__ void mac(Acc3 &a, uint64_t t, uint64_t k) {
    const uint32_t tl = (uint32_t)t, th = (uint32_t)(t >> 32), kl = (uint32_t)k, kh = (uint32_t)(k >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t p, m, c;
    if constexpr (KU)
        asm("v_mad_u64_u32 %0, %2, %3, %4, 0\n\tv_mad_u64_u32 %1, %2, %3, %5, 0"
            : "=&v"(p), "=&v"(m), "=&s"(c) : "v"(tl), "s"(kl), "s"(kh));
    else
        asm("v_mad_u64_u32 %0, %2, %3, %4, 0\n\tv_mad_u64_u32 %1, %2, %3, %5, 0"
            : "=&v"(p), "=&v"(m), "=&s"(c) : "v"(tl), "v"(kl), "v"(kh));
    if constexpr (KU)
        asm("v_mad_u64_u32 %0, %3, %4, %5, %0\n\tv_mad_u64_u32 %1, %3, %6, 1, %1\n\tv_mad_u64_u32 %2, %3, %7, 1, %2"
            : "+v"(m), "+v"(a.al), "+v"(a.ml), "=&s"(c) : "v"(th), "s"(kl), "v"((uint32_t)p), "v"((uint32_t)(p >> 32)));
    else
        asm("v_mad_u64_u32 %0, %3, %4, %5, %0\n\tv_mad_u64_u32 %1, %3, %6, 1, %1\n\tv_mad_u64_u32 %2, %3, %7, 1, %2"
            : "+v"(m), "+v"(a.al), "+v"(a.ml), "=&s"(c) : "v"(th), "v"(kl), "v"((uint32_t)p), "v"((uint32_t)(p >> 32)));
    if constexpr (KU)
        asm("v_mad_u64_u32 %0, %2, %3, 1, %0\n\tv_mad_u64_u32 %1, %2, %4, 1, %1\n\tv_mad_u64_u32 %1, %2, %5, %6, %1"
            : "+v"(a.ml), "+v"(a.h), "=&s"(c) : "v"((uint32_t)m), "v"((uint32_t)(m >> 32)), "v"(th), "s"(kh));
    else
        asm("v_mad_u64_u32 %0, %2, %3, 1, %0\n\tv_mad_u64_u32 %1, %2, %4, 1, %1\n\tv_mad_u64_u32 %1, %2, %5, %6, %1"
            : "+v"(a.ml), "+v"(a.h), "=&s"(c) : "v"((uint32_t)m), "v"((uint32_t)(m >> 32)), "v"(th), "v"(kh));
#else
    const uint64_t p = (uint64_t)tl * kl;
    a.al += (uint32_t)p;
    a.ml += p >> 32;
    const uint64_t m = (uint64_t)th * kl + (uint64_t)tl * kh;
    a.ml += (uint32_t)m;
    a.h += (uint64_t)th * kh + (m >> 32);
#endif
}

// value + x for x < 2^64 (one more term, e.g. a correction).
__host__ __device__ __forceinline__ void acc_add(Acc3 &a, uint64_t x) {
    a.al += (uint32_t)x;
    a.ml += x >> 32;
}

__host__ __device__ __forceinline__ u128 acc_value(const Acc3 &a) {
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
