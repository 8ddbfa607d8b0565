// modarith.h -- exact 64-bit modular arithmetic for gfx950 VALU.
//
// CDNA4 has no 64x64 multiply; every product below is lowered by the compiler to
// 32-bit v_mad_u64_u32 / v_mul_hi_u32 / v_mul_lo_u32 chains.  Moduli are < 2^61,
// so the lazy ranges [0,2q) and [0,4q) used by the NTT never overflow 64 bits.
//
// Semantics follow the reference's small-modulus arithmetic
// (src/utils/uintarithsmallmod.h:95-230, device twins src/kernelutils.cuh:99-189):
//   * Shoup multiply by a precomputed operand (w, floor(w 2^64 / q)),
//   * Barrett reduction of 64- and 128-bit values with floor(2^128 / q).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace shc {

struct u128 {
    uint64_t lo, hi;
};

__device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) { return __umul64hi(a, b); }

__device__ __forceinline__ u128 mul64x64(uint64_t a, uint64_t b) {
    u128 r;
    r.lo = a * b;
    r.hi = __umul64hi(a, b);
    return r;
}

__device__ __forceinline__ void add128(u128 &acc, u128 v) {
    uint64_t lo = acc.lo + v.lo;
    acc.hi += v.hi + (lo < acc.lo ? 1ull : 0ull);
    acc.lo = lo;
}

// x * w mod q in [0, 2q), for any 64-bit x and w < q (Shoup, lazy).
__device__ __forceinline__ uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wq, uint64_t q) {
    uint64_t hi = mulhi64(x, wq);
    return x * w - hi * q;
}

// floor(x * y / 2^64) - e with e in {0, 1, 2}: the high word without the carries of
// the two middle partial products' low halves and of x_lo * y_lo (3 multiplies).
__host__ __device__ __forceinline__ uint64_t mulhi64_approx(uint64_t x, uint64_t y) {
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), yl = (uint32_t)y, yh = (uint32_t)(y >> 32);
    const uint64_t mid = ((uint64_t)xh * yl >> 32) + ((uint64_t)xl * yh >> 32);
    return (uint64_t)xh * yh + mid;
}

// x * w mod q in [0, 4q), for any 64-bit x, w < q < 2^62, wq = floor(w 2^64 / q) and
// neg_q = 2^64 - q.  The quotient estimate is Shoup's minus at most 2, so the remainder
// x w - Q q lies in [0, 4q) and equals its low 64 bits, computed as x w + Q neg_q with
// only the low halves of the cross products (9 multiplies in all).
__host__ __device__ __forceinline__ uint64_t shoup_lazier(uint64_t x, uint64_t w, uint64_t wq, uint64_t neg_q) {
    const uint64_t Q = mulhi64_approx(x, wq);
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32), wl = (uint32_t)w, wh = (uint32_t)(w >> 32);
    const uint32_t ql = (uint32_t)Q, qh = (uint32_t)(Q >> 32), nl = (uint32_t)neg_q, nh = (uint32_t)(neg_q >> 32);
    const uint64_t p = (uint64_t)xl * wl + (uint64_t)ql * nl;
    // the cross products only reach the high word: 32-bit arithmetic there
    uint32_t hi = (uint32_t)(p >> 32) + xl * wh + xh * wl + ql * nh + qh * nl;
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(hi));  // keeps the 32-bit add from being widened back to 64 bits
#endif
    return ((uint64_t)hi << 32) | (uint32_t)p;
}

__device__ __forceinline__ uint64_t shoup(uint64_t x, uint64_t w, uint64_t wq, uint64_t q) {
    uint64_t r = shoup_lazy(x, w, wq, q);
    return r >= q ? r - q : r;
}

__device__ __forceinline__ uint64_t csub(uint64_t x, uint64_t q) { return x >= q ? x - q : x; }

// x mod q for any 64-bit x, using the high word of floor(2^128/q) (barrettReduce64).
__device__ __forceinline__ uint64_t barrett64(uint64_t x, uint64_t q, uint64_t ratio_hi) {
    uint64_t est = mulhi64(x, ratio_hi);
    uint64_t r = x - est * q;
    return r >= q ? r - q : r;
}

// (hi:lo) mod q for any 128-bit value (barrettReduce128, uintarithsmallmod.h:95-126).
__device__ __forceinline__ uint64_t barrett128(u128 x, uint64_t q, uint64_t ratio_lo, uint64_t ratio_hi) {
    uint64_t carry = mulhi64(x.lo, ratio_lo);
    u128 t2 = mul64x64(x.lo, ratio_hi);
    uint64_t tmp1 = t2.lo + carry;
    uint64_t tmp3 = t2.hi + (tmp1 < carry ? 1ull : 0ull);
    t2 = mul64x64(x.hi, ratio_lo);
    uint64_t s = tmp1 + t2.lo;
    carry = t2.hi + (s < tmp1 ? 1ull : 0ull);
    uint64_t est = x.hi * ratio_hi + tmp3 + carry;
    uint64_t r = x.lo - est * q;
    return r >= q ? r - q : r;
}

__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q, uint64_t ratio_lo,
                                           uint64_t ratio_hi) {
    return barrett128(mul64x64(a, b), q, ratio_lo, ratio_hi);
}

__device__ __forceinline__ uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t submod(uint64_t a, uint64_t b, uint64_t q) {
    return a >= b ? a - b : a + q - b;
}

__device__ __forceinline__ uint64_t negmod(uint64_t a, uint64_t q) { return a ? q - a : 0; }

}  // namespace shc
