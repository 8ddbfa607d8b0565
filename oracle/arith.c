/*
 * arith.c -- modular arithmetic, prime selection and primitive roots.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

typedef unsigned __int128 u128;

uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q) {
    /* exponentiateUintMod (utils/uintarithsmallmod.h) by square and multiply */
    uint64_t r = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) r = or_mulmod(r, a, q);
        a = or_mulmod(a, a, q);
        e >>= 1;
    }
    return r;
}

uint64_t or_invmod(uint64_t a, uint64_t q) {
    /* extended Euclid, as tryInvertUintMod (utils/uintarithsmallmod.h:311) via xgcd (utils/numth.h) */
    if (q < 2) return 0;
    a %= q;
    if (a == 0) return 0;
    __int128 t = 0, nt = 1, r = q, nr = a;
    while (nr) {
        __int128 quo = r / nr, tmp;
        tmp = t - quo * nt; t = nt; nt = tmp;
        tmp = r - quo * nr; r = nr; nr = tmp;
    }
    if (r != 1) return 0;
    if (t < 0) t += q;
    return (uint64_t)t;
}

uint64_t or_shoup(uint64_t w, uint64_t q) {
    /* MultiplyUIntModOperand::setQuotient (utils/uintarithsmallmod.h:170-178) */
    return (uint64_t)(((u128)w << 64) / q);
}

void or_barrett_ratio(uint64_t q, uint64_t out[3]) {
    /* Modulus::setValue (modulus.cpp:28-38): floor(2^128 / q) in two words plus remainder */
    u128 hi = ((u128)1 << 64) / q;            /* top word of 2^128 / q                  */
    u128 rem_hi = ((u128)1 << 64) % q;        /* remainder after the top word          */
    u128 num = rem_hi << 64;                  /* rem_hi * 2^64 < q * 2^64 fits in 128   */
    u128 lo = num / q;
    u128 rem = num % q;
    out[0] = (uint64_t)lo;
    out[1] = (uint64_t)hi;
    out[2] = (uint64_t)rem;
}

uint32_t or_reverse_bits(uint32_t x, int bits) {
    if (bits == 0) return 0;
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

int or_is_prime(uint64_t v) {
    /* isPrime (utils/numth.cpp:163-259) decides primality by Miller-Rabin; the
       deterministic base set below is exact for every 64-bit input.            */
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (v < 2) return 0;
    for (int i = 0; i < 12; i++) {
        if (v == small[i]) return 1;
        if (v % small[i] == 0) return 0;
    }
    uint64_t d = v - 1;
    int r = 0;
    while (!(d & 1)) { d >>= 1; r++; }
    for (int i = 0; i < 12; i++) {
        uint64_t x = or_powmod(small[i], d, v);
        if (x == 1 || x == v - 1) continue;
        int ok = 0;
        for (int k = 1; k < r; k++) {
            x = or_mulmod(x, x, v);
            if (x == v - 1) { ok = 1; break; }
        }
        if (!ok) return 0;
    }
    return 1;
}

int or_get_primes(uint64_t factor, int bit_size, size_t count, uint64_t *out) {
    /* getPrimes (utils/numth.cpp:261-283) */
    uint64_t value = ((((uint64_t)1) << bit_size) - 1) / factor * factor + 1;
    uint64_t lower = ((uint64_t)1) << (bit_size - 1);
    size_t found = 0;
    while (found < count && value > lower) {
        if (or_is_prime(value)) out[found++] = value;
        value -= factor;
    }
    return found == count ? 0 : -1;
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}

int or_coeff_modulus_create(size_t n, const int *bit_sizes, size_t count, uint64_t plain_modulus,
                            uint64_t *out) {
    /* CoeffModulus::Create (modulus.cpp:80-121 and 123-165): primes of each bit size are
       found in descending order, and handed out from the back of that list, so equal
       bit sizes receive ascending primes in order of appearance.                      */
    uint64_t factor = 2 * (uint64_t)n;
    if (plain_modulus) factor *= plain_modulus / gcd64(plain_modulus, factor);
    for (size_t i = 0; i < count; i++) {
        int b = bit_sizes[i];
        if (b < 2 || b > 60) return -1;
        size_t total = 0, before = 0;
        for (size_t j = 0; j < count; j++) {
            if (bit_sizes[j] == b) {
                total++;
                if (j < i) before++;
            }
        }
        uint64_t primes[64];
        if (total > 64) return -1;
        if (or_get_primes(factor, b, total, primes)) return -1;
        out[i] = primes[total - 1 - before];
    }
    return 0;
}

int or_minimal_primitive_root(uint64_t degree, uint64_t q, uint64_t *root) {
    /* tryMinimalPrimitiveRoot (utils/numth.cpp:335-362): the smallest element of
       order exactly `degree` (a power of two).  The reference starts from a random
       primitive root; every primitive root generates the same set of odd powers,
       so the minimum does not depend on the start.                              */
    if ((q - 1) % degree) return -1;
    uint64_t quo = (q - 1) / degree;
    uint64_t g = 0;
    for (uint64_t c = 2; c < q && c < 1000000; c++) {
        uint64_t r = or_powmod(c, quo, q);
        if (or_powmod(r, degree >> 1, q) == q - 1) { g = r; break; }
    }
    if (!g) return -1;
    uint64_t g2 = or_mulmod(g, g, q), cur = g, best = g;
    for (uint64_t i = 0; i < degree; i += 2) {
        if (cur < best) best = cur;
        cur = or_mulmod(cur, g2, q);
    }
    *root = best;
    return 0;
}
