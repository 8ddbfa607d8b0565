/*
 * sampler.c -- the samplers' word streams and distributions, restated for the checks.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The reference samples with Blake2xb on the
 * CPU (src/utils/rlwe.cpp:21-160, src/randomgen.cpp) and curand on the GPU
 * (src/utils/rlwe_cuda.cu:25-150); neither stream can be reproduced, so this build has
 * its own, documented streams, restated here word for word:
 *   - keyed: ChaCha20 (D. J. Bernstein, "ChaCha, a variant of Salsa20", 2008; the block
 *     function of RFC 8439 section 2.3 with state words 12-15 = 64-bit block counter,
 *     64-bit nonce), word i of seed s = words 2k, 2k+1 of block i / 8 under nonce s,
 *     k = i % 8;
 *   - splitmix64: word i of seed s = mix(s + (i + 1) * 0x9E3779B97F4A7C15).
 * Distributions: ternary = floor(3 w / 2^64) - 1; centred binomial = popcount of bits
 * 0-20 minus popcount of bits 21-41 (the reference's CBD, src/utils/rlwe.cpp:70-105,
 * uses 42 bits the same way); uniform mod q = (word 2i * 2^64 + word 2i+1) mod q.
 * Pinned by: the RFC 8439 section 2.3.2 block test vector and OpenSSL's ChaCha20
 * keystream (tests/test_oracle_sampler.py); the distributions by their closed forms.
 */
#include "oracle.h"

static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static void quarter(uint32_t *x, int a, int b, int c, int d) {
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 16);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 12);
    x[a] += x[b]; x[d] ^= x[a]; x[d] = rotl32(x[d], 8);
    x[c] += x[d]; x[b] ^= x[c]; x[b] = rotl32(x[b], 7);
}

void or_chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, uint32_t out[16]) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; i++) in[4 + i] = key[i];
    in[12] = (uint32_t)counter;
    in[13] = (uint32_t)(counter >> 32);
    in[14] = (uint32_t)nonce;
    in[15] = (uint32_t)(nonce >> 32);
    uint32_t x[16];
    for (int i = 0; i < 16; i++) x[i] = in[i];
    for (int r = 0; r < 10; r++) {
        quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13); quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
        quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12); quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

static uint64_t splitmix_word(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t stream_word(const uint32_t *key, uint64_t seed, uint64_t i) {
    if (!key) return splitmix_word(seed, i);
    uint32_t b[16];
    or_chacha20_block(key, i / 8, seed, b);
    const int k = (int)(i % 8);
    return (uint64_t)b[2 * k] | ((uint64_t)b[2 * k + 1] << 32);
}

int or_sample_poly(int kind, const uint32_t *key, uint64_t seed, int logn, const uint64_t *moduli, size_t L,
                   uint64_t *out) {
    const size_t n = (size_t)1 << logn;
    if (kind == 2) {
        for (size_t i = 0; i < L * n; i++) {
            const unsigned __int128 x =
                ((unsigned __int128)stream_word(key, seed, 2 * i) << 64) | stream_word(key, seed, 2 * i + 1);
            out[i] = (uint64_t)(x % moduli[i / n]);
        }
        return 0;
    }
    if (kind != 0 && kind != 1) return -1;
    for (size_t c = 0; c < n; c++) {
        const uint64_t w = stream_word(key, seed, c);
        int64_t s;
        if (kind == 0) s = (int64_t)(((unsigned __int128)w * 3) >> 64) - 1;
        else s = (int64_t)__builtin_popcountll(w & 0x1FFFFFull) - (int64_t)__builtin_popcountll((w >> 21) & 0x1FFFFFull);
        for (size_t l = 0; l < L; l++) out[l * n + c] = s < 0 ? moduli[l] - (uint64_t)(-s) : (uint64_t)s;
    }
    return 0;
}
