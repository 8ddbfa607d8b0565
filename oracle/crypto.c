/*
 * crypto.c -- encryption, decryption, key-switching key generation and the
 * BatchEncoder, restated on explicit randomness.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference draws u, e and a from Blake2xb streams (CPU) or curand (GPU); those
 * streams cannot be reproduced bit for bit, so these functions take the sampled
 * polynomials as inputs (coefficient form, one integer per coefficient reduced into
 * every limb) and restate the arithmetic that follows the sampling.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static uint64_t addm(uint64_t a, uint64_t b, uint64_t q) { uint64_t s = a + b; return s >= q ? s - q : s; }
static uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

int or_encrypt_zero_asymmetric(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *pk,
                               size_t pk_limbs, const uint64_t *u, const uint64_t *e0, const uint64_t *e1,
                               uint64_t *out) {
    /* encryptZeroAsymmetric (src/utils/rlwe.cpp:170-237): c_j = NTT(u) * pk_j (NTT form),
       back to coefficients for BFV, plus e_j (NTT'd for CKKS).  Level with L primes
       moduli[0..L); pk has pk_limbs limbs per poly (key level), of which the first L
       are used.  out: [2][L][N]. */
    size_t n = (size_t)1 << logn;
    or_ntt_tables t;
    uint64_t *un = malloc(n * 8), *e = malloc(n * 8);
    const uint64_t *es[2] = {e0, e1};
    for (size_t i = 0; i < L; i++) {
        uint64_t q = moduli[i];
        if (or_ntt_tables_init(&t, logn, q)) return -1;
        memcpy(un, u + i * n, n * 8);
        or_ntt_forward(&t, un);
        for (int j = 0; j < 2; j++) {
            uint64_t *o = out + (j * L + i) * n;
            const uint64_t *p = pk + (j * pk_limbs + i) * n;
            for (size_t c = 0; c < n; c++) o[c] = or_mulmod(un[c], p[c], q);
            if (scheme != 2) or_ntt_inverse(&t, o);
            memcpy(e, es[j] + i * n, n * 8);
            if (scheme == 2) or_ntt_forward(&t, e);
            for (size_t c = 0; c < n; c++) o[c] = addm(o[c], e[c], q);
        }
        or_ntt_tables_free(&t);
    }
    free(un);
    free(e);
    return 0;
}

int or_encrypt_zero_symmetric(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *sk,
                              size_t sk_limbs, const uint64_t *a, const uint64_t *e, uint64_t *out) {
    /* encryptZeroSymmetric (src/utils/rlwe.cpp:239-331): c1 = a (sampled as NTT form),
       c0 = -(a * s + e); for coefficient-form output both are inverse transformed.
       sk: NTT form, sk_limbs limbs (first L used).  out: [2][L][N]. */
    size_t n = (size_t)1 << logn;
    or_ntt_tables t;
    uint64_t *en = malloc(n * 8);
    for (size_t i = 0; i < L; i++) {
        uint64_t q = moduli[i];
        if (or_ntt_tables_init(&t, logn, q)) return -1;
        uint64_t *c0 = out + i * n, *c1 = out + (L + i) * n;
        memcpy(c1, a + i * n, n * 8);
        for (size_t c = 0; c < n; c++) c0[c] = or_mulmod(sk[i * n + c], c1[c], q);
        memcpy(en, e + i * n, n * 8);
        if (scheme == 2) or_ntt_forward(&t, en);
        else or_ntt_inverse(&t, c0);
        for (size_t c = 0; c < n; c++) {
            uint64_t v = addm(c0[c], en[c], q);
            c0[c] = v ? q - v : 0;
        }
        if (scheme != 2) or_ntt_inverse(&t, c1);
        or_ntt_tables_free(&t);
    }
    free(en);
    return 0;
}

int or_generate_kswitch_key(int logn, const uint64_t *key_moduli, size_t K, const uint64_t *sk,
                            const uint64_t *new_key, const uint64_t *a, const uint64_t *e, uint64_t *out) {
    /* KeyGenerator::generateOneKswitchKey (src/keygenerator.cpp:294-329): for each digit
       j < K-1, a symmetric encryption of zero at key level (NTT form) with randomness
       a[j], e[j], then limb j of c0 gets + new_key_j * (q_special mod q_j).
       sk, new_key: [K][N] NTT form; a, e: [K-1][K][N]; out: [K-1][2][K][N]. */
    size_t n = (size_t)1 << logn;
    for (size_t j = 0; j + 1 < K; j++) {
        uint64_t *dst = out + j * 2 * K * n;
        if (or_encrypt_zero_symmetric(2, logn, key_moduli, K, sk, K, a + j * K * n, e + j * K * n, dst)) return -1;
        uint64_t q = key_moduli[j], factor = key_moduli[K - 1] % q;
        for (size_t c = 0; c < n; c++)
            dst[j * n + c] = addm(dst[j * n + c], or_mulmod(new_key[j * n + c], factor, q), q);
    }
    return 0;
}

int or_decrypt_phase(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *ct, size_t size,
                     const uint64_t *sk, size_t sk_limbs, uint64_t *out) {
    /* Decryptor::dotProductCtSkArray (src/decryptor.cpp:206-330): sum_i c_i s^i,
       NTT form for CKKS input, coefficient form for BFV.  out: [L][N]. */
    size_t n = (size_t)1 << logn;
    or_ntt_tables t;
    uint64_t *acc = malloc(n * 8), *tmp = malloc(n * 8), *spow = malloc(n * 8);
    for (size_t i = 0; i < L; i++) {
        uint64_t q = moduli[i];
        if (or_ntt_tables_init(&t, logn, q)) return -1;
        memset(acc, 0, n * 8);
        for (size_t c = 0; c < n; c++) spow[c] = 1;
        for (size_t k = 0; k < size; k++) {
            memcpy(tmp, ct + (k * L + i) * n, n * 8);
            if (scheme != 2) or_ntt_forward(&t, tmp);
            for (size_t c = 0; c < n; c++) acc[c] = addm(acc[c], or_mulmod(tmp[c], spow[c], q), q);
            for (size_t c = 0; c < n; c++) spow[c] = or_mulmod(spow[c], sk[i * n + c], q);
        }
        if (scheme != 2) or_ntt_inverse(&t, acc);
        memcpy(out + i * n, acc, n * 8);
        or_ntt_tables_free(&t);
    }
    free(acc); free(tmp); free(spow);
    return 0;
}

static uint64_t prod_mod(const uint64_t *base, size_t n, size_t skip, uint64_t p) {
    uint64_t r = 1 % p;
    for (size_t k = 0; k < n; k++)
        if (k != skip) r = (uint64_t)((u128)r * (base[k] % p) % p);
    return r;
}

int or_bfv_decrypt_scale_round(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *phase,
                               uint64_t *out) {
    /* RNSTool::decryptScaleAndRound (src/utils/rns.cpp:1039-1095): x * (t gamma) mod q,
       fast conversion to {t, gamma}, times -q^-1, centred gamma correction, times
       gamma^-1 mod t.  phase: [L][N] coefficient form; out: [N] mod t. */
    uint64_t msk, gamma, B[70];
    size_t bsz;
    if (or_behz_bases(logn, q, L, t, &msk, &gamma, B, &bsz)) return -1;
    size_t n = (size_t)1 << logn;
    uint64_t tg[2] = {t, gamma};
    uint64_t qhat_inv[64], mat[2][64], neg_inv_q[2];
    for (size_t j = 0; j < L; j++) {
        qhat_inv[j] = or_invmod(prod_mod(q, L, j, q[j]), q[j]);
        for (int k = 0; k < 2; k++) mat[k][j] = prod_mod(q, L, j, tg[k]);
    }
    for (int k = 0; k < 2; k++) {
        uint64_t inv = or_invmod(prod_mod(q, L, (size_t)-1, tg[k]), tg[k]);
        neg_inv_q[k] = inv ? tg[k] - inv : 0;
    }
    uint64_t inv_gamma_t = or_invmod(gamma % t, t);
    for (size_t c = 0; c < n; c++) {
        uint64_t y[64];
        for (size_t j = 0; j < L; j++) {
            uint64_t tgj = (uint64_t)((u128)(t % q[j]) * (gamma % q[j]) % q[j]);
            uint64_t x = (uint64_t)((u128)phase[j * n + c] * tgj % q[j]);
            y[j] = (uint64_t)((u128)x * qhat_inv[j] % q[j]);
        }
        uint64_t r[2];
        for (int k = 0; k < 2; k++) {
            u128 acc = 0;
            for (size_t j = 0; j < L; j++) acc = (acc + (u128)y[j] * mat[k][j]) % tg[k];
            r[k] = (uint64_t)((u128)acc * neg_inv_q[k] % tg[k]);
        }
        uint64_t d;
        if (r[1] > (gamma >> 1)) d = (r[0] + (gamma - r[1]) % t) % t;
        else d = (r[0] + t - r[1] % t) % t;
        if (d) d = (uint64_t)((u128)d * inv_gamma_t % t);
        out[c] = d;
    }
    return 0;
}

static int bfv_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0, int subtract);

int or_bfv_add_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0) {
    return bfv_plain_scaled(logn, q, L, t, plain, plain_count, c0, 0);
}

int or_bfv_sub_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0) {
    return bfv_plain_scaled(logn, q, L, t, plain, plain_count, c0, 1);
}

static int bfv_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0, int subtract) {
    /* multiplyAdd / multiplySubPlainWithScalingVariant (src/utils/scalingvariant.cpp:
       53-131): c0 +-= m * floor(Q/t) + floor((m * (Q mod t) + floor((t+1)/2)) / t)
       (mod q_j). */
    size_t n = (size_t)1 << logn;
    /* Q mod t and floor(Q/t) mod q_j by long division of the product */
    uint64_t prod[70] = {1};
    size_t words = 1;
    for (size_t i = 0; i < L; i++) {
        uint64_t carry = 0;
        for (size_t w = 0; w < words; w++) {
            u128 p = (u128)prod[w] * q[i] + carry;
            prod[w] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        if (carry) prod[words++] = carry;
    }
    uint64_t quot[70];
    u128 rem = 0;
    for (size_t w = words; w-- > 0;) {
        u128 cur = (rem << 64) | prod[w];
        quot[w] = (uint64_t)(cur / t);
        rem = cur % t;
    }
    uint64_t q_mod_t = (uint64_t)rem, delta[64];
    for (size_t j = 0; j < L; j++) {
        u128 r = 0;
        for (size_t w = words; w-- > 0;) r = ((r << 64) | quot[w]) % q[j];
        delta[j] = (uint64_t)r;
    }
    uint64_t half = (t + 1) >> 1;
    for (size_t c = 0; c < plain_count && c < n; c++) {
        uint64_t m = plain[c];
        uint64_t fix = (uint64_t)(((u128)m * q_mod_t + half) / t);
        for (size_t j = 0; j < L; j++) {
            uint64_t v = (uint64_t)(((u128)m * delta[j] + fix) % q[j]);
            c0[j * n + c] = subtract ? subm(c0[j * n + c], v, q[j]) : addm(c0[j * n + c], v, q[j]);
        }
    }
    return 0;
}

void or_batch_index_map(int logn, uint64_t *map) {
    /* BatchEncoder::populateMatrixRepsIndexMap (src/batchencoder.cpp:64-84) */
    uint64_t n = (uint64_t)1 << logn, row = n >> 1, m = n << 1, pos = 1;
    for (uint64_t i = 0; i < row; i++) {
        uint64_t i1 = (pos - 1) >> 1, i2 = (m - pos - 1) >> 1;
        map[i] = or_reverse_bits((uint32_t)i1, logn);
        map[row | i] = or_reverse_bits((uint32_t)i2, logn);
        pos = (pos * 3) & (m - 1);
    }
}

int or_batch_encode(int logn, uint64_t t, const uint64_t *values, size_t count, uint64_t *plain) {
    /* BatchEncoder::encode (src/batchencoder.cpp:104-132): scatter through the index
       map, then the inverse NTT modulo t. */
    size_t n = (size_t)1 << logn;
    uint64_t *map = malloc(n * 8);
    or_batch_index_map(logn, map);
    memset(plain, 0, n * 8);
    for (size_t i = 0; i < count && i < n; i++) plain[map[i]] = values[i];
    free(map);
    return or_ntt_batch(logn, t, plain, 1, 1);
}

int or_batch_decode(int logn, uint64_t t, const uint64_t *plain, uint64_t *values) {
    /* BatchEncoder::decode (src/batchencoder.cpp:167-200) */
    size_t n = (size_t)1 << logn;
    uint64_t *map = malloc(n * 8), *tmp = malloc(n * 8);
    or_batch_index_map(logn, map);
    memcpy(tmp, plain, n * 8);
    int rc = or_ntt_batch(logn, t, tmp, 1, 0);
    for (size_t i = 0; i < n; i++) values[i] = tmp[map[i]];
    free(map);
    free(tmp);
    return rc;
}
