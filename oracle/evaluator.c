/*
 * evaluator.c -- tensor product, key switching, rescale / modulus switch.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Layouts: ciphertext [poly][limb][N]; key-switching key [dnum][2][K][N] where
 * dnum = K-1 digits (one per data prime at the first level), each a size-2
 * ciphertext at key level in NTT form (KSwitchKeys::data(), src/kswitchkeys.h).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static uint64_t addm(uint64_t a, uint64_t b, uint64_t q) { uint64_t s = a + b; return s >= q ? s - q : s; }
static uint64_t subm(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

int or_ckks_multiply(int logn, const uint64_t *moduli, size_t L, const uint64_t *a, size_t a_size,
                     const uint64_t *b, size_t b_size, uint64_t *out) {
    /* Evaluator::ckksMultiply (src/evaluator.cpp:626-702): out_k = sum_{i+j=k} a_i * b_j,
       dyadic in NTT form, mod q_limb. */
    size_t n = (size_t)1 << logn, d = L * n, dest = a_size + b_size - 1;
    memset(out, 0, dest * d * 8);
    for (size_t i = 0; i < a_size; i++)
        for (size_t j = 0; j < b_size; j++)
            for (size_t l = 0; l < L; l++) {
                uint64_t q = moduli[l];
                const uint64_t *x = a + i * d + l * n, *y = b + j * d + l * n;
                uint64_t *o = out + (i + j) * d + l * n;
                for (size_t c = 0; c < n; c++) o[c] = addm(o[c], or_mulmod(x[c], y[c], q), q);
            }
    return 0;
}

int or_switch_key(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, const uint64_t *target,
                  const uint64_t *key, uint64_t *ct) {
    /* Evaluator::switchKeyInplace (src/evaluator.cpp:2310-2653), CKKS and BFV branches.
       target: L limbs (CKKS: NTT form; BFV: coefficient form).  ct: 2 polys x L limbs,
       updated in place: ct_k += ModDown(sum_j ModUp(d_j) * key[j][k]).              */
    size_t n = (size_t)1 << logn, R = L + 1;
    const uint64_t qk = key_moduli[K - 1], qk_half = qk >> 1;
    or_ntt_tables *tabs = calloc(K, sizeof(or_ntt_tables));
    uint64_t *t_target = malloc(L * n * 8);
    uint64_t *prod = calloc(2 * R * n, 8);   /* [k][R][n] */
    uint64_t *t_ntt = malloc(n * 8);
    u128 *acc = malloc(2 * n * sizeof(u128));
    int rc = 0;
    for (size_t i = 0; i < K; i++)
        if (or_ntt_tables_init(&tabs[i], logn, key_moduli[i])) { rc = -1; goto done; }
    memcpy(t_target, target, L * n * 8);
    if (scheme == 2) /* CKKS: digits are the coefficient form of the target */
        for (size_t j = 0; j < L; j++) or_ntt_inverse(&tabs[j], t_target + j * n);

    for (size_t i = 0; i < R; i++) {
        size_t ki = (i == L) ? K - 1 : i;
        uint64_t qi = key_moduli[ki];
        memset(acc, 0, 2 * n * sizeof(u128));
        for (size_t j = 0; j < L; j++) {
            /* operand = NTT_i(d_j mod q_i); CKKS reuses the NTT-form input when i == j */
            if (scheme == 2 && i == j) {
                memcpy(t_ntt, target + j * n, n * 8);
            } else {
                for (size_t c = 0; c < n; c++) t_ntt[c] = t_target[j * n + c] % qi;
                or_ntt_forward(&tabs[ki], t_ntt);
            }
            for (size_t k = 0; k < 2; k++) {
                const uint64_t *kp = key + ((j * 2 + k) * K + ki) * n;
                for (size_t c = 0; c < n; c++) acc[k * n + c] += (u128)t_ntt[c] * kp[c];
            }
        }
        for (size_t k = 0; k < 2; k++)
            for (size_t c = 0; c < n; c++) prod[(k * R + i) * n + c] = (uint64_t)(acc[k * n + c] % qi);
    }

    /* ModDown by the special prime with rounding (src/evaluator.cpp:2584-2648) */
    for (size_t k = 0; k < 2; k++) {
        uint64_t *t_last = prod + (k * R + L) * n;
        or_ntt_inverse(&tabs[K - 1], t_last);
        for (size_t c = 0; c < n; c++) t_last[c] = addm(t_last[c], qk_half % qk, qk);
        for (size_t j = 0; j < L; j++) {
            uint64_t qi = key_moduli[j];
            uint64_t half_mod = qk_half % qi;
            uint64_t inv_qk = or_invmod(qk % qi, qi);
            uint64_t *p = prod + (k * R + j) * n;
            for (size_t c = 0; c < n; c++) t_ntt[c] = subm(t_last[c] % qi, half_mod, qi);
            if (scheme == 2) or_ntt_forward(&tabs[j], t_ntt);
            else if (scheme == 1) or_ntt_inverse(&tabs[j], p);
            uint64_t *dst = ct + (k * L + j) * n;
            for (size_t c = 0; c < n; c++) {
                uint64_t v = or_mulmod(subm(p[c], t_ntt[c], qi), inv_qk, qi);
                dst[c] = addm(dst[c], v, qi);
            }
        }
    }
done:
    for (size_t i = 0; i < K; i++) or_ntt_tables_free(&tabs[i]);
    free(tabs); free(t_target); free(prod); free(t_ntt); free(acc);
    return rc;
}

int or_relinearize(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, uint64_t *ct3,
                   const uint64_t *relin_key) {
    /* Evaluator::relinearizeInternal (src/evaluator.cpp:1113-1163) for a size-3 input:
       switch c2 with relin key index 0, drop c2.  ct3 is [3][L][N]; on return its
       first two polys hold the size-2 result. */
    size_t n = (size_t)1 << logn;
    return or_switch_key(scheme, logn, key_moduli, K, L, ct3 + 2 * L * n, relin_key, ct3);
}

int or_apply_galois_keyswitch(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, uint64_t *ct,
                              uint32_t elt, const uint64_t *galois_key) {
    /* Evaluator::applyGaloisInplace (src/evaluator.cpp:2210-2250): c0 <- sigma(c0),
       target <- sigma(c1), c1 <- 0, switchKeyInplace(target, galois key).  ct [2][L][N]. */
    size_t n = (size_t)1 << logn;
    uint64_t *tmp = malloc(L * n * 8), *target = malloc(L * n * 8);
    for (size_t j = 0; j < L; j++) {
        if (scheme == 2) {
            or_apply_galois_ntt(logn, elt, ct + j * n, tmp + j * n);
            or_apply_galois_ntt(logn, elt, ct + (L + j) * n, target + j * n);
        } else {
            or_apply_galois(logn, elt, key_moduli[j], ct + j * n, tmp + j * n);
            or_apply_galois(logn, elt, key_moduli[j], ct + (L + j) * n, target + j * n);
        }
    }
    memcpy(ct, tmp, L * n * 8);
    memset(ct + L * n, 0, L * n * 8);
    int rc = or_switch_key(scheme, logn, key_moduli, K, L, target, galois_key, ct);
    free(tmp);
    free(target);
    return rc;
}

int or_rescale_to_next(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *in, size_t polys,
                       uint64_t *out) {
    /* Evaluator::modSwitchScaleToNext (src/evaluator.cpp:1165-1240) via
       RNSTool::divideAndRoundqLastNttInplace (CKKS, src/utils/rns.cpp:838-877) or
       divideAndRoundqLastInplace (BFV, src/utils/rns.cpp:805-836):
       c_i <- (c_i - ([c_L + floor(q_L/2)]_{q_L} - floor(q_L/2))) * q_L^-1 mod q_i.
       in: [polys][L][N], out: [polys][L-1][N]. */
    size_t n = (size_t)1 << logn;
    if (L < 2) return -1;
    const uint64_t qL = moduli[L - 1], half = qL >> 1;
    or_ntt_tables tl, ti;
    uint64_t *last = malloc(n * 8), *tmp = malloc(n * 8);
    if (or_ntt_tables_init(&tl, logn, qL)) return -1;
    for (size_t p = 0; p < polys; p++) {
        const uint64_t *src = in + p * L * n;
        memcpy(last, src + (L - 1) * n, n * 8);
        if (scheme == 2) or_ntt_inverse(&tl, last);
        for (size_t c = 0; c < n; c++) last[c] = addm(last[c], half, qL);
        for (size_t i = 0; i + 1 < L; i++) {
            uint64_t qi = moduli[i];
            uint64_t half_mod = half % qi, inv = or_invmod(qL % qi, qi);
            for (size_t c = 0; c < n; c++) tmp[c] = subm(last[c] % qi, half_mod, qi);
            if (scheme == 2) {
                or_ntt_tables_init(&ti, logn, qi);
                or_ntt_forward(&ti, tmp);
                or_ntt_tables_free(&ti);
            }
            uint64_t *dst = out + (p * (L - 1) + i) * n;
            for (size_t c = 0; c < n; c++) dst[c] = or_mulmod(subm(src[i * n + c], tmp[c], qi), inv, qi);
        }
    }
    or_ntt_tables_free(&tl);
    free(last); free(tmp);
    return 0;
}

int or_multiply_plain_ntt(int logn, const uint64_t *moduli, size_t L, const uint64_t *a, size_t polys,
                          const uint64_t *plain, uint64_t *out) {
    /* Evaluator::multiplyPlainNtt (src/evaluator.cpp:1932-1990): dyadic product of every
       NTT-form polynomial with the NTT-form plaintext, limb by limb. */
    size_t n = (size_t)1 << logn;
    for (size_t p = 0; p < polys; p++)
        for (size_t i = 0; i < L; i++)
            for (size_t c = 0; c < n; c++)
                out[(p * L + i) * n + c] = or_mulmod(a[(p * L + i) * n + c], plain[i * n + c], moduli[i]);
    return 0;
}

int or_bfv_multiply_plain(int logn, const uint64_t *moduli, size_t L, uint64_t t, const uint64_t *a, size_t polys,
                          const uint64_t *plain, uint64_t *out) {
    /* EvaluatorCuda::multiplyPlainNormal (src/evaluator_cuda.cu:1726-1807): lift m to
       m - t when m >= ceil(t/2) (gMultiplyPlainNormalUtilA/B), NTT both operands, dyadic
       product, inverse NTT; fully reduced.  Coefficient-form a: [polys][L][N]. */
    size_t n = (size_t)1 << logn;
    uint64_t *lift = malloc(n * 8);
    const uint64_t thr = (t + 1) >> 1;
    or_ntt_tables tb;
    for (size_t i = 0; i < L; i++) {
        uint64_t q = moduli[i];
        if (or_ntt_tables_init(&tb, logn, q)) { free(lift); return -1; }
        for (size_t c = 0; c < n; c++) {
            uint64_t m = plain[c] % q;
            lift[c] = plain[c] >= thr ? subm(m, t % q, q) : m;
        }
        or_ntt_forward(&tb, lift);
        for (size_t p = 0; p < polys; p++) {
            uint64_t *dst = out + (p * L + i) * n;
            memcpy(dst, a + (p * L + i) * n, n * 8);
            or_ntt_forward(&tb, dst);
            for (size_t c = 0; c < n; c++) dst[c] = or_mulmod(dst[c], lift[c], q);
            or_ntt_inverse(&tb, dst);
        }
        or_ntt_tables_free(&tb);
    }
    free(lift);
    return 0;
}
