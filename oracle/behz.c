/*
 * behz.c -- BFV multiplication (BEHZ) and the RNS tool it relies on.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates RNSTool (src/utils/rns.cpp:581-1037) and Evaluator::bfvMultiply
 * (src/evaluator.cpp:385-624).  Every base conversion is the exact integer formula
 *   conv_{I->p}(x) = sum_j [x_j * Ihat_j^-1]_{I_j} * (Ihat_j mod p)  mod p
 * (BaseConverter::fastConvertArray, src/utils/rns.cpp:408-449), computed with
 * 128-bit products, so the words equal the reference's.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static int bits_of_product(const uint64_t *q, size_t L) {
    /* significant bit count of prod(q) (getSignificantBitCountUint of baseProd) */
    uint64_t acc[70];
    size_t words = 1;
    memset(acc, 0, sizeof(acc));
    acc[0] = 1;
    for (size_t i = 0; i < L; i++) {
        uint64_t carry = 0;
        for (size_t w = 0; w < words; w++) {
            u128 p = (u128)acc[w] * q[i] + carry;
            acc[w] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
        if (carry) acc[words++] = carry;
    }
    int top = 63;
    while (top >= 0 && !((acc[words - 1] >> top) & 1)) top--;
    return (int)(64 * (words - 1)) + top + 1;
}

static int bitlen(uint64_t v) { int b = 0; while (v) { b++; v >>= 1; } return b; }

/* prod_{k != skip} base[k] mod p  (skip = (size_t)-1 for the full product) */
static uint64_t prod_mod(const uint64_t *base, size_t n, size_t skip, uint64_t p) {
    uint64_t r = 1 % p;
    for (size_t k = 0; k < n; k++)
        if (k != skip) r = (uint64_t)((u128)r * (base[k] % p) % p);
    return r;
}

int or_behz_bases(int logn, const uint64_t *q, size_t L, uint64_t t, uint64_t *m_sk, uint64_t *gamma, uint64_t *B,
                  size_t *B_size) {
    /* RNSTool::initialize (src/utils/rns.cpp:587-646): |B| = |q|, plus one when
       32 + bits(t) + bits(prod q) >= 61 |q| + 61; primes from getPrimes(2N, 61, |B| + 2):
       m_sk first, gamma second, then B.                                                */
    size_t bsz = L;
    if (32 + bitlen(t) + bits_of_product(q, L) >= 61 * (int)L + 61) bsz++;
    uint64_t primes[70];
    if (or_get_primes((uint64_t)2 << logn, 61, bsz + 2, primes)) return -1;
    *m_sk = primes[0];
    *gamma = primes[1];
    for (size_t i = 0; i < bsz; i++) B[i] = primes[2 + i];
    *B_size = bsz;
    return 0;
}

/* BaseConverter (src/utils/rns.cpp:551-573): Ihat_j^-1 mod I_j and (Ihat_j mod p) */
typedef struct {
    size_t ni, no;
    uint64_t in[70], out[70], inv[70];
    uint64_t mat[70][70];
} conv_t;

static void conv_init(conv_t *c, const uint64_t *ibase, size_t ni, const uint64_t *obase, size_t no) {
    c->ni = ni;
    c->no = no;
    memcpy(c->in, ibase, ni * 8);
    memcpy(c->out, obase, no * 8);
    for (size_t j = 0; j < ni; j++) c->inv[j] = or_invmod(prod_mod(ibase, ni, j, ibase[j]), ibase[j]);
    for (size_t i = 0; i < no; i++)
        for (size_t j = 0; j < ni; j++) c->mat[i][j] = prod_mod(ibase, ni, j, obase[i]);
}

static void conv_apply(const conv_t *c, const uint64_t *x, uint64_t *out) {
    uint64_t y[70];
    for (size_t j = 0; j < c->ni; j++) y[j] = (uint64_t)((u128)x[j] * c->inv[j] % c->in[j]);
    for (size_t i = 0; i < c->no; i++) {
        u128 acc = 0;
        for (size_t j = 0; j < c->ni; j++) acc = (acc + (u128)y[j] * c->mat[i][j]) % c->out[i];
        out[i] = (uint64_t)acc;
    }
}

typedef struct {
    size_t L, bsz, bsk;        /* |q|, |B|, |Bsk| = |B| + 1 */
    uint64_t q[64], B[66], Bsk[67], m_sk, gamma, m_tilde, t;
    conv_t q_bskm, q_bsk, B_q, B_msk;
    uint64_t q_mod_bsk[67], inv_q_bsk[67], inv_mt_bsk[67], neg_inv_q_mt, inv_prodB_msk, prodB_q[64];
} behz_t;

static int behz_init(behz_t *b, int logn, const uint64_t *q, size_t L, uint64_t t) {
    memset(b, 0, sizeof(*b));
    b->L = L;
    b->t = t;
    memcpy(b->q, q, L * 8);
    if (or_behz_bases(logn, q, L, t, &b->m_sk, &b->gamma, b->B, &b->bsz)) return -1;
    memcpy(b->Bsk, b->B, b->bsz * 8);
    b->Bsk[b->bsz] = b->m_sk;
    b->bsk = b->bsz + 1;
    b->m_tilde = (uint64_t)1 << 32;
    uint64_t obase[68];
    memcpy(obase, b->Bsk, b->bsk * 8);
    obase[b->bsk] = b->m_tilde;
    conv_init(&b->q_bskm, q, L, obase, b->bsk + 1);
    conv_init(&b->q_bsk, q, L, b->Bsk, b->bsk);
    conv_init(&b->B_q, b->B, b->bsz, q, L);
    conv_init(&b->B_msk, b->B, b->bsz, &b->m_sk, 1);
    for (size_t i = 0; i < b->bsk; i++) {
        uint64_t p = b->Bsk[i];
        b->q_mod_bsk[i] = prod_mod(q, L, (size_t)-1, p);                    /* prod_q_mod_Bsk_      */
        b->inv_q_bsk[i] = or_invmod(b->q_mod_bsk[i], p);                    /* inv_prod_q_mod_Bsk_  */
        b->inv_mt_bsk[i] = or_invmod(b->m_tilde % p, p);                    /* inv_m_tilde_mod_Bsk_ */
    }
    uint64_t mt = b->m_tilde;
    b->neg_inv_q_mt = (mt - or_invmod(prod_mod(q, L, (size_t)-1, mt), mt)) % mt; /* neg_inv_prod_q_mod_m_tilde_ */
    b->inv_prodB_msk = or_invmod(prod_mod(b->B, b->bsz, (size_t)-1, b->m_sk), b->m_sk);
    for (size_t i = 0; i < L; i++) b->prodB_q[i] = prod_mod(b->B, b->bsz, (size_t)-1, q[i]);
    return 0;
}

/* (1)-(2): fastbconvmTilde then smMrq (rns.cpp:939-960, 886-918): one coefficient
   x (base q) -> base Bsk, with the q-overflow removed by Montgomery reduction. */
static void extend_to_bsk(const behz_t *b, const uint64_t *x, uint64_t *out) {
    uint64_t xm[64], conv[68];
    for (size_t j = 0; j < b->L; j++) xm[j] = (uint64_t)((u128)x[j] * (b->m_tilde % b->q[j]) % b->q[j]);
    conv_apply(&b->q_bskm, xm, conv);
    const uint64_t mt = b->m_tilde;
    const uint64_t r = (uint64_t)((u128)conv[b->bsk] * b->neg_inv_q_mt % mt);
    for (size_t i = 0; i < b->bsk; i++) {
        uint64_t p = b->Bsk[i];
        /* centred r: r - m_tilde when r >= m_tilde / 2 */
        uint64_t rc = (r >= (mt >> 1)) ? (r + p - mt) : r;
        uint64_t v = (uint64_t)(((u128)rc * b->q_mod_bsk[i] + conv[i]) % p);
        out[i] = (uint64_t)((u128)v * b->inv_mt_bsk[i] % p);
    }
}

/* (7): fastFloor (rns.cpp:920-937): x in q u Bsk -> floor(x / q) in Bsk */
static void fast_floor(const behz_t *b, const uint64_t *xq, const uint64_t *xbsk, uint64_t *out) {
    uint64_t conv[68];
    conv_apply(&b->q_bsk, xq, conv);
    for (size_t i = 0; i < b->bsk; i++) {
        uint64_t p = b->Bsk[i];
        uint64_t d = (xbsk[i] + p - conv[i]) % p;
        out[i] = (uint64_t)((u128)d * b->inv_q_bsk[i] % p);
    }
}

/* (8): fastbconvSk (rns.cpp:849-884): Bsk -> q with the Shenoy-Kumaresan correction */
static void bconv_sk(const behz_t *b, const uint64_t *xbsk, uint64_t *out) {
    const uint64_t msk = b->m_sk;
    uint64_t tmp;
    conv_apply(&b->B_q, xbsk, out);
    conv_apply(&b->B_msk, xbsk, &tmp);
    uint64_t alpha = (uint64_t)((u128)((tmp + msk - xbsk[b->bsz]) % msk) * b->inv_prodB_msk % msk);
    for (size_t i = 0; i < b->L; i++) {
        uint64_t qi = b->q[i], pb = b->prodB_q[i];
        if (alpha > (msk >> 1)) {
            uint64_t na = alpha ? msk - alpha : 0;
            out[i] = (uint64_t)(((u128)(na % qi) * pb + out[i]) % qi);
        } else {
            out[i] = (uint64_t)(((u128)(alpha % qi) * ((qi - pb) % qi) + out[i]) % qi);
        }
    }
}

int or_bfv_multiply(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *a, const uint64_t *b_ct,
                    uint64_t *out) {
    /* Evaluator::bfvMultiply (src/evaluator.cpp:385-624) for two size-2 ciphertexts in
       coefficient form, [2][L][N] each; out: [3][L][N]. */
    behz_t *B = malloc(sizeof(behz_t));
    if (behz_init(B, logn, q, L, t)) { free(B); return -1; }
    size_t n = (size_t)1 << logn, bsk = B->bsk;
    or_ntt_tables *tq = calloc(L, sizeof(or_ntt_tables)), *tb = calloc(bsk, sizeof(or_ntt_tables));
    for (size_t i = 0; i < L; i++) or_ntt_tables_init(&tq[i], logn, q[i]);
    for (size_t i = 0; i < bsk; i++) or_ntt_tables_init(&tb[i], logn, B->Bsk[i]);
    /* [ct][poly][limb][n] in base q and base Bsk */
    uint64_t *Aq = malloc(2 * 2 * L * n * 8), *Ab = malloc(2 * 2 * bsk * n * 8);
    const uint64_t *src[2] = {a, b_ct};
    uint64_t x[64], y[68];
    for (int c = 0; c < 2; c++)
        for (int p = 0; p < 2; p++) {
            const uint64_t *poly = src[c] + p * L * n;
            uint64_t *oq = Aq + (c * 2 + p) * L * n, *ob = Ab + (c * 2 + p) * bsk * n;
            memcpy(oq, poly, L * n * 8);
            for (size_t k = 0; k < n; k++) {
                for (size_t j = 0; j < L; j++) x[j] = poly[j * n + k];
                extend_to_bsk(B, x, y);
                for (size_t i = 0; i < bsk; i++) ob[i * n + k] = y[i];
            }
            for (size_t j = 0; j < L; j++) or_ntt_forward(&tq[j], oq + j * n);
            for (size_t i = 0; i < bsk; i++) or_ntt_forward(&tb[i], ob + i * n);
        }
    /* (4) tensor in both bases, (5) back to coefficients */
    uint64_t *Tq = calloc(3 * L * n, 8), *Tb = calloc(3 * bsk * n, 8);
    for (int base = 0; base < 2; base++) {
        size_t nl = base ? bsk : L;
        const uint64_t *m = base ? B->Bsk : q;
        uint64_t *A = base ? Ab : Aq, *T = base ? Tb : Tq;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (size_t l = 0; l < nl; l++) {
                    const uint64_t *u = A + (0 * 2 + i) * nl * n + l * n, *v = A + (1 * 2 + j) * nl * n + l * n;
                    uint64_t *o = T + (i + j) * nl * n + l * n;
                    for (size_t k = 0; k < n; k++) o[k] = (uint64_t)(((u128)u[k] * v[k] + o[k]) % m[l]);
                }
        for (int d = 0; d < 3; d++)
            for (size_t l = 0; l < nl; l++) or_ntt_inverse(base ? &tb[l] : &tq[l], T + d * nl * n + l * n);
    }
    /* (6)-(8): times t, divide-and-floor by q into Bsk, Shenoy-Kumaresan back to q */
    uint64_t xb[68], fl[68], res[64];
    for (int d = 0; d < 3; d++)
        for (size_t k = 0; k < n; k++) {
            for (size_t j = 0; j < L; j++) x[j] = (uint64_t)((u128)Tq[(d * L + j) * n + k] * (t % q[j]) % q[j]);
            for (size_t i = 0; i < bsk; i++)
                xb[i] = (uint64_t)((u128)Tb[(d * bsk + i) * n + k] * (t % B->Bsk[i]) % B->Bsk[i]);
            fast_floor(B, x, xb, fl);
            bconv_sk(B, fl, res);
            for (size_t j = 0; j < L; j++) out[(d * L + j) * n + k] = res[j];
        }
    for (size_t i = 0; i < L; i++) or_ntt_tables_free(&tq[i]);
    for (size_t i = 0; i < bsk; i++) or_ntt_tables_free(&tb[i]);
    free(tq); free(tb); free(Aq); free(Ab); free(Tq); free(Tb); free(B);
    return 0;
}
