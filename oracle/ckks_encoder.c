/*
 * ckks_encoder.c -- CKKS canonical-embedding encoder.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates CKKSEncoder (src/ckks.cpp:32-84 roots and index map, :98-277 encode,
 * :474-580 decode) and ComplexRoots (src/utils/croots.cpp:15-60) in double precision.
 * Encode rounds real parts to integers, so outputs may differ from the reference's in
 * rare near-half cases; decode is compared within a tolerance.
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef double complex cplx;

static cplx croot(size_t m, size_t index) {
    /* ComplexRoots::getRoot: roots e^{2 pi i k/m} from a table of the first m/8 + 1,
       extended by the symmetries of the unit circle */
    index &= m - 1;
    if (index <= m / 8) {
        double a = 2 * M_PI * (double)index / (double)m;
        return cos(a) + I * sin(a);
    }
    if (index <= m / 4) {
        cplx r = croot(m, m / 4 - index);
        return cimag(r) + I * creal(r);
    }
    if (index <= m / 2) return -conj(croot(m, m / 2 - index));
    if (index <= 3 * m / 4) return -croot(m, index - m / 2);
    return conj(croot(m, m - index));
}

static void roots(int logn, cplx *rp, cplx *irp) {
    size_t n = (size_t)1 << logn, m = n << 1;
    rp[0] = irp[0] = 1;
    if (m >= 8) {
        for (size_t i = 1; i < n; i++) {
            rp[i] = croot(m, or_reverse_bits((uint32_t)i, logn));
            irp[i] = conj(croot(m, or_reverse_bits((uint32_t)(i - 1), logn) + 1));
        }
    } else if (m == 4) {
        rp[1] = I;
        irp[1] = -I;
    }
}

static void fft_from_rev(cplx *a, int logn, const cplx *r, double scalar) {
    /* DWTHandler::transformFromRev with complex arithmetic and the scalar folded into
       the last layer (src/utils/dwthandler.h:215-372) */
    size_t n = (size_t)1 << logn, gap = 1, ridx = 0;
    for (size_t m = n >> 1; m > 1; m >>= 1, gap <<= 1)
        for (size_t i = 0; i < m; i++) {
            cplx w = r[++ridx];
            for (size_t j = 0; j < gap; j++) {
                cplx *x = a + 2 * gap * i + j, *y = x + gap;
                cplx u = *x, v = *y;
                *x = u + v;
                *y = (u - v) * w;
            }
        }
    cplx w = r[++ridx] * scalar;
    for (size_t j = 0; j < gap; j++) {
        cplx u = a[j], v = a[j + gap];
        a[j] = (u + v) * scalar;
        a[j + gap] = (u - v) * w;
    }
}

static void fft_to_rev(cplx *a, int logn, const cplx *r) {
    /* DWTHandler::transformToRev (src/utils/dwthandler.h:88-204) */
    size_t n = (size_t)1 << logn, gap = n >> 1, ridx = 0;
    for (size_t m = 1; m < n; m <<= 1, gap >>= 1)
        for (size_t i = 0; i < m; i++) {
            cplx w = r[++ridx];
            for (size_t j = 0; j < gap; j++) {
                cplx *x = a + 2 * gap * i + j, *y = x + gap;
                cplx u = *x, v = *y * w;
                *x = u + v;
                *y = u - v;
            }
        }
}

int or_ckks_encode(int logn, const uint64_t *moduli, size_t L, const double *values, size_t count, double scale,
                   uint64_t *out) {
    /* CKKSEncoder::encodeInternal (src/ckks.cpp:98-277); values: count (re, im) pairs,
       count <= N/2; out: NTT-form plaintext [L][N].  Coefficients up to 128 bits. */
    size_t n = (size_t)1 << logn, slots = n >> 1;
    if (count > slots) return -1;
    cplx *a = calloc(n, sizeof(cplx)), *rp = malloc(n * sizeof(cplx)), *irp = malloc(n * sizeof(cplx));
    uint64_t *map = malloc(n * 8);
    or_batch_index_map(logn, map);
    roots(logn, rp, irp);
    for (size_t i = 0; i < count; i++) {
        cplx v = values[2 * i] + I * values[2 * i + 1];
        a[map[i]] = v;
        a[map[i + slots]] = conj(v);
    }
    fft_from_rev(a, logn, irp, scale / (double)n);
    const double two64 = 18446744073709551616.0;
    int rc = 0;
    for (size_t c = 0; c < n; c++) {
        double d = round(creal(a[c]));
        int neg = signbit(d);
        d = fabs(d);
        if (d >= two64 * two64) { rc = -2; break; }
        uint64_t hi = (uint64_t)(d / two64), lo = (uint64_t)fmod(d, two64);
        for (size_t j = 0; j < L; j++) {
            uint64_t q = moduli[j];
            uint64_t r = (uint64_t)((((u128)hi << 64) | lo) % q);
            out[j * n + c] = neg ? (r ? q - r : 0) : r;
        }
    }
    for (size_t j = 0; j < L && !rc; j++) or_ntt_batch(logn, moduli[j], out + j * n, 1, 0);
    free(a); free(rp); free(irp); free(map);
    return rc;
}

int or_ckks_decode(int logn, const uint64_t *moduli, size_t L, const uint64_t *plain, double scale, double *values) {
    /* CKKSEncoder::decodeInternal (src/ckks.cpp:474-580): inverse NTT, CRT compose into
       [0, Q), centred lift (x >= (Q+1)/2 is negative), times 1/scale, forward DWT,
       gather; values: N/2 (re, im) pairs. */
    size_t n = (size_t)1 << logn, slots = n >> 1;
    uint64_t *c = malloc(L * n * 8), *map = malloc(n * 8);
    cplx *a = malloc(n * sizeof(cplx)), *rp = malloc(n * sizeof(cplx)), *irp = malloc(n * sizeof(cplx));
    memcpy(c, plain, L * n * 8);
    for (size_t j = 0; j < L; j++) or_ntt_batch(logn, moduli[j], c + j * n, 1, 1);
    /* Q and the punctured products as multi-word integers (L words each suffice) */
    size_t W = L + 1;
    uint64_t *Q = calloc(W, 8), *qhat = calloc(L * W, 8), *acc = malloc(W * 8), *tmp = malloc(W * 8);
    uint64_t *qhat_inv = malloc(L * 8);
    Q[0] = 1;
    for (size_t j = 0; j < L; j++) {
        uint64_t carry = 0;
        for (size_t w = 0; w < W; w++) {
            u128 p = (u128)Q[w] * moduli[j] + carry;
            Q[w] = (uint64_t)p;
            carry = (uint64_t)(p >> 64);
        }
    }
    for (size_t j = 0; j < L; j++) {
        uint64_t *h = qhat + j * W;
        h[0] = 1;
        for (size_t k = 0; k < L; k++) {
            if (k == j) continue;
            uint64_t carry = 0;
            for (size_t w = 0; w < W; w++) {
                u128 p = (u128)h[w] * moduli[k] + carry;
                h[w] = (uint64_t)p;
                carry = (uint64_t)(p >> 64);
            }
        }
        u128 r = 0;
        for (size_t w = W; w-- > 0;) r = ((r << 64) | h[w]) % moduli[j];
        qhat_inv[j] = or_invmod((uint64_t)r, moduli[j]);
    }
    const double two64 = 18446744073709551616.0, inv_scale = 1.0 / scale;
    for (size_t i = 0; i < n; i++) {
        memset(acc, 0, W * 8);
        for (size_t j = 0; j < L; j++) {
            uint64_t y = or_mulmod(c[j * n + i], qhat_inv[j], moduli[j]);
            /* acc += y * qhat_j, then reduce below Q */
            uint64_t carry = 0;
            for (size_t w = 0; w < W; w++) {
                u128 p = (u128)qhat[j * W + w] * y + carry;
                tmp[w] = (uint64_t)p;
                carry = (uint64_t)(p >> 64);
            }
            unsigned char cy = 0;
            for (size_t w = 0; w < W; w++) {
                u128 s = (u128)acc[w] + tmp[w] + cy;
                acc[w] = (uint64_t)s;
                cy = (unsigned char)(s >> 64);
            }
            for (;;) { /* while acc >= Q: acc -= Q */
                int ge = 1;
                for (size_t w = W; w-- > 0;) {
                    if (acc[w] != Q[w]) { ge = acc[w] > Q[w]; break; }
                }
                if (!ge) break;
                unsigned char br = 0;
                for (size_t w = 0; w < W; w++) {
                    u128 d = (u128)acc[w] - Q[w] - br;
                    acc[w] = (uint64_t)d;
                    br = (unsigned char)((d >> 64) != 0);
                }
            }
        }
        /* centred lift: acc >= (Q+1)/2 -> acc - Q */
        int upper = 0;
        {
            uint64_t *h2 = tmp;  /* (Q + 1) / 2 */
            unsigned char cy = 1;
            for (size_t w = 0; w < W; w++) {
                u128 s = (u128)Q[w] + cy;
                h2[w] = (uint64_t)s;
                cy = (unsigned char)(s >> 64);
            }
            for (size_t w = 0; w < W; w++) h2[w] = (h2[w] >> 1) | (w + 1 < W ? h2[w + 1] << 63 : 0);
            upper = 1;
            for (size_t w = W; w-- > 0;) {
                if (acc[w] != h2[w]) { upper = acc[w] > h2[w]; break; }
            }
        }
        double v = 0, s = inv_scale;
        for (size_t w = 0; w < L; w++, s *= two64) {
            if (upper) {
                if (acc[w] > Q[w]) v += (double)(acc[w] - Q[w]) * s;
                else v -= (double)(Q[w] - acc[w]) * s;
            } else {
                v += (double)acc[w] * s;
            }
        }
        a[i] = v;
    }
    or_batch_index_map(logn, map);
    roots(logn, rp, irp);
    fft_to_rev(a, logn, rp);
    for (size_t i = 0; i < slots; i++) {
        values[2 * i] = creal(a[map[i]]);
        values[2 * i + 1] = cimag(a[map[i]]);
    }
    free(c); free(map); free(a); free(rp); free(irp); free(Q); free(qhat); free(acc); free(tmp); free(qhat_inv);
    return 0;
}
