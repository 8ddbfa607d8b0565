/*
 * ntt.c -- negacyclic NTT tables and transforms, Galois automorphisms.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

int or_ntt_tables_init(or_ntt_tables *t, int logn, uint64_t q) {
    /* NTTTables::initialize (utils/ntt.cpp:17-66) */
    memset(t, 0, sizeof(*t));
    t->logn = logn;
    t->n = (uint64_t)1 << logn;
    t->q = q;
    if (or_minimal_primitive_root(2 * t->n, q, &t->root)) return -1;
    t->inv_root = or_invmod(t->root, q);
    size_t n = t->n;
    t->root_powers = malloc(n * 8);
    t->root_quot = malloc(n * 8);
    t->inv_root_powers = malloc(n * 8);
    t->inv_root_quot = malloc(n * 8);
    uint64_t p = t->root;
    for (size_t i = 1; i < n; i++) {
        t->root_powers[or_reverse_bits((uint32_t)i, logn)] = p;
        p = or_mulmod(p, t->root, q);
    }
    t->root_powers[0] = 1;
    p = t->inv_root;
    for (size_t i = 1; i < n; i++) {
        t->inv_root_powers[or_reverse_bits((uint32_t)(i - 1), logn) + 1] = p;
        p = or_mulmod(p, t->inv_root, q);
    }
    t->inv_root_powers[0] = 1;
    for (size_t i = 0; i < n; i++) {
        t->root_quot[i] = or_shoup(t->root_powers[i], q);
        t->inv_root_quot[i] = or_shoup(t->inv_root_powers[i], q);
    }
    t->inv_n = or_invmod(n % q, q);
    t->inv_n_quot = or_shoup(t->inv_n, q);
    return 0;
}

void or_ntt_tables_free(or_ntt_tables *t) {
    free(t->root_powers);
    free(t->root_quot);
    free(t->inv_root_powers);
    free(t->inv_root_quot);
    memset(t, 0, sizeof(*t));
}

void or_ntt_forward(const or_ntt_tables *t, uint64_t *a) {
    /* DWTHandler::transformToRev (utils/dwthandler.h:88-204): Cooley-Tukey layers
       m = 1, 2, ..., n/2 with root root_powers[m + i] for group i; the reference's
       lazy ranges are replaced by full reduction, which leaves the result mod q
       unchanged.                                                                  */
    const uint64_t q = t->q;
    size_t n = t->n, gap = n >> 1;
    size_t ridx = 0;
    for (size_t m = 1; m < n; m <<= 1, gap >>= 1) {
        for (size_t i = 0; i < m; i++) {
            uint64_t r = t->root_powers[++ridx];
            uint64_t *x = a + 2 * gap * i, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                uint64_t u = x[j], v = or_mulmod(y[j], r, q);
                uint64_t s = u + v; if (s >= q) s -= q;
                uint64_t d = u + q - v; if (d >= q) d -= q;
                x[j] = s;
                y[j] = d;
            }
        }
    }
}

void or_ntt_inverse(const or_ntt_tables *t, uint64_t *a) {
    /* DWTHandler::transformFromRev (utils/dwthandler.h:215-372): Gentleman-Sande
       layers m = n/2, ..., 1 with inv_root_powers read sequentially, then n^-1
       (the reference folds n^-1 into the last layer; the product mod q is equal). */
    const uint64_t q = t->q;
    size_t n = t->n, gap = 1;
    size_t ridx = 0;
    for (size_t m = n >> 1; m >= 1; m >>= 1, gap <<= 1) {
        for (size_t i = 0; i < m; i++) {
            uint64_t r = t->inv_root_powers[++ridx];
            uint64_t *x = a + 2 * gap * i, *y = x + gap;
            for (size_t j = 0; j < gap; j++) {
                uint64_t u = x[j], v = y[j];
                uint64_t s = u + v; if (s >= q) s -= q;
                uint64_t d = u + q - v; if (d >= q) d -= q;
                x[j] = s;
                y[j] = or_mulmod(d, r, q);
            }
        }
        if (m == 1) break;
    }
    for (size_t j = 0; j < n; j++) a[j] = or_mulmod(a[j], t->inv_n, q);
}

int or_ntt_batch(int logn, uint64_t q, uint64_t *data, size_t count, int inverse) {
    or_ntt_tables t;
    if (or_ntt_tables_init(&t, logn, q)) return -1;
    for (size_t c = 0; c < count; c++) {
        if (inverse) or_ntt_inverse(&t, data + c * t.n);
        else or_ntt_forward(&t, data + c * t.n);
    }
    or_ntt_tables_free(&t);
    return 0;
}

uint32_t or_galois_elt_from_step(int logn, int step) {
    /* GaloisTool::getEltFromStep (utils/galois.cpp:44-86): 3^step mod 2N; a negative
       step s maps to N/2 - |s|; step 0 gives the conjugation element 2N - 1.       */
    uint64_t n = (uint64_t)1 << logn, m = 2 * n;
    if (step == 0) return (uint32_t)(m - 1);
    int neg = step < 0;
    uint64_t pos = (uint64_t)(neg ? -step : step);
    if (pos >= (n >> 1)) return 0;
    uint64_t e = neg ? (n >> 1) - pos : pos;
    uint64_t g = 1;
    while (e--) g = (g * 3) & (m - 1);
    return (uint32_t)g;
}

void or_apply_galois(int logn, uint32_t elt, uint64_t q, const uint64_t *in, uint64_t *out) {
    /* GaloisTool::applyGalois (utils/galois.cpp:143-162): X^i -> (-1)^{floor(i elt / N)} X^{i elt mod N} */
    uint64_t n = (uint64_t)1 << logn;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t raw = i * elt;
        uint64_t idx = raw & (n - 1);
        uint64_t v = in[i];
        if ((raw >> logn) & 1) v = v ? q - v : 0;
        out[idx] = v;
    }
}

void or_apply_galois_ntt(int logn, uint32_t elt, const uint64_t *in, uint64_t *out) {
    /* GaloisTool::generateTableNtt + applyGaloisNtt (utils/galois.cpp:13-30, 164-175) */
    uint64_t n = (uint64_t)1 << logn;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t rev = or_reverse_bits((uint32_t)(n + i), logn + 1);
        uint64_t raw = ((uint64_t)elt * rev) >> 1;
        raw &= n - 1;
        out[i] = in[or_reverse_bits((uint32_t)raw, logn)];
    }
}

int or_ntt_tables_export(int logn, uint64_t q, uint64_t *rp, uint64_t *rq, uint64_t *irp, uint64_t *irq,
                         uint64_t *inv_n) {
    or_ntt_tables t;
    if (or_ntt_tables_init(&t, logn, q)) return -1;
    memcpy(rp, t.root_powers, t.n * 8);
    memcpy(rq, t.root_quot, t.n * 8);
    memcpy(irp, t.inv_root_powers, t.n * 8);
    memcpy(irq, t.inv_root_quot, t.n * 8);
    inv_n[0] = t.inv_n;
    inv_n[1] = t.inv_n_quot;
    or_ntt_tables_free(&t);
    return 0;
}
