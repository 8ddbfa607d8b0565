/*
 * oracle.h -- CPU restatement of the reference's RNS polynomial core.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker that the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links or calls it.
 *
 * Every function restates the mathematics of the reference CPU path (a SEAL-4.0
 * port under the reference src/ directory) and cites the file:line it follows.  Arithmetic is done
 * with exact 128-bit integers (unsigned __int128) and fully reduced residues: the
 * reference's lazy ranges never change a fully reduced word, so the restatement
 * reproduces the reference's outputs word for word.
 *
 * Layouts follow the reference: a polynomial in RNS form is [limb][coeff], a
 * ciphertext is [poly][limb][coeff], all uint64.
 *
 * Parity status of each group is stated in its section below ("pinned by").
 */
#ifndef SEALHIP_ORACLE_H
#define SEALHIP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- arithmetic
 * Pinned by: closed forms (Barrett ratio floor(2^128/q), Shoup floor(w 2^64/q)),
 * brute force in tests, reference KATs (test/utils/uintarithsmallmod.cpp).     */
uint64_t or_mulmod(uint64_t a, uint64_t b, uint64_t q);
uint64_t or_powmod(uint64_t a, uint64_t e, uint64_t q);
/* returns 0 when not invertible (reference tryInvertUintMod, uintarithsmallmod.h:311) */
uint64_t or_invmod(uint64_t a, uint64_t q);
uint64_t or_shoup(uint64_t w, uint64_t q);              /* floor(w * 2^64 / q) */
void or_barrett_ratio(uint64_t q, uint64_t out[3]);     /* modulus.cpp:15-42 */
uint32_t or_reverse_bits(uint32_t x, int bits);

/* ---------------------------------------------------------------- primes
 * Pinned by: reference KAT test/modulus.cpp (CoeffModTest), BFVDefault table,
 * and the reference library (oracle/_ref) in tests.                            */
int or_is_prime(uint64_t v);
/* getPrimes (utils/numth.cpp:261-283): descending from (2^b-1)/factor*factor+1 */
int or_get_primes(uint64_t factor, int bit_size, size_t count, uint64_t *out);
/* CoeffModulus::Create (modulus.cpp:80-121); plain_modulus 0 = the variant without t */
int or_coeff_modulus_create(size_t n, const int *bit_sizes, size_t count, uint64_t plain_modulus,
                            uint64_t *out);
/* tryMinimalPrimitiveRoot (utils/numth.cpp:335-362) */
int or_minimal_primitive_root(uint64_t degree, uint64_t q, uint64_t *root);

/* ---------------------------------------------------------------- NTT
 * Pinned by: reference KATs (test/utils/ntt.cpp:52-99), O(N^2) negacyclic DFT
 * at psi^(2 bitrev(k)+1), round trip, and oracle/_ref.                        */
typedef struct {
    int logn;
    uint64_t n;
    uint64_t q;
    uint64_t root, inv_root;
    uint64_t *root_powers;       /* [bitrev(i)] = psi^i, [0] = 1           (utils/ntt.cpp:36-44) */
    uint64_t *root_quot;         /* Shoup quotients of root_powers                              */
    uint64_t *inv_root_powers;   /* [bitrev(i-1)+1] = psi^-i, [0] = 1       (utils/ntt.cpp:47-54) */
    uint64_t *inv_root_quot;
    uint64_t inv_n, inv_n_quot;
} or_ntt_tables;

int or_ntt_tables_init(or_ntt_tables *t, int logn, uint64_t q);
void or_ntt_tables_free(or_ntt_tables *t);
/* nttNegacyclicHarvey (utils/ntt.cpp:165-187): natural order in, bit-reversed out, in [0,q) */
void or_ntt_forward(const or_ntt_tables *t, uint64_t *a);
/* inverseNttNegacyclicHarvey (utils/ntt.cpp:195-213): bit-reversed in, natural out, in [0,q) */
void or_ntt_inverse(const or_ntt_tables *t, uint64_t *a);
/* convenience: count polynomials of length 2^logn, all modulo q */
int or_ntt_batch(int logn, uint64_t q, uint64_t *data, size_t count, int inverse);
/* copies of the tables (root powers, their Shoup quotients, n^-1 and its quotient) */
int or_ntt_tables_export(int logn, uint64_t q, uint64_t *rp, uint64_t *rq, uint64_t *irp, uint64_t *irq,
                         uint64_t *inv_n);

/* ---------------------------------------------------------------- Galois
 * Pinned by: reference KATs (test/utils/galois.cpp:26-109) and oracle/_ref.    */
uint32_t or_galois_elt_from_step(int logn, int step);        /* utils/galois.cpp:44-86   */
void or_apply_galois(int logn, uint32_t elt, uint64_t q, const uint64_t *in, uint64_t *out); /* :143-162 */
void or_apply_galois_ntt(int logn, uint32_t elt, const uint64_t *in, uint64_t *out);         /* :13-30, 164-175 */

/* ---------------------------------------------------------------- evaluator (evaluator.c)
 * Pinned by: oracle/_ref (the reference's own Evaluator) on identical injected
 * ciphertexts and keys at N = 2^10..2^13, decrypt-equality with real keys, and the
 * closed forms in the tests (tensor = schoolbook product; rescale = CRT round).     */
/* Evaluator::ckksMultiply (src/evaluator.cpp:626-702); out has a_size + b_size - 1 polys */
int or_ckks_multiply(int logn, const uint64_t *moduli, size_t L, const uint64_t *a, size_t a_size,
                     const uint64_t *b, size_t b_size, uint64_t *out);
/* Evaluator::switchKeyInplace (src/evaluator.cpp:2310-2653); scheme 1 = BFV, 2 = CKKS.
   key: [K-1][2][K][N]; target: L limbs; ct: [2][L][N] updated in place. */
int or_switch_key(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, const uint64_t *target,
                  const uint64_t *key, uint64_t *ct);
/* Evaluator::applyGaloisInplace (src/evaluator.cpp:2210-2250): ct [2][L][N] in place */
int or_apply_galois_keyswitch(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, uint64_t *ct,
                              uint32_t elt, const uint64_t *galois_key);
/* Evaluator::relinearizeInternal (src/evaluator.cpp:1113-1163): ct3 [3][L][N] -> first 2 polys */
int or_relinearize(int scheme, int logn, const uint64_t *key_moduli, size_t K, size_t L, uint64_t *ct3,
                   const uint64_t *relin_key);
/* ---------------------------------------------------------------- BFV (behz.c)
 * Pinned by: oracle/_ref on injected words (BFV multiply words), decrypt-equality. */
/* RNSTool bases (src/utils/rns.cpp:587-646): m_sk, gamma and B for base q of size L */
int or_behz_bases(int logn, const uint64_t *q, size_t L, uint64_t t, uint64_t *m_sk, uint64_t *gamma, uint64_t *B,
                  size_t *B_size);
/* Evaluator::bfvMultiply (src/evaluator.cpp:385-624): a, b [2][L][N] coefficient form -> out [3][L][N] */
int or_bfv_multiply(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *a, const uint64_t *b,
                    uint64_t *out);

/* ---------------------------------------------------------------- encryption (crypto.c)
 * Randomness is an explicit input (see crypto.c).  Pinned by: decrypt-equality with
 * the reference Decryptor (oracle/_ref), the reference Decryptor and BatchEncoder
 * words, closed forms (phase = Delta m + small noise).                             */
int or_encrypt_zero_asymmetric(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *pk,
                               size_t pk_limbs, const uint64_t *u, const uint64_t *e0, const uint64_t *e1,
                               uint64_t *out);
int or_encrypt_zero_symmetric(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *sk,
                              size_t sk_limbs, const uint64_t *a, const uint64_t *e, uint64_t *out);
int or_generate_kswitch_key(int logn, const uint64_t *key_moduli, size_t K, const uint64_t *sk,
                            const uint64_t *new_key, const uint64_t *a, const uint64_t *e, uint64_t *out);
int or_decrypt_phase(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *ct, size_t size,
                     const uint64_t *sk, size_t sk_limbs, uint64_t *out);
int or_bfv_decrypt_scale_round(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *phase,
                               uint64_t *out);
int or_bfv_add_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0);
/* c0 -= Delta m (multiplySubPlainWithScalingVariant, src/utils/scalingvariant.cpp:94-131).
   Pinned by: the reference subPlainInplace on the same words (tests/test_oracle_evaluator.py). */
int or_bfv_sub_plain_scaled(int logn, const uint64_t *q, size_t L, uint64_t t, const uint64_t *plain,
                            size_t plain_count, uint64_t *c0);
void or_batch_index_map(int logn, uint64_t *map);
int or_batch_encode(int logn, uint64_t t, const uint64_t *values, size_t count, uint64_t *plain);
int or_batch_decode(int logn, uint64_t t, const uint64_t *plain, uint64_t *values);

/* ---------------------------------------------------------------- CKKS encoder (ckks_encoder.c)
 * Pinned by: the reference CKKSEncoder (oracle/_ref): encode words equal up to rare
 * +-1 rounding cases, decode within 2^-30 relative; round trip encode -> decode.     */
int or_ckks_encode(int logn, const uint64_t *moduli, size_t L, const double *values, size_t count, double scale,
                   uint64_t *out);
int or_ckks_decode(int logn, const uint64_t *moduli, size_t L, const uint64_t *plain, double scale, double *values);

/* Evaluator::modSwitchScaleToNext (src/evaluator.cpp:1165-1240): in [polys][L][N] -> out [polys][L-1][N] */
int or_rescale_to_next(int scheme, int logn, const uint64_t *moduli, size_t L, const uint64_t *in, size_t polys,
                       uint64_t *out);
/* Evaluator::multiplyPlainNtt (src/evaluator.cpp:1932): [polys][L][N] x plain [L][N].
   Pinned by: the reference multiplyPlainInplace (CKKS, NTT form) on the same words. */
int or_multiply_plain_ntt(int logn, const uint64_t *moduli, size_t L, const uint64_t *a, size_t polys,
                          const uint64_t *plain, uint64_t *out);
/* GPU multiplyPlainNormal (src/evaluator_cuda.cu:1726-1807): the plaintext is lifted
   (m >= ceil(t/2) -> m - t) for every plaintext, monomials included.  Pinned by: the
   reference CPU multiplyPlainInplace on non-monomial plaintexts (equal words), and on a
   monomial below ceil(t/2) (no lift on either path). */
int or_bfv_multiply_plain(int logn, const uint64_t *moduli, size_t L, uint64_t t, const uint64_t *a, size_t polys,
                          const uint64_t *plain, uint64_t *out);

/* ---------------------------------------------------------------- samplers
 * The build's documented streams (the reference's Blake2xb / curand streams cannot
 * be reproduced).  Pinned by: RFC 8439 section 2.3.2 and OpenSSL's ChaCha20.       */
void or_chacha20_block(const uint32_t key[8], uint64_t counter, uint64_t nonce, uint32_t out[16]);
/* kind 0 ternary, 1 centred binomial, 2 uniform; key NULL = splitmix64 stream.
 * out: [L][N] coefficient form. */
int or_sample_poly(int kind, const uint32_t *key, uint64_t seed, int logn, const uint64_t *moduli, size_t L,
                   uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
