/*
 * sealhip.h -- C ABI of the MI355X (gfx950) RNS polynomial core.
 *
 * The reference has no C ABI: its GPU operations are reached through the C++
 * classes of src/troy_cuda.cuh and, one level down, through the kernel_util::k*
 * launchers (src/kernelutils.cuh:448-706).  This header is the thin layer that the
 * C++ class API (include/troyhip.hpp) and any FFI (ctypes, cgo, JNI) bind; each entry
 * point names the reference interface it replaces.
 *
 * Conventions
 *   - Device data uses the reference layout: a ciphertext is [poly][limb][coeff] of
 *     uint64, N coefficients per limb (src/ciphertext_cuda.cuh:60-67).
 *   - Residues are fully reduced, in [0, q_i), on entry and on return.
 *   - Pointers named d_* are device pointers; the library never allocates or frees
 *     them.  Work is ordered on `stream` (a hipStream_t; NULL = default stream).
 *   - Every function returns SHC_OK or a negative status.  shc_last_error() gives
 *     the message of the calling thread's last failure.  Status values map to the
 *     reference's exception types.
 */
#ifndef SEALHIP_H
#define SEALHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SHC_OK = 0,
    SHC_INVALID_ARGUMENT = -1, /* std::invalid_argument */
    SHC_LOGIC_ERROR = -2,      /* std::logic_error      */
    SHC_OUT_OF_RANGE = -3,     /* std::out_of_range     */
    SHC_RUNTIME_ERROR = -4,    /* std::runtime_error    */
    SHC_HIP_ERROR = -5,        /* HIP runtime failure (reported as std::runtime_error) */
    SHC_TIMEOUT = 1            /* shc_server_wait: the request is not done yet (not an error) */
};

enum { SHC_SCHEME_NONE = 0, SHC_SCHEME_BFV = 1, SHC_SCHEME_CKKS = 2, SHC_SCHEME_BGV = 3 };

typedef struct shc_context shc_context;

/* Message of this thread's last failed call ("" if none). */
const char *shc_last_error(void);

/* Select the device and initialise the runtime.
 * Replaces KernelProvider::initialize (src/kernelprovider.cuh:18-21). */
int shc_init(int device);

/* Build and upload the tables of one parameter set: NTT tables of every prime of
 * `coeff_modulus` (key level, special prime last), and the scheme's RNS tables.
 * Replaces SEALContextCuda(parms) (src/context_cuda.cuh:172-189, src/context_cuda.cu:5-32)
 * together with the CPU SEALContext it is built from (src/context.cpp:138-522). */
int shc_context_create(int scheme, uint32_t poly_modulus_degree, const uint64_t *coeff_modulus,
                       uint32_t coeff_modulus_size, uint64_t plain_modulus, shc_context **out);
void shc_context_destroy(shc_context *ctx);

/* CoeffModulus::Create (src/modulus.cpp:80-165): primes = 1 mod 2N (times t when
 * plain_modulus != 0) of the requested bit sizes, written to out[count]. */
int shc_coeff_modulus_create(uint32_t poly_modulus_degree, const int *bit_sizes, uint32_t count,
                             uint64_t plain_modulus, uint64_t *out);

/* In-place negacyclic NTT of d_data[poly_count][limb_count][N]; limb i is reduced
 * modulo coeff_modulus[i].  Natural order in, bit-reversed out.
 * Replaces kernel_util::kNttNegacyclicHarvey (src/kernelutils.cuh:578-603). */
int shc_ntt_forward(const shc_context *ctx, uint64_t *d_data, uint32_t poly_count, uint32_t limb_count,
                    void *stream);

/* Inverse of shc_ntt_forward (bit-reversed in, natural out, times N^-1).
 * Replaces kernel_util::kInverseNttNegacyclicHarvey (src/kernelutils.cuh:607-632). */
int shc_ntt_inverse(const shc_context *ctx, uint64_t *d_data, uint32_t poly_count, uint32_t limb_count,
                    void *stream);

/* A transform job: poly_count polynomials of limb_count limbs at d_data, in ctx's ring. */
typedef struct shc_ntt_job {
    const shc_context *ctx;
    uint64_t *d_data;
    uint32_t poly_count, limb_count;
} shc_ntt_job;
/* shc_ntt_forward / shc_ntt_inverse over the jobs of several contexts -- ring degrees and
 * limb counts mixed -- with one launch per pass for every job at N = 2^13 .. 2^17 (jobs at
 * smaller N run their own launches); at most 16 jobs.  The reference transforms one ring
 * per call (kNttNegacyclicHarvey, src/kernelutils.cuh:578-632); this is its batched form
 * for ciphertexts of different degrees. */
int shc_ntt_forward_multi(const shc_ntt_job *jobs, uint32_t n, void *stream);
int shc_ntt_inverse_multi(const shc_ntt_job *jobs, uint32_t n, void *stream);

/* ------------------------------------------------------------------ evaluator
 * Batched entry points: `count` ciphertexts packed back to back, all at the same
 * level with `limb_count` = L data primes.  Scratch memory is caller-owned; the
 * *_workspace_size queries give its size in bytes for a given shape. */

/* CKKS tensor product of size-2 ciphertexts in NTT form:
 * d_out[c] = (a0 b0, a0 b1 + a1 b0, a1 b1) for c < count; d_a, d_b: [count][2][L][N],
 * d_out: [count][3][L][N].
 * Replaces EvaluatorCuda::ckksMultiply (src/evaluator_cuda.cu:384-432). */
int shc_ckks_multiply(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out,
                      uint32_t count, uint32_t limb_count, void *stream);

int shc_multiply_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limb_count, size_t *bytes);

/* Scheme-dispatched multiply of size-2 ciphertexts: CKKS tensor in NTT form, or BFV
 * BEHZ multiplication in coefficient form (same layouts as shc_ckks_multiply).
 * Replaces EvaluatorCuda::multiplyInplace (src/evaluator_cuda.cu:262-281) with
 * bfvMultiply (src/evaluator_cuda.cu:283-382). */
int shc_multiply(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out, uint32_t count,
                 uint32_t limb_count, void *d_workspace, size_t workspace_bytes, void *stream);

int shc_keyswitch_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limb_count, size_t *bytes);

/* Key switching: for each ciphertext c, (c0, c1) += ModDown(sum_j ModUp(d_j) * key[j]),
 * where d_j are the limbs of the target polynomial (CKKS: NTT form; BFV: coefficient
 * form).  d_target + c * target_stride and d_ct + c * ct_stride (strides in words) give
 * ciphertext c's target and its (c0, c1).  d_key is a key-switching key at key level,
 * [K-1][2][K][N] in NTT form (KSwitchKeys::data(), src/kswitchkeys_cuda.cuh).
 * Replaces EvaluatorCuda::switchKeyInplace (src/evaluator_cuda.cu:1164-1363). */
int shc_switch_key(const shc_context *ctx, const uint64_t *d_target, size_t target_stride, uint64_t *d_ct,
                   size_t ct_stride, const uint64_t *d_key, uint32_t count, uint32_t limb_count, void *d_workspace,
                   size_t workspace_bytes, void *stream);

/* Relinearize size-3 ciphertexts d_ct[count][3][L][N] in place: the size-2 result is
 * left in polys 0 and 1 of each ciphertext; poly 2 is used as scratch (its words are
 * unspecified afterwards).
 * Replaces EvaluatorCuda::relinearizeInplace (src/evaluator_cuda.cuh:101-105,
 * src/evaluator_cuda.cu:703-744). */
int shc_relinearize(const shc_context *ctx, uint64_t *d_ct, const uint64_t *d_relin_key, uint32_t count,
                    uint32_t limb_count, void *d_workspace, size_t workspace_bytes, void *stream);

/* shc_relinearize without a workspace argument: the key-switch scratch is a fixed
 * 128 MiB device array inside the library (one per device, held until the process
 * ends), so the call allocates nothing and takes no allocating library routine.  No
 * counterpart in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Fails with
 * SHC_OUT_OF_RANGE when the key switch of `count` ciphertexts at this level needs more
 * scratch than that (shc_keyswitch_workspace_size > 134217728 bytes; e.g. one
 * ciphertext at N=2^16 with 24 data primes).  Every call on a device shares that
 * device's array: calls that overlap on one device (other streams, other threads) race;
 * the array keeps the last call's words, which no call reads before writing. */
int shc_relinearize_noalloc(const shc_context *ctx, uint64_t *d_ct, const uint64_t *d_relin_key, uint32_t count,
                            uint32_t limb_count, void *stream);

/* ---- launch-shape autotuner ----------------------------------------------------
 * The key switch's two dominant kernels (the ModUp columns pass and the fused ModUp
 * rows pass + key products) have launch-shape choices whose best value depends on the
 * level, the batch and the degree: target primes per columns-pass block (1..L+1) and
 * ciphertexts per rows-pass block (4, 2 or 1).  Zero means the built-in choice.  A
 * context keeps a table of shapes by (limb_count, count); every key switch of a shape
 * in the table uses it.  No counterpart in the reference, which launches every kernel
 * as <<<ceil(n/256), 256>>> (src/kernelutils.cu:3). */
typedef struct {
    uint32_t cols_targets; /* target primes per ModUp columns-pass block; 0: built-in */
    uint32_t rows_group;   /* ciphertexts per fused rows-pass block (4, 2, 1); 0: built-in */
} shc_ks_shape;

/* Times `reps` relinearizations of d_ct [count][3][L][N] (scratch: its words are
 * overwritten) with each candidate shape -- the built-in choice and every
 * (cols_targets, rows_group) pair -- and stores the fastest in the context's table for
 * (limb_count, count) (the built-in choice unless another is more than 1% faster).  Synchronises `stream`; key switches on this context that run
 * during the call may take any candidate shape (their words do not depend on it). */
int shc_autotune_relinearize(const shc_context *ctx, uint64_t *d_ct, const uint64_t *d_relin_key, uint32_t count,
                             uint32_t limb_count, void *d_workspace, size_t workspace_bytes, int reps, void *stream,
                             shc_ks_shape *best, float *best_ms, float *builtin_ms);
/* Sets (shape NULL or all zero: removes) / reads the table entry of (limb_count, count). */
int shc_keyswitch_shape_set(const shc_context *ctx, uint32_t limb_count, uint32_t count, const shc_ks_shape *shape);
int shc_keyswitch_shape_get(const shc_context *ctx, uint32_t limb_count, uint32_t count, shc_ks_shape *shape);
/* The table as text, one line "limb_count count cols_targets rows_group" per entry:
 * export writes it (NUL-terminated) to buf when capacity allows and its size with the
 * NUL to *length; import adds a table written by export (lines are validated first). */
int shc_keyswitch_shapes_export(const shc_context *ctx, char *buf, size_t capacity, size_t *length);
int shc_keyswitch_shapes_import(const shc_context *ctx, const char *text);

/* ---- level groups -------------------------------------------------------------
 * A batch of ciphertexts at uneven levels, as groups of one level each (a work table
 * of up to 16 groups per launch; more are taken 16 at a time).  Each grouped call gives
 * every group the words of the one-level call on it, and runs each stage with one
 * launch for all groups: CKKS at N = 2^15 and 2^16 (the fused key switch); other
 * degrees and BFV key switches run group by group.  No counterpart in the reference
 * (one ciphertext per call, src/evaluator_cuda.cu:262-281). */
typedef struct {
    uint64_t *d_ct;      /* the group's ciphertexts back to back */
    uint32_t count;      /* ciphertexts in the group */
    uint32_t limb_count; /* their level (data primes) */
} shc_ct_group;

/* CKKS tensor products of groups: out[g] = a[g] x b[g] (size 2 x size 2 -> size 3). */
int shc_ckks_multiply_groups(const shc_context *ctx, const shc_ct_group *a, const shc_ct_group *b,
                             const shc_ct_group *out, uint32_t ngroups, void *stream);
/* Relinearize size-3 ciphertext groups in place (as shc_relinearize on each group). */
int shc_relinearize_groups_workspace_size(const shc_context *ctx, const shc_ct_group *cts, uint32_t ngroups,
                                          size_t *bytes);
int shc_relinearize_groups(const shc_context *ctx, const shc_ct_group *cts, uint32_t ngroups,
                           const uint64_t *d_relin_key, void *d_workspace, size_t workspace_bytes, void *stream);
/* Rescale / modulus-switch polys 0 and 1 of each ciphertext of in[g] (in_size polys per
 * ciphertext in memory: 2, or 3 straight after a relinearization) into out[g] (size 2,
 * one level down; must not overlap in). */
int shc_rescale_groups_workspace_size(const shc_context *ctx, const shc_ct_group *in, uint32_t ngroups,
                                      size_t *bytes);
int shc_rescale_to_next_groups(const shc_context *ctx, const shc_ct_group *in, uint32_t in_size,
                               const shc_ct_group *out, uint32_t ngroups, void *d_workspace, size_t workspace_bytes,
                               void *stream);

/* GaloisTool::getEltFromStep (src/utils/galois_cuda.cu:44-86): 3^step mod 2N, a
 * negative step s maps to N/2 - |s|, step 0 gives 2N - 1 (complex conjugation). */
int shc_galois_elt_from_step(uint32_t poly_modulus_degree, int step, uint32_t *elt);

/* d_out[g] = sigma_elt(d_in[g]) for `poly_count` polynomials of L limbs, in NTT form
 * (ntt_form != 0) or coefficient form.  d_in and d_out must not overlap.
 * Replaces GaloisToolCuda::applyGaloisNtt / applyGalois (src/utils/galois_cuda.cuh:21-25). */
int shc_apply_galois(const shc_context *ctx, const uint64_t *d_in, uint64_t *d_out, uint32_t poly_count,
                     uint32_t limb_count, uint32_t galois_elt, int ntt_form, void *stream);

int shc_galois_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limb_count, size_t *bytes);

/* Apply sigma_elt to size-2 ciphertexts d_ct[count][2][L][N] in place and switch back to
 * the original key with d_galois_key ([K-1][2][K][N], the key of element `galois_elt`,
 * i.e. GaloisKeys::data()[(elt - 1) / 2]).  CKKS ciphertexts are in NTT form, BFV in
 * coefficient form.
 * Replaces EvaluatorCuda::applyGaloisInplace (src/evaluator_cuda.cu:2011-2104); rotate
 * rows / vector compose it with shc_galois_elt_from_step (src/evaluator_cuda.cu:2106-2163). */
int shc_apply_galois_keyswitch(const shc_context *ctx, uint64_t *d_ct, const uint64_t *d_galois_key,
                               uint32_t galois_elt, uint32_t count, uint32_t limb_count, void *d_workspace,
                               size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------ keys, encryption
 * Sampling is counter based, so a seed reproduces a polynomial exactly.  A context with
 * a sampler key draws every stream from ChaCha20 under that 256-bit key, the seed being
 * the 64-bit nonce (word i of seed s = words 2(i % 8), 2(i % 8) + 1 of keystream block
 * i / 8, little endian): a cryptographic generator, which is what keys and encryptions
 * need.  A context without one (the default of shc_context_create) draws word i of
 * seed s as splitmix64's i-th output -- fast and reproducible, for explicitly seeded
 * tests and benchmarks only: its streams have 64 bits of entropy.  The reference
 * samples with Blake2xb (CPU) or curand (GPU); distributions match (ternary; centred
 * binomial, sigma ~ 3.24; uniform mod q_i), streams do not.  Every random input can
 * also be passed explicitly (shc_encrypt_zero_asymmetric). */
enum { SHC_SAMPLE_TERNARY = 0, SHC_SAMPLE_CBD = 1, SHC_SAMPLE_UNIFORM = 2 };

/* Sets the context's sampler key (32 bytes, read as 8 little-endian words), or clears
 * it (key = NULL: splitmix64 streams).  Not synchronised with calls that sample on
 * this context: set it before them.  No counterpart in the reference (its KeyGenerator
 * seeds Blake2xb from std::random_device, src/randomgen.cpp). */
int shc_context_set_sampler_key(shc_context *ctx, const uint8_t *key);

/* d_out[limb][N] (coefficient form) for the first limb_count key-level primes.
 * Replaces samplePolyTernary / samplePolyCbd / samplePolyUniform (src/utils/rlwe_cuda.cu:25-150). */
int shc_sample_poly(const shc_context *ctx, int kind, uint64_t seed, uint32_t limb_count, uint64_t *d_out,
                    void *stream);

/* Secret key [K][N] in NTT form.  Replaces KeyGeneratorCuda(context) (src/keygenerator_cuda.cuh). */
int shc_keygen_secret(const shc_context *ctx, uint64_t seed, uint64_t *d_secret_key, void *stream);
/* Public key [2][K][N] in NTT form.  Replaces KeyGeneratorCuda::createPublicKey. */
int shc_keygen_public(const shc_context *ctx, const uint64_t *d_secret_key, uint64_t seed, uint64_t *d_public_key,
                      void *stream);
int shc_keygen_workspace_size(const shc_context *ctx, size_t *bytes);
/* Relinearization key [K-1][2][K][N] (key for s^2).  Replaces KeyGeneratorCuda::createRelinKeys. */
int shc_keygen_relin(const shc_context *ctx, const uint64_t *d_secret_key, uint64_t seed, uint64_t *d_relin_key,
                     void *d_workspace, size_t workspace_bytes, void *stream);
/* Galois key of one element [K-1][2][K][N].  Replaces KeyGeneratorCuda::createGaloisKeys. */
int shc_keygen_galois(const shc_context *ctx, const uint64_t *d_secret_key, uint32_t galois_elt, uint64_t seed,
                      uint64_t *d_galois_key, void *d_workspace, size_t workspace_bytes, void *stream);

int shc_encrypt_workspace_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
/* Encryption of zero at level L (limb_count) from explicit randomness u, e0, e1
 * ([L+1][N] coefficient form: one level above, or L when L is the key level), then
 * divide-and-round to L limbs; CKKS output in NTT form, BFV in coefficient form.
 * Replaces EncryptorCuda::encryptZeroInternal (src/encryptor_cuda.cu:92-160). */
int shc_encrypt_zero_asymmetric(const shc_context *ctx, const uint64_t *d_public_key, const uint64_t *d_u,
                                const uint64_t *d_e0, const uint64_t *d_e1, uint32_t limb_count, uint64_t *d_out,
                                void *d_workspace, size_t workspace_bytes, void *stream);
/* Public-key encryption of d_plain (BFV: [N] coefficients mod t; CKKS: [L][N] in NTT
 * form; NULL encrypts zero) with randomness drawn from `seed`.
 * Replaces EncryptorCuda::encrypt (src/encryptor_cuda.cuh:123, src/encryptor_cuda.cu:162-237). */
int shc_encrypt(const shc_context *ctx, const uint64_t *d_public_key, const uint64_t *d_plain, uint32_t limb_count,
                uint64_t seed, uint64_t *d_out, void *d_workspace, size_t workspace_bytes, void *stream);
/* `count` public-key encryptions in one pass: ciphertext g (d_out + g * 2 * L * N) is
 * the one shc_encrypt gives for seed + g and plaintext d_plain + g * (BFV: N, CKKS:
 * L * N) words (d_plain NULL: encryptions of zero).  The workspace needs room for one
 * ciphertext at least; shc_encrypt_many_workspace_size(count) lets the whole batch
 * share each launch, a smaller one runs it in slices.  The reference encrypts one
 * ciphertext per EncryptorCuda::encrypt call (src/encryptor_cuda.cu:162-237). */
int shc_encrypt_many_workspace_size(const shc_context *ctx, uint32_t limb_count, uint32_t count, size_t *bytes);
int shc_encrypt_many(const shc_context *ctx, const uint64_t *d_public_key, const uint64_t *d_plain, uint32_t count,
                     uint32_t limb_count, uint64_t seed, uint64_t *d_out, void *d_workspace, size_t workspace_bytes,
                     void *stream);
/* Symmetric encryption of zero with the secret key (utils/rlwe_cuda.cu:262-331). */
int shc_encrypt_zero_symmetric(const shc_context *ctx, const uint64_t *d_secret_key, uint32_t limb_count,
                               uint64_t seed, uint64_t *d_out, void *stream);

int shc_decrypt_workspace_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
/* Decrypt a size-2 or size-3 ciphertext d_ct[size][L][N].  BFV: d_out [N] mod t;
 * CKKS: d_out [L][N], the NTT-form plaintext.
 * Replaces DecryptorCuda::decrypt (src/decryptor_cuda.cuh:66, src/decryptor_cuda.cu:92-328). */
int shc_decrypt(const shc_context *ctx, const uint64_t *d_ct, uint32_t ct_size, const uint64_t *d_secret_key,
                uint32_t limb_count, uint64_t *d_out, void *d_workspace, size_t workspace_bytes, void *stream);

/* `count` ciphertexts d_ct[count][size][L][N] decrypted in one pass into d_out[count][N]
 * (BFV) or d_out[count][L][N] (CKKS), the words of shc_decrypt per ciphertext.  BFV
 * needs size * L * N words of workspace at least (shc_decrypt_many_workspace_size(count)
 * takes the batch at once); CKKS none.  The reference decrypts one ciphertext per
 * DecryptorCuda::decrypt call (src/decryptor_cuda.cu:92-328). */
int shc_decrypt_many_workspace_size(const shc_context *ctx, uint32_t limb_count, uint32_t count, size_t *bytes);
int shc_decrypt_many(const shc_context *ctx, const uint64_t *d_ct, uint32_t ct_size, uint32_t count,
                     const uint64_t *d_secret_key, uint32_t limb_count, uint64_t *d_out, void *d_workspace,
                     size_t workspace_bytes, void *stream);

/* BatchEncoder: d_values[count <= N] -> d_plain[N] mod t, and back (d_scratch: N words).
 * Replaces BatchEncoderCuda::encode / decode (src/batchencoder_cuda.cuh:61-75). */
int shc_batch_encode(const shc_context *ctx, const uint64_t *d_values, uint32_t count, uint64_t *d_plain,
                     void *stream);
int shc_batch_decode(const shc_context *ctx, const uint64_t *d_plain, uint64_t *d_values, uint64_t *d_scratch,
                     void *stream);
/* Batches of plaintexts in one pass: plaintext g's `count` values at d_values + g *
 * values_stride into d_plain[g][N]; decode d_plain[plaintexts][N] into
 * d_values[plaintexts][N] through d_scratch[plaintexts][N].  The reference encodes and
 * decodes one plaintext per BatchEncoderCuda call. */
int shc_batch_encode_many(const shc_context *ctx, const uint64_t *d_values, uint32_t count, size_t values_stride,
                          uint32_t plaintexts, uint64_t *d_plain, void *stream);
int shc_batch_decode_many(const shc_context *ctx, const uint64_t *d_plain, uint32_t plaintexts, uint64_t *d_values,
                          uint64_t *d_scratch, void *stream);

/* CKKSEncoder on the device (fp64): d_values holds `count` (re, im) double pairs,
 * count <= N/2; d_plain is the NTT-form plaintext [L][N] at the level with
 * limb_count primes.  Encode rounds to integers (coefficients up to 128 bits).
 * Replaces CKKSEncoderCuda::encode / decode (src/ckks_cuda.cuh:21-84). */
int shc_ckks_encode_workspace_size(const shc_context *ctx, size_t *bytes);
int shc_ckks_encode(const shc_context *ctx, const double *d_values, uint32_t count, double scale, uint32_t limb_count,
                    uint64_t *d_plain, void *d_workspace, size_t workspace_bytes, void *stream);
/* shc_ckks_encode without waiting for the device: the range check that shc_ckks_encode
 * raises (encoded values not finite or too large for the coefficient modulus,
 * CKKSEncoder::encodeInternal, src/ckks.cpp:167-171) is written to the device word
 * *d_status instead -- 0 when encoded, SHC_INVALID_ARGUMENT when rejected, the plaintext
 * then all zero -- for the caller to read when it needs to.  Argument errors known on
 * the host are still returned.  The call launches kernels only, so it can be captured
 * in a HIP graph.  No counterpart in the reference. */
int shc_ckks_encode_deferred(const shc_context *ctx, const double *d_values, uint32_t count, double scale,
                             uint32_t limb_count, uint64_t *d_plain, int32_t *d_status, void *d_workspace,
                             size_t workspace_bytes, void *stream);
/* Batches of plaintexts, two LDS-tiled DWT passes for the whole batch: plaintext p's
 * `count` (re, im) pairs at d_values + p * values_stride (doubles), its NTT-form words at
 * d_plain + p * limb_count * N.  d_status: NULL raises the range check like
 * shc_ckks_encode (one wait for the batch), else one int32 per plaintext as
 * shc_ckks_encode_deferred.  Decode writes plaintext p's N/2 pairs at d_values + p * N.
 * No counterpart in the reference (one plaintext per call). */
int shc_ckks_encode_many_workspace_size(const shc_context *ctx, uint32_t plaintexts, size_t *bytes);
int shc_ckks_encode_many(const shc_context *ctx, const double *d_values, uint32_t count, size_t values_stride,
                         uint32_t plaintexts, double scale, uint32_t limb_count, uint64_t *d_plain,
                         int32_t *d_status, void *d_workspace, size_t workspace_bytes, void *stream);
int shc_ckks_decode_many_workspace_size(const shc_context *ctx, uint32_t limb_count, uint32_t plaintexts,
                                        size_t *bytes);
int shc_ckks_decode_many(const shc_context *ctx, const uint64_t *d_plain, uint32_t limb_count, uint32_t plaintexts,
                         double scale, double *d_values, void *d_workspace, size_t workspace_bytes, void *stream);
int shc_ckks_decode_workspace_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
int shc_ckks_decode(const shc_context *ctx, const uint64_t *d_plain, uint32_t limb_count, double scale,
                    double *d_values, void *d_workspace, size_t workspace_bytes, void *stream);
/* CKKSEncoderCuda::encodePolynomial / decodePolynomial (src/ckks_cuda.cuh:34-40, 84-87;
 * src/ckks_cuda.cu:455-572, 983-1053): the coefficients themselves, with no FFT.  Encode
 * takes count <= N doubles at d_values, multiplies them by scale, rounds half away from
 * zero and writes the [L][N] NTT-form plaintext (workspace: shc_ckks_encode_workspace_size;
 * waits for the range check like shc_ckks_encode).  Decode writes the N centred
 * coefficients divided by scale (workspace: shc_ckks_decode_workspace_size). */
int shc_ckks_encode_polynomial(const shc_context *ctx, const double *d_values, uint32_t count, double scale,
                               uint32_t limb_count, uint64_t *d_plain, void *d_workspace, size_t workspace_bytes,
                               void *stream);
int shc_ckks_decode_polynomial(const shc_context *ctx, const uint64_t *d_plain, uint32_t limb_count, double scale,
                               double *d_values, void *d_workspace, size_t workspace_bytes, void *stream);

int shc_rescale_workspace_size(const shc_context *ctx, uint32_t count, uint32_t poly_count, uint32_t limb_count,
                               size_t *bytes);

/* Divide-and-round by the last prime, dropping one limb: CKKS rescale (NTT form) or
 * BFV modulus switch (coefficient form).  d_in: [count][poly_count][L][N],
 * d_out: [count][poly_count][L-1][N] (must not overlap d_in).
 * Replaces EvaluatorCuda::rescaleToNext / modSwitchToNext (src/evaluator_cuda.cu:749-824). */
int shc_rescale_to_next(const shc_context *ctx, const uint64_t *d_in, uint64_t *d_out, uint32_t count,
                        uint32_t poly_count, uint32_t limb_count, void *d_workspace, size_t workspace_bytes,
                        void *stream);

/* ---- coefficient-wise evaluator operations ------------------------------------
 * Polynomials are [polys][limb_count][N] (polys = ciphertexts x size); out may alias
 * an input.  Both operands in the same form (coefficient or NTT). */

/* out = a + b.  Replaces EvaluatorCuda::addInplace (src/evaluator_cuda.cu:156-208;
 * CPU src/evaluator.cpp:167-266). */
int shc_add(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out, uint32_t polys,
            uint32_t limb_count, void *stream);
/* out = a - b.  Replaces EvaluatorCuda::subInplace (src/evaluator_cuda.cu:210-281). */
int shc_sub(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_out, uint32_t polys,
            uint32_t limb_count, void *stream);
/* out = -a.  Replaces EvaluatorCuda::negateInplace (src/evaluator_cuda.cu:145-154). */
int shc_negate(const shc_context *ctx, const uint64_t *d_a, uint64_t *d_out, uint32_t polys, uint32_t limb_count,
               void *stream);
/* out[p] = a[p] * plain (dyadic; a and plain in NTT form, plain [limb_count][N]).
 * Replaces EvaluatorCuda::multiplyPlainNtt (src/evaluator_cuda.cu:1811-1867). */
int shc_multiply_plain_ntt(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_plain, uint64_t *d_out,
                           uint32_t polys, uint32_t limb_count, void *stream);
/* d_out[p] = sum_t d_plains[t] * d_cts[t][p] for NTT-form polynomials: d_cts
 * [terms][polys][limb_count][N], d_plains [terms][limb_count][N], d_out [polys][L][N]
 * (not overlapping d_cts); 128-bit sums, one reduction per 64 terms.  The inner sums of
 * a diagonal matrix-vector product (app/matmul.cuh:98-121: multiplyPlain then add per
 * diagonal) in one pass. */
int shc_multiply_plain_accumulate(const shc_context *ctx, const uint64_t *d_cts, const uint64_t *d_plains,
                                  uint32_t terms, uint32_t polys, uint32_t limb_count, uint64_t *d_out, void *stream);
/* `outputs` such sums over the same d_cts at once: d_plains [outputs][terms][limb_count][N],
 * d_out [outputs][polys][limb_count][N] (not overlapping d_cts); output o equals
 * shc_multiply_plain_accumulate with d_plains + o * terms * limb_count * N.  Each ciphertext
 * word is read once for up to 8 outputs (all the giant steps' inner sums of a
 * baby-step giant-step matrix-vector product in a few passes).  No counterpart in the
 * reference (its diagonal products read each rotation once per diagonal). */
int shc_multiply_plain_accumulate_many(const shc_context *ctx, const uint64_t *d_cts, const uint64_t *d_plains,
                                       uint32_t terms, uint32_t outputs, uint32_t polys, uint32_t limb_count,
                                       uint64_t *d_out, void *stream);
/* Ciphertexts [count][size][limb_count][N]: out = ct with poly 0 +- plain ([limb_count][N],
 * same form as ct; subtract != 0 subtracts).  CKKS add/subPlainInplace
 * (src/evaluator_cuda.cu:1544-1690). */
int shc_add_plain(const shc_context *ctx, const uint64_t *d_ct, const uint64_t *d_plain, uint64_t *d_out,
                  uint32_t count, uint32_t size, uint32_t limb_count, int subtract, void *stream);
/* BFV: out = ct with c0 +- Delta m, m = plain [N] mod t (coefficient form), SEAL's
 * rounding of q m / t.  BFV add/subPlainInplace through multiplyAdd/SubPlainWith-
 * ScalingVariant (src/utils/scalingvariant.cpp:53-131). */
int shc_bfv_add_plain(const shc_context *ctx, const uint64_t *d_ct, const uint64_t *d_plain, uint64_t *d_out,
                      uint32_t count, uint32_t size, uint32_t limb_count, int subtract, void *stream);
/* BFV: out[p] = a[p] * m for coefficient-form a [polys][limb_count][N] and m = plain [N]
 * mod t lifted as m - t for m >= ceil(t/2).  Replaces EvaluatorCuda::multiplyPlainNormal
 * (src/evaluator_cuda.cu:1744-1807). */
int shc_bfv_multiply_plain_workspace_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
int shc_bfv_multiply_plain(const shc_context *ctx, const uint64_t *d_a, const uint64_t *d_plain, uint64_t *d_out,
                           uint32_t polys, uint32_t limb_count, void *d_workspace, size_t workspace_bytes,
                           void *stream);

/* BFV: d_out [limb_count][N] = NTT of m lifted as in shc_bfv_multiply_plain, m = d_plain
 * [N] mod t.  Replaces EvaluatorCuda::transformToNttInplace(PlaintextCuda&, ParmsID)
 * (src/evaluator_cuda.cu:1869-1935). */
int shc_bfv_plain_to_ntt(const shc_context *ctx, const uint64_t *d_plain, uint32_t limb_count, uint64_t *d_out,
                         void *stream);

/* ---- hoisted rotations --------------------------------------------------------
 * Sums of rotations of one ciphertext (the diagonal matrix-vector pattern of
 * app/matmul.cuh:98-121: rotateVectorInplace / rotateRowsInplace then addInplace) with
 * c1's ModUp done once for all rotations.  The words equal the sum of separate
 * shc_apply_galois_keyswitch results.  A Galois key is first prepared for a level
 * (permuted copy plus a sign-pattern correction, shc_hoisted_key_size bytes).  A prepared
 * key is bound to its level and Galois element: its last two words name them, and
 * shc_rotate_sum / shc_rotate_many reject (SHC_INVALID_ARGUMENT) a key prepared for
 * another limb_count or element. */
int shc_hoisted_key_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
int shc_hoisted_key_workspace_size(const shc_context *ctx, uint32_t limb_count, size_t *bytes);
int shc_hoisted_key_prepare(const shc_context *ctx, const uint64_t *d_galois_key, uint32_t galois_elt,
                            uint32_t limb_count, uint64_t *d_hoisted_key, void *d_workspace, size_t workspace_bytes,
                            void *stream);
/* d_out[c] = sum_r rotate_{galois_elts[r]}(d_in[c]) for size-2 ciphertexts
 * [count][2][L][N] (CKKS: NTT form; BFV: coefficient form); d_out must not overlap d_in.
 * galois_elts, d_hoisted_keys and d_galois_keys are host arrays of `rotations` entries
 * (device pointers for the keys; the Galois keys serve inputs with a zero digit, where
 * hoisting is not exact).  Synchronises `stream` once, so it cannot be graph-captured. */
int shc_rotate_sum_workspace_size(const shc_context *ctx, uint32_t count, uint32_t limb_count, size_t *bytes);
int shc_rotate_sum(const shc_context *ctx, const uint64_t *d_in, uint32_t count, uint32_t limb_count,
                   const uint32_t *galois_elts, const uint64_t *const *d_hoisted_keys,
                   const uint64_t *const *d_galois_keys, uint32_t rotations, uint64_t *d_out, void *d_workspace,
                   size_t workspace_bytes, void *stream);

/* d_out[r][c] = rotate_{galois_elts[r]}(d_in[c]) (d_out: [rotations][count][2][L][N], not
 * overlapping d_in), the words of separate shc_apply_galois_keyswitch calls, with one
 * ModUp of c1 for all rotations -- the baby steps of a BSGS matrix-vector product.  Same
 * arguments and workspace as shc_rotate_sum. */
int shc_rotate_many(const shc_context *ctx, const uint64_t *d_in, uint32_t count, uint32_t limb_count,
                    const uint32_t *galois_elts, const uint64_t *const *d_hoisted_keys,
                    const uint64_t *const *d_galois_keys, uint32_t rotations, uint64_t *d_out, void *d_workspace,
                    size_t workspace_bytes, void *stream);

/* Instrumentation (no counterpart in the reference; for benchmarks and profilers):
 * arms `ctx` so that its next key switch (shc_relinearize, shc_switch_key,
 * shc_apply_galois_keyswitch) records the caller-created hipEvent_t `begin_event`
 * right before and `end_event` right after its ModUp rows-pass kernel -- the dominant
 * kernel of a key switch -- on that call's stream, then disarms.  For N < 2^13 the
 * events bracket the whole ModUp forward NTT; a batch large enough to run in slices
 * reports its first slice.  shc_rotate_sum / shc_rotate_many (hoisted key switches)
 * record them around their first pass of products with the prepared keys (up to four
 * rotations over the shared ModUp limbs), their dominant kernel.  Arming is
 * thread-safe: of key switches that run concurrently on one context (on different
 * streams), at most one takes the armed probe. */
int shc_probe_keyswitch(shc_context *ctx, void *begin_event, void *end_event);

/* ---- serving ------------------------------------------------------------------
 * A continuous-batching request server (no counterpart in the reference, whose
 * serving-shaped caller, app/matmul.cuh:98-121, runs one request at a time on the
 * default stream).  The server allocates one device arena of `device_bytes` and one
 * pinned host staging arena of `staging_bytes` at creation, and nothing afterwards.
 * It drives `lanes` (1-4) HIP streams of default priority, one host thread each.
 * Requests are submitted with host buffers (input copied into staging before
 * shc_server_submit returns; the result written to h_out before the request is
 * reported done).  A free lane takes the type of the oldest waiting request and admits
 * every waiting request of that type, up to `max_batch`, whose batch fits the device
 * memory free at that moment (when none fits, it waits for another lane to release
 * memory); the batch runs on the lane's stream (upload, batched operations, download).
 * Batches of different lanes overlap on the device.  A request's words do not depend on
 * the batch it ran in.  Contexts, keys and diagonals a type refers to stay owned by the
 * caller and must outlive the server.  Safe to call from several host threads. */
typedef struct shc_server shc_server;

typedef struct {
    double submit_s, admit_s, done_s; /* seconds on the server's clock (shc_server_clock) */
    uint32_t batch_size, lane;        /* the batch the request ran in, and its lane */
    uint64_t batch_id;
} shc_request_info;

typedef struct {
    uint64_t submitted, completed, failed, batches;
    uint32_t max_batches_in_flight; /* most batches that ran at once */
    size_t peak_device_bytes;       /* most device arena bytes taken at once */
} shc_server_stats;

int shc_server_create(int device, size_t device_bytes, size_t staging_bytes, uint32_t lanes, uint32_t max_batch,
                      shc_server **out);
/* Stops the lanes after their running batches; requests still waiting are dropped. */
void shc_server_destroy(shc_server *server);

/* Request type "diagonal matrix-vector product" (app/matmul.cuh:98-121 with baby-step
 * giant-step rotations): input and output one size-2 ciphertext [2][L][N] (CKKS: NTT
 * form; BFV: coefficient form); d_diagonals [giant][baby][L][N] the pre-rotated
 * diagonals in NTT form; d_baby_keys / d_baby_hoisted: Galois keys of steps
 * 1..baby-1 and their forms prepared for level L (shc_hoisted_key_prepare);
 * d_giant_keys: Galois keys of steps baby, 2 baby, ..., (giant-1) baby (host arrays of
 * device pointers; copied). */
int shc_server_add_matvec(shc_server *server, const shc_context *ctx, uint32_t limb_count, uint32_t baby,
                          uint32_t giant, const uint64_t *d_diagonals, const uint64_t *const *d_baby_keys,
                          const uint64_t *const *d_baby_hoisted, const uint64_t *const *d_giant_keys,
                          uint32_t *type);
/* Request type "multiply, relinearize[, rescale]" (EvaluatorCuda::multiplyInplace then
 * relinearizeInplace, src/evaluator_cuda.cu:262-281, 703-744): input two size-2
 * ciphertexts a, b back to back [2][2][L][N], output [2][L][N] ([2][L-1][N] after the
 * rescale / modulus switch when rescale != 0). */
int shc_server_add_mulrelin(shc_server *server, const shc_context *ctx, uint32_t limb_count,
                            const uint64_t *d_relin_key, int rescale, uint32_t *type);
/* Words of one request's input and output, and the device bytes a batch of one takes. */
int shc_server_request_size(const shc_server *server, uint32_t type, size_t *in_words, size_t *out_words,
                            size_t *device_bytes);
/* Queues a request; *id names it.  A buffer inside a block from shc_server_host_alloc is
 * used in place: the input is uploaded from it (it must stay unchanged until the request
 * is done) and the output downloaded into it.  Other buffers go through the staging
 * arena: the input is copied before the call returns (blocking while staging is full)
 * and the output copied out before the request is reported done. */
int shc_server_submit(shc_server *server, uint32_t type, const uint64_t *h_in, uint64_t *h_out, uint64_t *id);
/* Pinned host memory (hipHostMalloc on this call) that requests can use in place; freed
 * by shc_server_host_free or with the server. */
int shc_server_host_alloc(shc_server *server, size_t bytes, void **ptr);
int shc_server_host_free(shc_server *server, void *ptr);
/* Waits up to timeout_s seconds (< 0: no limit) for request id: SHC_OK when done (the
 * request is then forgotten; info, if not NULL, describes it), SHC_TIMEOUT when not
 * yet, or the error the request's batch failed with. */
int shc_server_wait(shc_server *server, uint64_t id, double timeout_s, shc_request_info *info);
/* Seconds since the server was created (the clock of shc_request_info). */
double shc_server_clock(const shc_server *server);
int shc_server_get_stats(const shc_server *server, shc_server_stats *out);

#ifdef __cplusplus
}
#endif
#endif
