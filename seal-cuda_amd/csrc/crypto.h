// crypto.h -- sampling, key generation, encryption, decryption and batch encoding.
//
// Randomness: every sampler is counter based, so any coefficient can be drawn
// independently by any thread and a seed reproduces a polynomial exactly.  A context
// with a sampler key (Context::sampler_keyed) takes word i of the stream with seed s
// from the ChaCha20 keystream under that 256-bit key (nonce s, block i / 8, words
// 2(i % 8) and 2(i % 8) + 1 little endian) -- a cryptographic generator, for keys and
// encryptions; without one, word i is splitmix64's i-th output,
// mix(s + (i + 1) * 0x9E3779B97F4A7C15), for explicitly seeded tests and benchmarks.
// The reference samples with Blake2xb (CPU, src/utils/rlwe.cpp:21-160) or curand (GPU,
// src/utils/rlwe_cuda.cu:25-150); the distributions are the same, the streams differ.
#pragma once
#include "context.h"

namespace shc {

enum SampleKind { SAMPLE_TERNARY = 0, SAMPLE_CBD = 1, SAMPLE_UNIFORM = 2 };

// out: [L][N] in coefficient form; limb l reduced modulo primes[l] (a ternary or CBD
// sample is one integer per coefficient written into every limb).
void sample_poly(const Context &c, SampleKind kind, uint64_t seed, const DevPrime *primes, uint32_t L, uint64_t *out,
                 hipStream_t st);

// Secret key: ternary at key level, NTT form [K][N].
void keygen_secret(const Context &c, uint64_t seed, uint64_t *sk, hipStream_t st);
// c0 = -(a s + e), c1 = a at L limbs of `primes`; NTT form when ntt_form, else both
// inverse transformed (src/utils/rlwe.cpp:239-331).  sk: NTT form, at least L limbs.
void encrypt_zero_symmetric(const Context &c, const uint64_t *sk, const DevPrime *primes, uint32_t L, bool ntt_form,
                            uint64_t seed, uint64_t *out, hipStream_t st);
// Public key = symmetric zero at key level in NTT form: [2][K][N].
void keygen_public(const Context &c, const uint64_t *sk, uint64_t seed, uint64_t *pk, hipStream_t st);
// Key-switching key to new_key ([K][N] NTT): [K-1][2][K][N] (src/keygenerator.cpp:294-329).
void keygen_kswitch(const Context &c, const uint64_t *sk, const uint64_t *new_key, uint64_t seed, uint64_t *key,
                    hipStream_t st);
size_t keygen_workspace(const Context &c);
void keygen_relin(const Context &c, const uint64_t *sk, uint64_t seed, uint64_t *key, void *ws, size_t ws_bytes,
                  hipStream_t st);
void keygen_galois(const Context &c, const uint64_t *sk, uint32_t elt, uint64_t seed, uint64_t *key, void *ws,
                   size_t ws_bytes, hipStream_t st);

// Asymmetric zero at level L from explicit randomness u, e0, e1 ([L+1][N], coefficient
// form, level L+1 = L data primes + the next prime up the chain), then divide-and-round
// down to L limbs (src/encryptor.cpp:88-150, src/utils/rlwe.cpp:170-237).
size_t encrypt_workspace(const Context &c, uint32_t L);
void encrypt_zero_asymmetric(const Context &c, const uint64_t *pk, const uint64_t *u, const uint64_t *e0,
                             const uint64_t *e1, uint32_t L, uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st);
// Samples u, e0, e1 from `seed` (into the workspace) and encrypts `plain`: BFV plain is
// [N] mod t (coefficients), CKKS plain is [L][N] in NTT form; out: [2][L][N].
void encrypt(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t L, uint64_t seed, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st);
// `count` encryptions: ciphertext g is encrypt()'s for seed + g and plain + g * (BFV: N,
// CKKS: L * N) words (plain == nullptr: zeros), out [count][2][L][N].  The workspace
// holds one ciphertext's samples and products at least; larger ones let more
// ciphertexts share each launch (encrypt_many_workspace(c, L, count) takes all at once).
size_t encrypt_many_workspace(const Context &c, uint32_t L, uint32_t count);
void encrypt_many(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t count, uint32_t L,
                  uint64_t seed, uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st);
// c0 += Delta m with SEAL's rounding fix, at level L (src/utils/scalingvariant.cpp:48-86).
void bfv_add_plain_scaled(const Context &c, uint64_t *c0, const uint64_t *plain, uint32_t L, hipStream_t st);
// c0 +- Delta m (multiplyAdd / multiplySubPlainWithScalingVariant, scalingvariant.cpp:53-131).
void bfv_scale_add(const Context &c, uint64_t *c0, const uint64_t *plain, uint32_t L, bool subtract, hipStream_t st);

// Decryption: phase = sum_i c_i s^i for ct [size][L][N]; BFV: out [N] mod t after
// scale-and-round; CKKS: out [L][N] NTT-form plaintext (src/decryptor.cpp:79-330).
size_t decrypt_workspace(const Context &c, uint32_t L);
void decrypt(const Context &c, const uint64_t *ct, uint32_t size, const uint64_t *sk, uint32_t L, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st);
// `count` ciphertexts [count][size][L][N] -> out [count][N] (BFV) or [count][L][N]
// (CKKS); BFV needs at least size * L * N words of workspace (3 * L * N * count from
// decrypt_many_workspace takes the batch at once), CKKS none.
size_t decrypt_many_workspace(const Context &c, uint32_t L, uint32_t count);
void decrypt_many(const Context &c, const uint64_t *ct, uint32_t size, uint32_t count, const uint64_t *sk, uint32_t L,
                  uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st);

// BatchEncoder (src/batchencoder.cpp:104-200): values [count <= N] -> plain [N] mod t,
// and back.  decode's output has N slots.
void batch_encode(const Context &c, const uint64_t *values, uint32_t count, uint64_t *plain, hipStream_t st);
void batch_decode(const Context &c, const uint64_t *plain, uint64_t *values, uint64_t *scratch, hipStream_t st);
// Batches: plaintext g from values + g * values_stride (count values each) into
// plain [plaintexts][N]; decode plain [plaintexts][N] into values [plaintexts][N]
// through scratch [plaintexts][N].
void batch_encode_many(const Context &c, const uint64_t *values, uint32_t count, size_t values_stride,
                       uint32_t plaintexts, uint64_t *plain, hipStream_t st);
void batch_decode_many(const Context &c, const uint64_t *plain, uint32_t plaintexts, uint64_t *values,
                       uint64_t *scratch, hipStream_t st);

}  // namespace shc
