// crypto.hip -- samplers, key generation, encryption, decryption, BatchEncoder.
#include "crypto.h"
#include "galois.h"
#include "modarith.h"
#include "ntt.h"
#include "rns_ops.h"

namespace shc {
namespace {

constexpr int TPB = 256;

inline uint32_t blocks_for(size_t n) {
    const size_t b = (n + TPB - 1) / TPB;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// i-th output of splitmix64 seeded with `seed`
__host__ __device__ __forceinline__ uint64_t stream_word(uint64_t seed, uint64_t i) {
    return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
}

// Independent sub-stream seeds (u, e0, e1, per-digit keys, ...)
inline uint64_t sub_seed(uint64_t seed, uint64_t k) { return mix64(seed ^ mix64(k + 0x5EA15EEDull)); }

// ---- keyed streams: ChaCha20 (20 rounds, 256-bit key, 64-bit block counter, 64-bit
// nonce; state words 12-15 = counter lo, hi, nonce lo, hi).  Word i of the stream of
// seed s is words 2k, 2k+1 (little endian) of block i / 8 under nonce s, k = i % 8.
struct SamplerKey {
    uint32_t k[8];
};

__host__ __device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

#define SHC_QR(a, b, c, d)                 \
    a += b; d ^= a; d = rotl32(d, 16);     \
    c += d; b ^= c; b = rotl32(b, 12);     \
    a += b; d ^= a; d = rotl32(d, 8);      \
    c += d; b ^= c; b = rotl32(b, 7);

__device__ __forceinline__ void chacha20_block(const SamplerKey &key, uint64_t counter, uint64_t nonce,
                                               uint32_t (&out)[16]) {
    uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key.k[0], key.k[1], key.k[2], key.k[3], key.k[4], key.k[5], key.k[6], key.k[7],
                      (uint32_t)counter, (uint32_t)(counter >> 32), (uint32_t)nonce, (uint32_t)(nonce >> 32)};
    uint32_t y[16];
#pragma unroll
    for (int i = 0; i < 16; i++) y[i] = x[i];
#pragma unroll 1
    for (int r = 0; r < 10; r++) {
        SHC_QR(y[0], y[4], y[8], y[12]) SHC_QR(y[1], y[5], y[9], y[13])
        SHC_QR(y[2], y[6], y[10], y[14]) SHC_QR(y[3], y[7], y[11], y[15])
        SHC_QR(y[0], y[5], y[10], y[15]) SHC_QR(y[1], y[6], y[11], y[12])
        SHC_QR(y[2], y[7], y[8], y[13]) SHC_QR(y[3], y[4], y[9], y[14])
    }
#pragma unroll
    for (int i = 0; i < 16; i++) out[i] = y[i] + x[i];
}
#undef SHC_QR

__device__ __forceinline__ uint64_t block_word(const uint32_t (&b)[16], int k) {
    return (uint64_t)b[2 * k] | ((uint64_t)b[2 * k + 1] << 32);
}

// A ternary or centred-binomial integer from one stream word (shared by both streams).
__device__ __forceinline__ int64_t small_from_word(int kind, uint64_t w) {
    if (kind == SAMPLE_TERNARY) return (int64_t)mulhi64(w, 3) - 1;  // floor(3 w / 2^64) in {0, 1, 2}
    // centred binomial with 21 + 21 bits: variance 21/2 (sigma ~ 3.24), |s| <= 21
    return (int64_t)__popcll(w & 0x1FFFFFull) - (int64_t)__popcll((w >> 21) & 0x1FFFFFull);
}

// One thread per keystream block: coefficients 8b .. 8b+7.
__global__ void sample_small_keyed_kernel(int kind, SamplerKey key, uint64_t seed,
                                          const DevPrime *__restrict__ primes, uint32_t L, int logn,
                                          uint64_t *__restrict__ out) {
    const size_t n = (size_t)1 << logn;
    const size_t blocks = (n + 7) / 8;
    for (size_t b = (size_t)blockIdx.x * TPB + threadIdx.x; b < blocks; b += (size_t)gridDim.x * TPB) {
        uint32_t ks[16];
        chacha20_block(key, b, seed, ks);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const size_t c = b * 8 + k;
            if (c >= n) break;
            const int64_t s = small_from_word(kind, block_word(ks, k));
            for (uint32_t l = 0; l < L; l++) {
                const uint64_t q = primes[l].q;
                out[(size_t)l * n + c] = s < 0 ? q - (uint64_t)(-s) : (uint64_t)s;
            }
        }
    }
}

// One thread per keystream block: outputs 4b .. 4b+3, output i from the 128-bit draw
// (word 2i high, word 2i+1 low) reduced modulo its limb's prime.
__global__ void sample_uniform_keyed_kernel(SamplerKey key, uint64_t seed, const DevPrime *__restrict__ primes,
                                            int logn, size_t total, uint64_t *__restrict__ out) {
    const size_t blocks = (total + 3) / 4;
    for (size_t b = (size_t)blockIdx.x * TPB + threadIdx.x; b < blocks; b += (size_t)gridDim.x * TPB) {
        uint32_t ks[16];
        chacha20_block(key, b, seed, ks);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const size_t i = b * 4 + k;
            if (i >= total) break;
            const DevPrime &P = primes[i >> logn];
            const u128 x{block_word(ks, 2 * k + 1), block_word(ks, 2 * k)};
            out[i] = barrett128(x, P.q, P.ratio_lo, P.ratio_hi);
        }
    }
}

__global__ void sample_small_kernel(int kind, uint64_t seed, const DevPrime *__restrict__ primes, uint32_t L,
                                    int logn, uint64_t *__restrict__ out) {
    const size_t n = (size_t)1 << logn;
    for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < n; c += (size_t)gridDim.x * TPB) {
        const int64_t s = small_from_word(kind, stream_word(seed, c));
        for (uint32_t l = 0; l < L; l++) {
            const uint64_t q = primes[l].q;
            out[(size_t)l * n + c] = s < 0 ? q - (uint64_t)(-s) : (uint64_t)s;
        }
    }
}

__global__ void sample_uniform_kernel(uint64_t seed, const DevPrime *__restrict__ primes, int logn, size_t total,
                                      uint64_t *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (size_t)gridDim.x * TPB) {
        const DevPrime &P = primes[i >> logn];
        // a 128-bit draw reduced modulo q: bias below q / 2^128
        const u128 x{stream_word(seed, 2 * i + 1), stream_word(seed, 2 * i)};
        out[i] = barrett128(x, P.q, P.ratio_lo, P.ratio_hi);
    }
}

enum Op { OP_MUL, OP_ADD, OP_NEG_MULADD, OP_MULADD, OP_NEG };

// out[l][c] (op) a[l][c], b[l][c] with limb strides; l uses primes[l].  out may alias
// a or b (each element is read before it is written by the same thread).
__global__ void limb_op_kernel(int op, const uint64_t *a, size_t a_ls, const uint64_t *b, size_t b_ls, uint64_t *out,
                               size_t o_ls, const DevPrime *__restrict__ primes, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (size_t)gridDim.x * TPB) {
        const size_t l = i >> logn, c = i & (n - 1);
        const DevPrime &P = primes[l];
        const uint64_t x = a[l * a_ls + c];
        uint64_t &o = out[l * o_ls + c];
        switch (op) {
            case OP_MUL: o = mulmod(x, b[l * b_ls + c], P.q, P.ratio_lo, P.ratio_hi); break;
            case OP_ADD: o = addmod(x, b[l * b_ls + c], P.q); break;
            case OP_MULADD: o = addmod(o, mulmod(x, b[l * b_ls + c], P.q, P.ratio_lo, P.ratio_hi), P.q); break;
            case OP_NEG: o = negmod(x, P.q); break;
            case OP_NEG_MULADD:
                o = negmod(addmod(mulmod(x, b[l * b_ls + c], P.q, P.ratio_lo, P.ratio_hi), o, P.q), P.q);
                break;
        }
    }
}

void limb_op(Op op, const uint64_t *a, size_t a_ls, const uint64_t *b, size_t b_ls, uint64_t *out, size_t o_ls,
             const DevPrime *primes, uint32_t L, int logn, hipStream_t st) {
    const size_t total = (size_t)L << logn;
    limb_op_kernel<<<blocks_for(total), TPB, 0, st>>>(op, a, a_ls, b, b_ls, out, o_ls, primes, logn, total);
    launch_check("limb_op");
}

// key[j][0][j][c] += new_key[j][c] * (P mod q_j)
__global__ void kswitch_fold_kernel(uint64_t *__restrict__ key, const uint64_t *__restrict__ new_key,
                                    const uint64_t *__restrict__ factors, const DevPrime *__restrict__ primes,
                                    uint32_t K, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (size_t)gridDim.x * TPB) {
        const size_t j = i >> logn, c = i & (n - 1);
        const uint64_t q = primes[j].q;
        uint64_t &dst = key[((j * 2) * K + j) * n + c];
        dst = addmod(dst, shoup(new_key[j * n + c], factors[2 * j], factors[2 * j + 1], q), q);
    }
}

// floor(x / t) and x mod t for x < t * 2^64 (quotient fits one word)
__device__ __forceinline__ uint64_t div128_small(u128 x, uint64_t t, uint64_t r_lo, uint64_t r_hi) {
    uint64_t carry = mulhi64(x.lo, r_lo);
    u128 t2 = mul64x64(x.lo, r_hi);
    uint64_t tmp1 = t2.lo + carry;
    uint64_t tmp3 = t2.hi + (tmp1 < carry ? 1ull : 0ull);
    t2 = mul64x64(x.hi, r_lo);
    uint64_t s = tmp1 + t2.lo;
    carry = t2.hi + (s < tmp1 ? 1ull : 0ull);
    uint64_t est = x.hi * r_hi + tmp3 + carry;
    const uint64_t r = x.lo - est * t;
    return r >= t ? est + 1 : est;
}

// c0[l][c] += m Delta_l + floor((m (Q mod t) + floor((t+1)/2)) / t)   (mod q_l)
__global__ void bfv_scale_add_kernel(uint64_t *__restrict__ c0, const uint64_t *__restrict__ plain, BehzDev B,
                                     const DevPrime *__restrict__ qp, int logn, bool subtract) {
    const size_t n = (size_t)1 << logn;
    const uint64_t t = B.t, half = (t + 1) >> 1;
    for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < n; c += (size_t)gridDim.x * TPB) {
        const uint64_t m = plain[c];
        u128 num = mul64x64(m, B.q_mod_t);
        add128(num, u128{half, 0});
        const uint64_t fix = div128_small(num, t, B.t_ratio_lo, B.t_ratio_hi);
        for (uint32_t l = 0; l < B.L; l++) {
            const DevPrime &Q = qp[l];
            const uint64_t v = addmod(shoup(m, B.delta[2 * l], B.delta[2 * l + 1], Q.q), barrett64(fix, Q.q, Q.ratio_hi),
                                      Q.q);
            c0[l * n + c] = subtract ? submod(c0[l * n + c], v, Q.q) : addmod(c0[l * n + c], v, Q.q);
        }
    }
}

// phase[l] = c0[l] + sum_{i>=1} ct_i[l] s^i (all operands in the same form)
__global__ void phase_kernel(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ cs, size_t c_stride,
                             uint32_t size, const uint64_t *__restrict__ sk, uint64_t *__restrict__ phase,
                             const DevPrime *__restrict__ primes, int logn, size_t total) {
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < total; i += (size_t)gridDim.x * TPB) {
        const DevPrime &P = primes[i >> logn];
        const uint64_t s = sk[i];
        uint64_t acc = 0, sp = s;
        for (uint32_t k = 1; k < size; k++) {
            acc = addmod(acc, mulmod(cs[(k - 1) * c_stride + i], sp, P.q, P.ratio_lo, P.ratio_hi), P.q);
            sp = mulmod(sp, s, P.q, P.ratio_lo, P.ratio_hi);
        }
        phase[i] = c0 ? addmod(acc, c0[i], P.q) : acc;
    }
}

// RNSTool::decryptScaleAndRound (src/utils/rns.cpp:1039-1095) per coefficient
__global__ void scale_round_kernel(const uint64_t *__restrict__ phase, uint64_t *__restrict__ out, BehzDev B,
                                   const DevPrime *__restrict__ qp, int logn) {
    const size_t n = (size_t)1 << logn;
    const uint64_t t = B.t, g = B.gamma;
    for (size_t c = (size_t)blockIdx.x * TPB + threadIdx.x; c < n; c += (size_t)gridDim.x * TPB) {
        u128 at{0, 0}, ag{0, 0};
        for (uint32_t j = 0; j < B.L; j++) {
            const uint64_t y = shoup(phase[j * n + c], B.tg_qhat_inv[2 * j], B.tg_qhat_inv[2 * j + 1], qp[j].q);
            add128(at, mul64x64(y, B.qhat_mod_t[j]));
            add128(ag, mul64x64(y, B.qhat_mod_g[j]));
        }
        uint64_t rt = barrett128(at, t, B.t_ratio_lo, B.t_ratio_hi);
        uint64_t rg = barrett128(ag, g, B.g_ratio_lo, B.g_ratio_hi);
        rt = shoup(rt, B.neg_inv_q_t, B.neg_inv_q_t_shoup, t);
        rg = shoup(rg, B.neg_inv_q_g, B.neg_inv_q_g_shoup, g);
        uint64_t d;
        if (rg > (g >> 1)) d = addmod(rt, barrett64(g - rg, t, B.t_ratio_hi), t);
        else d = submod(rt, barrett64(rg, t, B.t_ratio_hi), t);
        if (d) d = shoup(d, B.inv_gamma_t, B.inv_gamma_t_shoup, t);
        out[c] = d;
    }
}

// BatchEncoder::populateMatrixRepsIndexMap (src/batchencoder.cpp:64-84), evaluated per slot
__device__ __forceinline__ uint32_t slot_index(uint32_t i, int logn) {
    const uint64_t n = 1ull << logn, row = n >> 1, mask = 2 * n - 1;
    const uint32_t e = i < row ? i : i - (uint32_t)row;
    uint64_t pos = 1, b = 3;
    for (uint32_t k = e; k; k >>= 1) {
        if (k & 1) pos = (pos * b) & mask;
        b = (b * b) & mask;
    }
    const uint64_t idx = i < row ? (pos - 1) >> 1 : (2 * n - pos - 1) >> 1;
    return __brev((uint32_t)idx) >> (32 - logn);
}

__global__ void batch_scatter_kernel(const uint64_t *__restrict__ values, uint32_t count, uint64_t *__restrict__ plain,
                                     int logn) {
    const uint32_t n = 1u << logn;
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB)
        plain[slot_index(i, logn)] = i < count ? values[i] : 0;
}

__global__ void batch_gather_kernel(const uint64_t *__restrict__ ntt, uint64_t *__restrict__ values, int logn) {
    const uint32_t n = 1u << logn;
    for (uint32_t i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) values[i] = ntt[slot_index(i, logn)];
}

}  // namespace

void sample_poly(const Context &c, SampleKind kind, uint64_t seed, const DevPrime *primes, uint32_t L, uint64_t *out,
                 hipStream_t st) {
    if (c.sampler_keyed) {
        SamplerKey key;
        for (int i = 0; i < 8; i++) key.k[i] = c.sampler_key[i];
        if (kind == SAMPLE_UNIFORM) {
            const size_t total = (size_t)L << c.logn;
            sample_uniform_keyed_kernel<<<blocks_for((total + 3) / 4), TPB, 0, st>>>(key, seed, primes, c.logn, total,
                                                                                    out);
        } else {
            sample_small_keyed_kernel<<<blocks_for((c.n + 7) / 8), TPB, 0, st>>>((int)kind, key, seed, primes, L,
                                                                               c.logn, out);
        }
        launch_check("sample");
        return;
    }
    if (kind == SAMPLE_UNIFORM) {
        const size_t total = (size_t)L << c.logn;
        sample_uniform_kernel<<<blocks_for(total), TPB, 0, st>>>(seed, primes, c.logn, total, out);
    } else {
        sample_small_kernel<<<blocks_for(c.n), TPB, 0, st>>>((int)kind, seed, primes, L, c.logn, out);
    }
    launch_check("sample");
}

void keygen_secret(const Context &c, uint64_t seed, uint64_t *sk, hipStream_t st) {
    // KeyGenerator::generateSk (src/keygenerator.cpp:56-80): ternary, then NTT
    const uint32_t K = (uint32_t)c.key_count();
    sample_poly(c, SAMPLE_TERNARY, seed, c.primes(), K, sk, st);
    ntt_forward(sk, K, K, c.primes(), c.logn, NTT_FULL, st);
}

void encrypt_zero_symmetric(const Context &c, const uint64_t *sk, const DevPrime *primes, uint32_t L, bool ntt_form,
                            uint64_t seed, uint64_t *out, hipStream_t st) {
    const size_t n = c.n;
    uint64_t *c0 = out, *c1 = out + (size_t)L * n;
    sample_poly(c, SAMPLE_UNIFORM, sub_seed(seed, 0), primes, L, c1, st);
    sample_poly(c, SAMPLE_CBD, sub_seed(seed, 1), primes, L, c0, st);
    if (ntt_form) {
        ntt_forward(c0, L, L, primes, c.logn, NTT_FULL, st);
        limb_op(OP_NEG_MULADD, c1, n, sk, n, c0, n, primes, L, c.logn, st);  // c0 = -(a s + e)
    } else {
        // coefficient form: c0 = -(INTT(a s) + e), c1 = INTT(a); c1 holds a*s first
        limb_op(OP_MUL, c1, n, sk, n, c1, n, primes, L, c.logn, st);
        ntt_inverse(c1, L, L, primes, c.logn, NTT_FULL, st);
        limb_op(OP_ADD, c1, n, c0, n, c0, n, primes, L, c.logn, st);
        // c1 <- a again: regenerate the same uniform words and transform them
        sample_poly(c, SAMPLE_UNIFORM, sub_seed(seed, 0), primes, L, c1, st);
        ntt_inverse(c1, L, L, primes, c.logn, NTT_FULL, st);
        limb_op(OP_NEG, c0, n, nullptr, 0, c0, n, primes, L, c.logn, st);
    }
}

void keygen_public(const Context &c, const uint64_t *sk, uint64_t seed, uint64_t *pk, hipStream_t st) {
    // KeyGenerator::generatePk (src/keygenerator.cpp:82-106)
    encrypt_zero_symmetric(c, sk, c.primes(), (uint32_t)c.key_count(), true, seed, pk, st);
}

void keygen_kswitch(const Context &c, const uint64_t *sk, const uint64_t *new_key, uint64_t seed, uint64_t *key,
                    hipStream_t st) {
    const uint32_t K = (uint32_t)c.key_count();
    if (K < 2) fail(ERR_LOGIC, "keyswitching is not supported by the context");
    const size_t n = c.n;
    for (uint32_t j = 0; j + 1 < K; j++)
        encrypt_zero_symmetric(c, sk, c.primes(), K, true, sub_seed(seed, 100 + j), key + (size_t)j * 2 * K * n, st);
    const size_t total = (size_t)(K - 1) << c.logn;
    kswitch_fold_kernel<<<blocks_for(total), TPB, 0, st>>>(key, new_key, c.d_kswitch_factors.as<uint64_t>(),
                                                           c.primes(), K, c.logn, total);
    launch_check("kswitch_fold");
}

size_t keygen_workspace(const Context &c) { return c.key_count() * c.n * 8; }

void keygen_relin(const Context &c, const uint64_t *sk, uint64_t seed, uint64_t *key, void *ws, size_t ws_bytes,
                  hipStream_t st) {
    // KeyGenerator::createRelinKeys: key for s^2 (src/keygenerator.cpp:108-150)
    if (!ws || ws_bytes < keygen_workspace(c)) fail(ERR_INVALID_ARGUMENT, "workspace too small for key generation");
    const uint32_t K = (uint32_t)c.key_count();
    uint64_t *s2 = static_cast<uint64_t *>(ws);
    limb_op(OP_MUL, sk, c.n, sk, c.n, s2, c.n, c.primes(), K, c.logn, st);
    keygen_kswitch(c, sk, s2, seed, key, st);
}

void keygen_galois(const Context &c, const uint64_t *sk, uint32_t elt, uint64_t seed, uint64_t *key, void *ws,
                   size_t ws_bytes, hipStream_t st) {
    // KeyGenerator::createGaloisKeys: key for sigma_elt(s) in NTT form
    if (!ws || ws_bytes < keygen_workspace(c)) fail(ERR_INVALID_ARGUMENT, "workspace too small for key generation");
    const uint32_t K = (uint32_t)c.key_count();
    uint64_t *ss = static_cast<uint64_t *>(ws);
    apply_galois(c, sk, (size_t)K * c.n, ss, (size_t)K * c.n, 1, K, elt, true, st);
    keygen_kswitch(c, sk, ss, seed, key, st);
}

static uint32_t upper_limbs(const Context &c, uint32_t L) { return L + 1 <= c.key_count() ? L + 1 : L; }

size_t encrypt_workspace(const Context &c, uint32_t L) {
    const uint32_t L1 = upper_limbs(c, L);
    const size_t limb = c.n * 8;
    // u, e0, e1 (sampled), NTT(u), z[2], e scratch, rescale scratch
    return 7 * L1 * limb + rescale_workspace(c, 1, 2, L1) + 512;
}

void encrypt_zero_asymmetric(const Context &c, const uint64_t *pk, const uint64_t *u, const uint64_t *e0,
                             const uint64_t *e1, uint32_t L, uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st) {
    if (L < 1 || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "level out of range for encryption");
    if (!ws || ws_bytes < encrypt_workspace(c, L)) fail(ERR_INVALID_ARGUMENT, "workspace too small for encryption");
    const uint32_t L1 = upper_limbs(c, L), K = (uint32_t)c.key_count();
    const size_t n = c.n;
    const bool ckks = c.scheme == SCHEME_CKKS;
    const DevPrime *primes = c.primes();
    uint64_t *w = static_cast<uint64_t *>(ws) + 3 * (size_t)L1 * n;  // skip the sampling slots
    uint64_t *un = w, *z = w + (size_t)L1 * n, *et = w + 3 * (size_t)L1 * n;
    void *rws = w + 4 * (size_t)L1 * n;
    SHC_HIP(hipMemcpyAsync(un, u, (size_t)L1 * n * 8, hipMemcpyDeviceToDevice, st));
    ntt_forward(un, L1, L1, primes, c.logn, NTT_FULL, st);
    const uint64_t *es[2] = {e0, e1};
    uint64_t *dst = (L1 == L) ? out : z;
    for (int j = 0; j < 2; j++) {
        uint64_t *zj = dst + (size_t)j * L1 * n;
        limb_op(OP_MUL, un, n, pk + (size_t)j * K * n, n, zj, n, primes, L1, c.logn, st);
        if (ckks) {
            SHC_HIP(hipMemcpyAsync(et, es[j], (size_t)L1 * n * 8, hipMemcpyDeviceToDevice, st));
            ntt_forward(et, L1, L1, primes, c.logn, NTT_FULL, st);
            limb_op(OP_ADD, zj, n, et, n, zj, n, primes, L1, c.logn, st);
        } else {
            ntt_inverse(zj, L1, L1, primes, c.logn, NTT_FULL, st);
            limb_op(OP_ADD, zj, n, es[j], n, zj, n, primes, L1, c.logn, st);
        }
    }
    if (L1 != L) rescale_to_next(c, z, out, 1, 2, L1, rws, rescale_workspace(c, 1, 2, L1), st);
}

void bfv_scale_add(const Context &c, uint64_t *c0, const uint64_t *plain, uint32_t L, bool subtract, hipStream_t st) {
    const BehzDev &B = c.behz(L).dev;
    bfv_scale_add_kernel<<<blocks_for(c.n), TPB, 0, st>>>(c0, plain, B, c.primes(), c.logn, subtract);
    launch_check("bfv_scale_add");
}

void bfv_add_plain_scaled(const Context &c, uint64_t *c0, const uint64_t *plain, uint32_t L, hipStream_t st) {
    bfv_scale_add(c, c0, plain, L, false, st);
}

void encrypt(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t L, uint64_t seed, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st) {
    // Encryptor::encryptInternal (src/encryptor.cpp:152-233)
    if (!ws || ws_bytes < encrypt_workspace(c, L)) fail(ERR_INVALID_ARGUMENT, "workspace too small for encryption");
    const uint32_t L1 = upper_limbs(c, L);
    uint64_t *u = static_cast<uint64_t *>(ws), *e0 = u + (size_t)L1 * c.n, *e1 = e0 + (size_t)L1 * c.n;
    sample_poly(c, SAMPLE_TERNARY, sub_seed(seed, 0), c.primes(), L1, u, st);
    sample_poly(c, SAMPLE_CBD, sub_seed(seed, 1), c.primes(), L1, e0, st);
    sample_poly(c, SAMPLE_CBD, sub_seed(seed, 2), c.primes(), L1, e1, st);
    encrypt_zero_asymmetric(c, pk, u, e0, e1, L, out, ws, ws_bytes, st);
    if (!plain) return;
    if (c.scheme == SCHEME_BFV) bfv_add_plain_scaled(c, out, plain, L, st);
    else limb_op(OP_ADD, out, c.n, plain, c.n, out, c.n, c.primes(), L, c.logn, st);
}

size_t decrypt_workspace(const Context &c, uint32_t L) { return (size_t)4 * L * c.n * 8; }

void decrypt(const Context &c, const uint64_t *ct, uint32_t size, const uint64_t *sk, uint32_t L, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st) {
    if (size < 2 || size > 3) fail(ERR_INVALID_ARGUMENT, "ciphertext size must be 2 or 3");
    const size_t n = c.n, poly = (size_t)L * n, total = poly;
    const DevPrime *primes = c.primes();
    if (c.scheme == SCHEME_CKKS) {
        phase_kernel<<<blocks_for(total), TPB, 0, st>>>(ct, ct + poly, poly, size, sk, out, primes, c.logn, total);
        launch_check("phase");
        return;
    }
    if (!ws || ws_bytes < decrypt_workspace(c, L)) fail(ERR_INVALID_ARGUMENT, "workspace too small for decryption");
    uint64_t *tmp = static_cast<uint64_t *>(ws), *phase = tmp + 2 * poly;
    SHC_HIP(hipMemcpyAsync(tmp, ct + poly, (size - 1) * poly * 8, hipMemcpyDeviceToDevice, st));
    ntt_forward(tmp, (size - 1) * L, L, primes, c.logn, NTT_FULL, st);
    phase_kernel<<<blocks_for(total), TPB, 0, st>>>(nullptr, tmp, poly, size, sk, phase, primes, c.logn, total);
    launch_check("phase");
    ntt_inverse(phase, L, L, primes, c.logn, NTT_FULL, st);
    limb_op(OP_ADD, phase, n, ct, n, phase, n, primes, L, c.logn, st);
    scale_round_kernel<<<blocks_for(n), TPB, 0, st>>>(phase, out, c.behz(L).dev, primes, c.logn);
    launch_check("scale_round");
}

void batch_encode(const Context &c, const uint64_t *values, uint32_t count, uint64_t *plain, hipStream_t st) {
    if (!c.plain_prime()) fail(ERR_LOGIC, "encryption parameters are not valid for batching");
    if (count > c.n) fail(ERR_INVALID_ARGUMENT, "values_matrix size is too large");
    batch_scatter_kernel<<<blocks_for(c.n), TPB, 0, st>>>(values, count, plain, c.logn);
    launch_check("batch_scatter");
    ntt_inverse(plain, 1, 1, c.plain_prime(), c.logn, NTT_FULL, st);
}

void batch_decode(const Context &c, const uint64_t *plain, uint64_t *values, uint64_t *scratch, hipStream_t st) {
    if (!c.plain_prime()) fail(ERR_LOGIC, "encryption parameters are not valid for batching");
    SHC_HIP(hipMemcpyAsync(scratch, plain, c.n * 8, hipMemcpyDeviceToDevice, st));
    ntt_forward(scratch, 1, 1, c.plain_prime(), c.logn, NTT_FULL, st);
    batch_gather_kernel<<<blocks_for(c.n), TPB, 0, st>>>(scratch, values, c.logn);
    launch_check("batch_gather");
}

}  // namespace shc
