// crypto.hip -- samplers, key generation, encryption, decryption, BatchEncoder.
#include "crypto.h"
#include "galois.h"
#include "modarith.h"
#include "ntt.h"
#include "rns_ops.h"
#include <algorithm>

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
__host__ __device__ __forceinline__ uint64_t sub_seed(uint64_t seed, uint64_t k) {
    return mix64(seed ^ mix64(k + 0x5EA15EEDull));
}

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

// c0[l][c] += m Delta_l + floor((m (Q mod t) + floor((t+1)/2)) / t)   (mod q_l), for
// `count` pairs (c0 + g * c0_stride, plain + g * n).
__global__ void bfv_scale_add_kernel(uint64_t *__restrict__ c0_base, size_t c0_stride,
                                     const uint64_t *__restrict__ plain_base, uint32_t count, BehzDev B,
                                     const DevPrime *__restrict__ qp, int logn, bool subtract) {
    const size_t n = (size_t)1 << logn, total = (size_t)count << logn;
    const uint64_t t = B.t, half = (t + 1) >> 1;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t g = idx >> logn, c = idx & (n - 1);
        uint64_t *__restrict__ c0 = c0_base + g * c0_stride;
        const uint64_t m = plain_base[idx];
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

// phase[g][l] = c0[g][l] + sum_{k>=1} ct_k[g][l] s^k (all operands in the same form) for
// `count` ciphertexts of `poly` = L * N words per polynomial: c0 of ciphertext g at
// c0 + g * in_stride, its c_k at cs + g * in_stride + (k - 1) * c_stride, its phase at
// phase + g * poly.  c0 == nullptr: the sum alone.
__global__ void phase_kernel(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ cs, size_t c_stride,
                             size_t in_stride, uint32_t size, const uint64_t *__restrict__ sk,
                             uint64_t *__restrict__ phase, const DevPrime *__restrict__ primes, int logn, uint32_t L,
                             size_t total) {
    const size_t poly = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t r = uniform_row(idx, logn), g = r / L;
        const size_t i = idx - g * poly, in = g * in_stride + i;
        const DevPrime &P = primes[r - g * L];
        const uint64_t s = sk[i];
        uint64_t acc = 0, sp = s;
        for (uint32_t k = 1; k < size; k++) {
            acc = addmod(acc, mulmod(cs[(k - 1) * c_stride + in], sp, P.q, P.ratio_lo, P.ratio_hi), P.q);
            sp = mulmod(sp, s, P.q, P.ratio_lo, P.ratio_hi);
        }
        phase[idx] = c0 ? addmod(acc, c0[in], P.q) : acc;
    }
}

// phase_kernel with two words per thread and 16-byte accesses (every pointer and
// stride 16-byte aligned; a pair never straddles two limbs as N is even).
__global__ void phase_pairs_kernel(const uint64_t *__restrict__ c0, const uint64_t *__restrict__ cs, size_t c_stride,
                                   size_t in_stride, uint32_t size, const uint64_t *__restrict__ sk,
                                   uint64_t *__restrict__ phase, const DevPrime *__restrict__ primes, int logn,
                                   uint32_t L, size_t pairs) {
    const size_t poly = (size_t)L << logn;
    for (size_t p = (size_t)blockIdx.x * TPB + threadIdx.x; p < pairs; p += (size_t)gridDim.x * TPB) {
        const size_t w = 2 * p;
        uint32_t r = (uint32_t)(w >> logn);
        if (logn >= 7) r = __builtin_amdgcn_readfirstlane(r);  // a wave's 128 words in one limb
        const uint32_t g = r / L;
        const size_t i = w - g * poly, in = g * in_stride + i;
        const DevPrime &P = primes[r - g * L];
        const ulonglong2 s = *reinterpret_cast<const ulonglong2 *>(sk + i);
        uint64_t acc0 = 0, acc1 = 0, sp0 = s.x, sp1 = s.y;
        for (uint32_t k = 1; k < size; k++) {
            const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(cs + (k - 1) * c_stride + in);
            acc0 = addmod(acc0, mulmod(x.x, sp0, P.q, P.ratio_lo, P.ratio_hi), P.q);
            acc1 = addmod(acc1, mulmod(x.y, sp1, P.q, P.ratio_lo, P.ratio_hi), P.q);
            if (k + 1 < size) {
                sp0 = mulmod(sp0, s.x, P.q, P.ratio_lo, P.ratio_hi);
                sp1 = mulmod(sp1, s.y, P.q, P.ratio_lo, P.ratio_hi);
            }
        }
        if (c0) {
            const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(c0 + in);
            acc0 = addmod(acc0, x.x, P.q);
            acc1 = addmod(acc1, x.y, P.q);
        }
        *reinterpret_cast<ulonglong2 *>(phase + w) = make_ulonglong2(acc0, acc1);
    }
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// phase_kernel's launch, in pairs where the pointers allow
void launch_phase(const uint64_t *c0, const uint64_t *cs, size_t c_stride, size_t in_stride, uint32_t size,
                  const uint64_t *sk, uint64_t *phase, const DevPrime *primes, int logn, uint32_t L, size_t total,
                  hipStream_t st) {
    if (logn >= 1 && (!c0 || aligned16(c0)) && aligned16(cs) && aligned16(sk) && aligned16(phase)) {
        phase_pairs_kernel<<<blocks_for(total / 2), TPB, 0, st>>>(c0, cs, c_stride, in_stride, size, sk, phase, primes,
                                                                  logn, L, total / 2);
    } else {
        phase_kernel<<<blocks_for(total), TPB, 0, st>>>(c0, cs, c_stride, in_stride, size, sk, phase, primes, logn, L,
                                                        total);
    }
    launch_check("phase");
}

// RNSTool::decryptScaleAndRound (src/utils/rns.cpp:1039-1095) per coefficient of
// phase[h] + c0[h] (coefficient form, L = B.L limbs) for `count` ciphertexts: phase of
// ciphertext h at phase + h * L * N, its c0 at c0 + h * c0_stride, output at out + h * N.
__global__ void scale_round_kernel(const uint64_t *__restrict__ phase, const uint64_t *__restrict__ c0,
                                   size_t c0_stride, uint64_t *__restrict__ out, uint32_t count, BehzDev B,
                                   const DevPrime *__restrict__ qp, int logn) {
    const size_t n = (size_t)1 << logn, total = (size_t)count << logn;
    const uint64_t t = B.t, g = B.gamma;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t h = idx >> logn, c = idx & (n - 1);
        const uint64_t *ph = phase + h * B.L * n, *x0 = c0 + h * c0_stride;
        u128 at{0, 0}, ag{0, 0};
        for (uint32_t j = 0; j < B.L; j++) {
            const uint64_t v = addmod(ph[j * n + c], x0[j * n + c], qp[j].q);
            const uint64_t y = shoup(v, B.tg_qhat_inv[2 * j], B.tg_qhat_inv[2 * j + 1], qp[j].q);
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
        out[idx] = d;
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

// plain[g][slot_index(i)] = values[g * stride + i] (0 past count), g < plaintexts
__global__ void batch_scatter_kernel(const uint64_t *__restrict__ values, uint32_t count, size_t stride,
                                     uint64_t *__restrict__ plain, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t g = idx >> logn;
        const uint32_t i = (uint32_t)(idx & (n - 1));
        plain[g * n + slot_index(i, logn)] = i < count ? values[g * stride + i] : 0;
    }
}

// values[g][i] = ntt[g][slot_index(i)]
__global__ void batch_gather_kernel(const uint64_t *__restrict__ ntt, uint64_t *__restrict__ values, int logn,
                                    size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t g = idx >> logn;
        values[idx] = ntt[g * n + slot_index((uint32_t)(idx & (n - 1)), logn)];
    }
}


// ---- key-switching keys, every digit in one launch per stage -----------------------
// Digit j (blockIdx.y) of a key draws the streams encrypt_zero_symmetric draws for seed_j
// = sub_seed(seed, 100 + j): a (uniform, K limbs) from sub_seed(seed_j, 0) into
// key[j][1], e (centred binomial) from sub_seed(seed_j, 1) into key[j][0] -- the words of
// the single-stream samplers above, word for word.
template <bool KEYED>
__global__ void keygen_uniform_kernel(SamplerKey key, uint64_t seed, const DevPrime *__restrict__ primes, uint32_t K,
                                      int logn, uint64_t *__restrict__ out) {
    const size_t total = (size_t)K << logn;
    const uint64_t s = sub_seed(sub_seed(seed, 100 + (uint64_t)blockIdx.y), 0);
    uint64_t *c1 = out + ((size_t)(2 * blockIdx.y + 1) * K << logn);
    const size_t units = KEYED ? (total + 3) / 4 : total;
    for (size_t u = (size_t)blockIdx.x * TPB + threadIdx.x; u < units; u += (size_t)gridDim.x * TPB) {
        if (KEYED) {
            uint32_t ks[16];
            chacha20_block(key, u, s, ks);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const size_t i = u * 4 + k;
                if (i >= total) break;
                const DevPrime &P = primes[i >> logn];
                c1[i] = barrett128(u128{block_word(ks, 2 * k + 1), block_word(ks, 2 * k)}, P.q, P.ratio_lo, P.ratio_hi);
            }
        } else {
            const DevPrime &P = primes[u >> logn];
            c1[u] = barrett128(u128{stream_word(s, 2 * u + 1), stream_word(s, 2 * u)}, P.q, P.ratio_lo, P.ratio_hi);
        }
    }
}

template <bool KEYED>
__global__ void keygen_cbd_kernel(SamplerKey key, uint64_t seed, const DevPrime *__restrict__ primes, uint32_t K,
                                  int logn, uint64_t *__restrict__ out) {
    const size_t n = (size_t)1 << logn;
    const uint64_t s = sub_seed(sub_seed(seed, 100 + (uint64_t)blockIdx.y), 1);
    uint64_t *c0 = out + ((size_t)(2 * blockIdx.y) * K << logn);
    const size_t units = KEYED ? (n + 7) / 8 : n;
    for (size_t u = (size_t)blockIdx.x * TPB + threadIdx.x; u < units; u += (size_t)gridDim.x * TPB) {
        if (KEYED) {
            uint32_t ks[16];
            chacha20_block(key, u, s, ks);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const size_t c = u * 8 + k;
                if (c >= n) break;
                const int64_t v = small_from_word(SAMPLE_CBD, block_word(ks, k));
                for (uint32_t l = 0; l < K; l++) {
                    const uint64_t q = primes[l].q;
                    c0[(size_t)l * n + c] = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
                }
            }
        } else {
            const int64_t v = small_from_word(SAMPLE_CBD, stream_word(s, u));
            for (uint32_t l = 0; l < K; l++) {
                const uint64_t q = primes[l].q;
                c0[(size_t)l * n + u] = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
            }
        }
    }
}

// key[j][0][l] = -(key[j][1][l] s[l] + key[j][0][l]) for every digit j < K - 1
// (c0 = -(a s + NTT(e)), the operation order of encrypt_zero_symmetric's NTT form).
__global__ void keygen_combine_kernel(uint64_t *__restrict__ key, const uint64_t *__restrict__ sk,
                                      const DevPrime *__restrict__ primes, uint32_t K, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t r = uniform_row(idx, logn), j = r / K, l = r - j * K;
        const size_t c = idx & (n - 1);
        uint64_t *c0 = key + (((size_t)j * 2 * K + l) << logn) + c;
        const uint64_t a = c0[(size_t)K << logn];
        const DevPrime &P = primes[l];
        *c0 = negmod(addmod(mulmod(a, sk[((size_t)l << logn) + c], P.q, P.ratio_lo, P.ratio_hi), *c0, P.q), P.q);
    }
}

// ---- batched public-key encryption ------------------------------------------------
// Ciphertext g of a batch is the one encrypt() gives for the seed seed + g: u from the
// stream of sub_seed(seed + g, 0) (ternary), e0 and e1 from sub_seed(seed + g, 1 / 2)
// (centred binomial).  With L1 = L + 1 limbs before the divide-and-round by p = q_L
// (the special prime at the top level), the CKKS words are
//   c_j[i] = p^-1 (pk_j[i] NTT_i(u) + NTT_i(e_j - r_j))  (mod q_i, i < L),
//   r_j    = ([INTT_p(pk_j[L] NTT_p(u)) + e_j + floor(p/2)]_p mod q_i) - floor(p/2),
// which equals dividing-and-rounding pk_j NTT(u) + NTT(e_j) at L1 limbs (the transform
// is linear modulo each q_i): L1 + 2 + 2L transforms per ciphertext instead of
// 3 L1 + 2 + 2L.  BFV stays in coefficient form: c_j[i] = p^-1 (INTT_i(pk_j[i] NTT_i(u))
// + e_j - r_j), r_j from limb L of the same sum.

// u lifted to the L1 limbs of U[g]; e0, e1 as signed integers in Es[g][0 / 1].
template <bool KEYED>
__global__ void sample_encrypt_kernel(SamplerKey key, uint64_t seed, const DevPrime *__restrict__ primes, uint32_t L1,
                                      int logn, size_t total, uint64_t *__restrict__ U, int64_t *__restrict__ Es) {
    const size_t n = (size_t)1 << logn;
    // per ciphertext: 2^ulog keystream blocks (8 coefficients each) or coefficients
    const int ulog = KEYED ? (logn > 3 ? logn - 3 : 0) : logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t g = idx >> ulog, b = idx & (((size_t)1 << ulog) - 1);
        uint64_t *Ug = U + g * L1 * n;
        int64_t *Eg = Es + g * 2 * n;
        for (int k = 0; k < 3; k++) {
            const uint64_t s = sub_seed(seed + g, (uint64_t)k);
            const int kind = k == 0 ? SAMPLE_TERNARY : SAMPLE_CBD;
            if (KEYED) {
                uint32_t ks[16];
                chacha20_block(key, b, s, ks);
#pragma unroll
                for (int w = 0; w < 8; w++) {
                    const size_t c = b * 8 + w;
                    if (c >= n) break;
                    const int64_t v = small_from_word(kind, block_word(ks, w));
                    if (k == 0) {
                        for (uint32_t l = 0; l < L1; l++) {
                            const uint64_t q = primes[l].q;
                            Ug[(size_t)l * n + c] = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
                        }
                    } else {
                        Eg[(size_t)(k - 1) * n + c] = v;
                    }
                }
            } else {
                const int64_t v = small_from_word(kind, stream_word(s, b));
                if (k == 0) {
                    for (uint32_t l = 0; l < L1; l++) {
                        const uint64_t q = primes[l].q;
                        Ug[(size_t)l * n + b] = v < 0 ? q - (uint64_t)(-v) : (uint64_t)v;
                    }
                } else {
                    Eg[(size_t)(k - 1) * n + b] = v;
                }
            }
        }
    }
}

// dst[r][k] = src[r * src_stride + k] for rows of row_limbs * N words (total = rows * row)
__global__ void gather_rows_kernel(const uint64_t *__restrict__ src, size_t src_stride, uint64_t *__restrict__ dst,
                                   uint32_t row_limbs, int logn, size_t total) {
    const size_t row = (size_t)row_limbs << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t r = uniform_row(idx, logn) / row_limbs;
        dst[idx] = src[r * src_stride + (idx - r * row)];
    }
}

__device__ __forceinline__ uint64_t lift_small(int64_t v, uint64_t q) { return v < 0 ? q - (uint64_t)(-v) : (uint64_t)v; }

// SEAL's rounding correction of a divide-and-round by p, from the coefficient-form last
// limb x (< p): ([x + floor(p/2)]_p mod q_i) - floor(p/2) mod q_i.
__device__ __forceinline__ uint64_t divround_corr(uint64_t x, uint64_t p, const DevPrime &Q, const DivRoundConst &d) {
    uint64_t v = x + (p >> 1);
    v = v >= p ? v - p : v;
    return submod(barrett64(v, Q.q, Q.ratio_hi), d.half_mod, Q.q);
}

// P[g][j][l] = pk_j[l0 + l] * U[g][l0 + l] (mod q_{l0 + l}) for l < lw; U and pk limbs
// in NTT form, U with L1 limbs per ciphertext, P with lw limbs per polynomial.
__global__ void enc_products_kernel(const uint64_t *__restrict__ pk, uint32_t K, const uint64_t *__restrict__ U,
                                    uint32_t L1, uint32_t l0, uint32_t lw, uint64_t *__restrict__ P,
                                    const DevPrime *__restrict__ primes, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (n - 1);
        const uint32_t r = uniform_row(idx, logn), gj = r / lw, l = r - gj * lw + l0;
        const size_t g = gj >> 1, j = gj & 1;
        const DevPrime &Q = primes[l];
        P[idx] = mulmod(pk[(j * K + l) * n + c], U[(g * L1 + l) * n + c], Q.q, Q.ratio_lo, Q.ratio_hi);
    }
}

// CKKS, before the forward transform: out[g][j][i] = e_j - r_j (mod q_i), i < L (r_j = 0
// when nothing is dropped); Pl[g][j] = INTT_p(pk_j[L] NTT_p(u)), coefficient form.
__global__ void enc_error_kernel(const int64_t *__restrict__ Es, const uint64_t *__restrict__ Pl,
                                 uint64_t *__restrict__ out, const DevPrime *__restrict__ primes,
                                 const DivRoundConst *__restrict__ dc, uint64_t p, uint32_t L, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (n - 1);
        const uint32_t r = uniform_row(idx, logn), gj = r / L, i = r - gj * L;
        const DevPrime &Q = primes[i];
        const int64_t e = Es[(size_t)gj * n + c];
        uint64_t x = lift_small(e, Q.q);
        if (Pl) x = submod(x, divround_corr(addmod(Pl[(size_t)gj * n + c], lift_small(e, p), p), p, Q, dc[i]), Q.q);
        out[idx] = x;
    }
}

// CKKS, after the forward transform of out: out[g][j][i] = (pk_j[i] U[g][i] + out) times
// p^-1 when a prime was dropped, plus plain[g][i] for j = 0 when plain != nullptr.
__global__ void enc_finish_ckks_kernel(const uint64_t *__restrict__ pk, uint32_t K, const uint64_t *__restrict__ U,
                                       uint32_t L1, uint64_t *__restrict__ out, const uint64_t *__restrict__ plain,
                                       const DevPrime *__restrict__ primes, const DivRoundConst *__restrict__ dc,
                                       bool drop, uint32_t L, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (n - 1);
        const uint32_t r = uniform_row(idx, logn), gj = r / L, i = r - gj * L;
        const size_t g = gj >> 1, j = gj & 1;
        const DevPrime &Q = primes[i];
        uint64_t x = addmod(mulmod(pk[(j * K + i) * n + c], U[(g * L1 + i) * n + c], Q.q, Q.ratio_lo, Q.ratio_hi),
                            out[idx], Q.q);
        if (drop) x = shoup(x, dc[i].inv, dc[i].inv_shoup, Q.q);
        if (plain && j == 0) x = addmod(x, plain[(g * L + i) * n + c], Q.q);
        out[idx] = x;
    }
}

// BFV: out[g][j][i] = (P[g][j][i] + e_j - r_j) p^-1 (mod q_i), P = INTT(pk_j NTT(u)) with
// L1 limbs per polynomial; without a dropped prime, P[g][j][i] + e_j.
__global__ void enc_finish_bfv_kernel(const uint64_t *__restrict__ P, const int64_t *__restrict__ Es,
                                      uint64_t *__restrict__ out, const DevPrime *__restrict__ primes,
                                      const DivRoundConst *__restrict__ dc, uint64_t p, bool drop, uint32_t L,
                                      uint32_t L1, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (n - 1);
        const uint32_t r = uniform_row(idx, logn), gj = r / L, i = r - gj * L;
        const DevPrime &Q = primes[i];
        const int64_t e = Es[(size_t)gj * n + c];
        const uint64_t *Pg = P + (size_t)gj * L1 * n;
        uint64_t x = addmod(Pg[(size_t)i * n + c], lift_small(e, Q.q), Q.q);
        if (drop) {
            const uint64_t last = addmod(Pg[(size_t)L * n + c], lift_small(e, p), p);
            x = shoup(submod(x, divround_corr(last, p, Q, dc[i]), Q.q), dc[i].inv, dc[i].inv_shoup, Q.q);
        }
        out[idx] = x;
    }
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
    // every digit's symmetric encryption of zero (encrypt_zero_symmetric with seed
    // sub_seed(seed, 100 + j)) in one launch per stage
    const DevPrime *primes = c.primes();
    SamplerKey skey{};
    for (int i = 0; i < 8; i++) skey.k[i] = c.sampler_key[i];
    const size_t words = (size_t)K << c.logn;
    const dim3 ug(blocks_for(c.sampler_keyed ? (words + 3) / 4 : words), K - 1);
    const dim3 eg(blocks_for(c.sampler_keyed ? (c.n + 7) / 8 : c.n), K - 1);
    if (c.sampler_keyed) {
        keygen_uniform_kernel<true><<<ug, TPB, 0, st>>>(skey, seed, primes, K, c.logn, key);
        keygen_cbd_kernel<true><<<eg, TPB, 0, st>>>(skey, seed, primes, K, c.logn, key);
    } else {
        keygen_uniform_kernel<false><<<ug, TPB, 0, st>>>(skey, seed, primes, K, c.logn, key);
        keygen_cbd_kernel<false><<<eg, TPB, 0, st>>>(skey, seed, primes, K, c.logn, key);
    }
    launch_check("keygen sample");
    ntt_forward_grid(key, LimbGrid{1, 2 * K, K, K - 1}, primes, c.logn, NTT_FULL, st);
    const size_t cw = (size_t)(K - 1) * words;
    keygen_combine_kernel<<<blocks_for(cw), TPB, 0, st>>>(key, sk, primes, K, c.logn, cw);
    launch_check("keygen combine");
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
    bfv_scale_add_kernel<<<blocks_for(c.n), TPB, 0, st>>>(c0, 0, plain, 1, B, c.primes(), c.logn, subtract);
    launch_check("bfv_scale_add");
}

void bfv_add_plain_scaled(const Context &c, uint64_t *c0, const uint64_t *plain, uint32_t L, hipStream_t st) {
    bfv_scale_add(c, c0, plain, L, false, st);
}

namespace {

// Words of encrypt_many's workspace per ciphertext: U [L1][N], e0 and e1 [2][N], and
// the products P ([2][N] of the dropped limb for CKKS, [2][L1][N] for BFV).
size_t encrypt_many_words(const Context &c, uint32_t L) {
    const size_t L1 = upper_limbs(c, L);
    return (L1 + 2 + (c.scheme == SCHEME_BFV ? 2 * L1 : 2)) * c.n;
}

void encrypt_chunk(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t m, uint32_t L, uint64_t seed,
                   uint64_t *out, void *ws, hipStream_t st) {
    const uint32_t L1 = upper_limbs(c, L), K = (uint32_t)c.key_count();
    const bool drop = L1 != L, ckks = c.scheme == SCHEME_CKKS;
    const size_t n = c.n;
    const int logn = c.logn;
    const DevPrime *primes = c.primes();
    uint64_t *U = static_cast<uint64_t *>(ws);
    int64_t *Es = reinterpret_cast<int64_t *>(U + (size_t)m * L1 * n);
    uint64_t *P = reinterpret_cast<uint64_t *>(Es + (size_t)m * 2 * n);
    // drop: divide-and-round by q_L (row L1 of the rescale constants, entries i < L)
    const DivRoundConst *dc = drop ? c.rescale_consts(L1) : nullptr;
    const uint64_t p = drop ? c.moduli[L] : 0;

    SamplerKey key{};
    for (int i = 0; i < 8; i++) key.k[i] = c.sampler_key[i];
    if (c.sampler_keyed) {
        const size_t total = (size_t)m * ((n + 7) / 8);
        sample_encrypt_kernel<true><<<blocks_for(total), TPB, 0, st>>>(key, seed, primes, L1, logn, total, U, Es);
    } else {
        const size_t total = (size_t)m * n;
        sample_encrypt_kernel<false><<<blocks_for(total), TPB, 0, st>>>(key, seed, primes, L1, logn, total, U, Es);
    }
    launch_check("sample_encrypt");
    ntt_forward(U, m * L1, L1, primes, logn, NTT_FULL, st);
    const size_t out_words = (size_t)m * 2 * L * n;
    if (ckks) {
        if (drop) {
            const size_t pw = (size_t)m * 2 * n;
            enc_products_kernel<<<blocks_for(pw), TPB, 0, st>>>(pk, K, U, L1, L, 1, P, primes, logn, pw);
            launch_check("enc_products");
            ntt_inverse(P, m * 2, 1, primes + L, logn, NTT_FULL, st);
        }
        const EncArgs ea{Es, drop ? P : nullptr, pk, U, plain, out, dc, p, K, L1, L, drop};
        if (ntt_encrypt_ckks(ea, m, primes, logn, st)) return;
        enc_error_kernel<<<blocks_for(out_words), TPB, 0, st>>>(Es, drop ? P : nullptr, out, primes, dc, p, L, logn,
                                                                out_words);
        launch_check("enc_error");
        ntt_forward(out, m * 2 * L, L, primes, logn, NTT_FULL, st);
        enc_finish_ckks_kernel<<<blocks_for(out_words), TPB, 0, st>>>(pk, K, U, L1, out, plain, primes, dc, drop, L,
                                                                      logn, out_words);
        launch_check("enc_finish");
        return;
    }
    const size_t pw = (size_t)m * 2 * L1 * n;
    enc_products_kernel<<<blocks_for(pw), TPB, 0, st>>>(pk, K, U, L1, 0, L1, P, primes, logn, pw);
    launch_check("enc_products");
    ntt_inverse(P, m * 2 * L1, L1, primes, logn, NTT_FULL, st);
    enc_finish_bfv_kernel<<<blocks_for(out_words), TPB, 0, st>>>(P, Es, out, primes, dc, p, drop, L, L1, logn,
                                                                 out_words);
    launch_check("enc_finish");
    if (plain) {
        bfv_scale_add_kernel<<<blocks_for((size_t)m * n), TPB, 0, st>>>(out, (size_t)2 * L * n, plain, m,
                                                                        c.behz(L).dev, primes, logn, false);
        launch_check("bfv_scale_add");
    }
}

}  // namespace

size_t encrypt_many_workspace(const Context &c, uint32_t L, uint32_t count) {
    return (size_t)(count ? count : 1) * encrypt_many_words(c, L) * 8;
}

void encrypt_many(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t count, uint32_t L,
                  uint64_t seed, uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st) {
    // Encryptor::encryptInternal (src/encryptor.cpp:152-233), one seed per ciphertext
    if (L < 1 || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "level out of range for encryption");
    if (!count) return;
    const size_t per = encrypt_many_words(c, L) * 8;
    if (!ws || ws_bytes < per) fail(ERR_INVALID_ARGUMENT, "workspace too small for encryption");
    const uint32_t chunk = (uint32_t)std::min<size_t>(count, ws_bytes / per);
    const size_t out_words = (size_t)2 * L * c.n, plain_words = c.scheme == SCHEME_BFV ? c.n : (size_t)L * c.n;
    for (uint32_t g0 = 0; g0 < count; g0 += chunk) {
        const uint32_t m = std::min(chunk, count - g0);
        encrypt_chunk(c, pk, plain ? plain + g0 * plain_words : nullptr, m, L, seed + g0, out + g0 * out_words, ws, st);
    }
}

void encrypt(const Context &c, const uint64_t *pk, const uint64_t *plain, uint32_t L, uint64_t seed, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st) {
    // Encryptor::encryptInternal (src/encryptor.cpp:152-233)
    if (L < 1 || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "level out of range for encryption");
    if (!ws || ws_bytes < encrypt_workspace(c, L)) fail(ERR_INVALID_ARGUMENT, "workspace too small for encryption");
    encrypt_many(c, pk, plain, 1, L, seed, out, ws, ws_bytes, st);
}

size_t decrypt_workspace(const Context &c, uint32_t L) { return (size_t)4 * L * c.n * 8; }

size_t decrypt_many_workspace(const Context &c, uint32_t L, uint32_t count) {
    return c.scheme == SCHEME_CKKS ? 0 : (size_t)(count ? count : 1) * 3 * L * c.n * 8;
}

void decrypt_many(const Context &c, const uint64_t *ct, uint32_t size, uint32_t count, const uint64_t *sk, uint32_t L,
                  uint64_t *out, void *ws, size_t ws_bytes, hipStream_t st) {
    // Decryptor::decrypt (src/decryptor.cpp:79-153), one call per ciphertext there
    if (size < 2 || size > 3) fail(ERR_INVALID_ARGUMENT, "ciphertext size must be 2 or 3");
    if (!count) return;
    const size_t n = c.n, poly = (size_t)L * n, ct_words = size * poly;
    const DevPrime *primes = c.primes();
    if (c.scheme == SCHEME_CKKS) {
        const size_t total = (size_t)count * poly;
        launch_phase(ct, ct + poly, poly, ct_words, size, sk, out, primes, c.logn, L, total, st);
        return;
    }
    // BFV: phase = INTT(sum_k NTT(c_k) s^k) + c0, then scale-and-round, per chunk of m
    const size_t per = ct_words * 8;  // (size - 1) transformed polynomials + the phase
    if (!ws || ws_bytes < per) fail(ERR_INVALID_ARGUMENT, "workspace too small for decryption");
    const uint32_t chunk = (uint32_t)std::min<size_t>(count, ws_bytes / per);
    const BehzDev &B = c.behz(L).dev;
    for (uint32_t g0 = 0; g0 < count; g0 += chunk) {
        const uint32_t m = std::min(chunk, count - g0);
        const uint64_t *src = ct + g0 * ct_words;
        uint64_t *tmp = static_cast<uint64_t *>(ws), *phase = tmp + (size_t)m * (size - 1) * poly;
        const size_t tw = (size - 1) * poly, gw = (size_t)m * tw;
        gather_rows_kernel<<<blocks_for(gw), TPB, 0, st>>>(src + poly, ct_words, tmp, (size - 1) * L, c.logn, gw);
        launch_check("gather_rows");
        ntt_forward(tmp, m * (size - 1) * L, L, primes, c.logn, NTT_FULL, st);
        const size_t total = (size_t)m * poly;
        launch_phase(nullptr, tmp, poly, (size - 1) * poly, size, sk, phase, primes, c.logn, L, total, st);
        ntt_inverse(phase, m * L, L, primes, c.logn, NTT_FULL, st);
        scale_round_kernel<<<blocks_for((size_t)m * n), TPB, 0, st>>>(phase, src, ct_words, out + g0 * n, m, B, primes,
                                                                      c.logn);
        launch_check("scale_round");
    }
}

void decrypt(const Context &c, const uint64_t *ct, uint32_t size, const uint64_t *sk, uint32_t L, uint64_t *out,
             void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme == SCHEME_BFV && (!ws || ws_bytes < decrypt_workspace(c, L)))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for decryption");
    decrypt_many(c, ct, size, 1, sk, L, out, ws, ws_bytes, st);
}

void batch_encode_many(const Context &c, const uint64_t *values, uint32_t count, size_t values_stride,
                       uint32_t plaintexts, uint64_t *plain, hipStream_t st) {
    if (!c.plain_prime()) fail(ERR_LOGIC, "encryption parameters are not valid for batching");
    if (count > c.n) fail(ERR_INVALID_ARGUMENT, "values_matrix size is too large");
    if (plaintexts > 1 && values_stride < count) fail(ERR_INVALID_ARGUMENT, "values stride is below the value count");
    if (!plaintexts) return;
    const size_t total = (size_t)plaintexts * c.n;
    batch_scatter_kernel<<<blocks_for(total), TPB, 0, st>>>(values, count, values_stride, plain, c.logn, total);
    launch_check("batch_scatter");
    ntt_inverse(plain, plaintexts, 1, c.plain_prime(), c.logn, NTT_FULL, st);
}

void batch_decode_many(const Context &c, const uint64_t *plain, uint32_t plaintexts, uint64_t *values,
                       uint64_t *scratch, hipStream_t st) {
    if (!c.plain_prime()) fail(ERR_LOGIC, "encryption parameters are not valid for batching");
    if (!plaintexts) return;
    const size_t total = (size_t)plaintexts * c.n;
    SHC_HIP(hipMemcpyAsync(scratch, plain, total * 8, hipMemcpyDeviceToDevice, st));
    ntt_forward(scratch, plaintexts, 1, c.plain_prime(), c.logn, NTT_FULL, st);
    batch_gather_kernel<<<blocks_for(total), TPB, 0, st>>>(scratch, values, c.logn, total);
    launch_check("batch_gather");
}

void batch_encode(const Context &c, const uint64_t *values, uint32_t count, uint64_t *plain, hipStream_t st) {
    batch_encode_many(c, values, count, count, 1, plain, st);
}

void batch_decode(const Context &c, const uint64_t *plain, uint64_t *values, uint64_t *scratch, hipStream_t st) {
    batch_decode_many(c, plain, 1, values, scratch, st);
}

}  // namespace shc
