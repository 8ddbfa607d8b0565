// rotations.hip -- hoisted rotation and summation, exact against separate rotations.
//
// A rotation (Evaluator::applyGaloisInplace, src/evaluator.cpp:2210-2250; GPU
// src/evaluator_cuda.cu:2011-2104) switches the key of sigma(c1), whose digits are the
// canonical residues of sigma applied to c1's coefficients.  Hoisting transforms c1's
// digits once: T_ij = NTT_i(d_j mod q_i), and sigma(T_ij) (an NTT-domain permutation)
// is the transform of sigma(d_j) mod q_i -- except where sigma negates a coefficient.
// There the reference's digit is q_j - d (for d != 0), which is q_j mod q_i above the
// hoisted -d mod q_i.  So, per target prime i and with S_sigma the 0/1 polynomial of
// negated positions,
//   sum_j NTT_i(digit_j) key_j
//     = sigma(sum_j T_ij sigma^-1(key_j)) + NTT_i(S_sigma) sum_j (q_j mod q_i) key_j
//     = sigma(sum_j T_ij key'_j + corr')
// with key' = sigma^-1(key) and corr' = sigma^-1(NTT_i(S_sigma) sum_j (q_j mod q_i) key_j),
// prepared once per key and level.  A zero digit d_j[p] = 0 at a position sigma negates
// has the reference digit 0, not q_j: corr' counted (q_j mod q_i) X^sigma(p) there, whose
// preimage under sigma is -X^p.  Each such zero is put right in the products before the
// permutation, P += (q_j mod q_i) NTT_i(X^p) key'_j (the NTT of a monomial is a power of
// the root at each slot), and a ciphertext whose digits are all zero has products zero.
// The zeros are listed on the device, so no decision waits for the host.  Each rotation
// then costs one inner product, one permutation of the 2(L+1) product limbs and one
// ModDown instead of a full ModUp (L x L transforms for CKKS, L(L+1) for BFV).
#include "rotations.h"
#include "elementwise.h"
#include "galois.h"
#include "hostmath.h"
#include "accum.h"
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

inline size_t align256(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// neg[l][k] = 1 where sigma_elt negates the coefficient that lands at k (the source
// i = k elt^-1 mod N moves to i elt mod 2N, negated when that is >= N), in each limb.
__global__ void negated_positions_kernel(uint64_t *__restrict__ neg, uint32_t elt, uint32_t elt_inv, int logn,
                                         size_t total) {
    const uint32_t n = 1u << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t k = (uint32_t)(idx & (n - 1));
        const uint32_t i = (uint32_t)(((uint64_t)k * elt_inv) & (n - 1));
        neg[idx] = (((uint64_t)i * elt) >> logn) & 1;
    }
}

// C[k][i][c] = negntt[i][c] * sum_{j<L} (q_j mod q_i) key[j][k][ki][c] mod q_i, i <= L
// (ki = i, or the special prime's key limb K-1 for i == L).
__global__ void correction_kernel(const uint64_t *__restrict__ negntt, const uint64_t *__restrict__ key,
                                  uint64_t *__restrict__ C, const DevPrime *__restrict__ targets,
                                  const DevPrime *__restrict__ data, uint32_t L, uint32_t K, int logn, size_t total) {
    const size_t n = (size_t)1 << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (n - 1);
        const size_t r = idx >> logn;
        const uint32_t i = (uint32_t)(r % (L + 1));
        const uint32_t k = (uint32_t)(r / (L + 1));
        const DevPrime &Q = targets[i];
        const uint32_t ki = i == L ? K - 1 : i;
        Acc3 acc{0, 0, 0};
        for (uint32_t j = 0; j < L; j++) {
            const uint64_t cij = barrett64(data[j].q, Q.q, Q.ratio_hi);
            mac(acc, cij, key[(((size_t)j * 2 + k) * K + ki) * n + c]);
        }
        const uint64_t s = acc_reduce(acc, Q);
        C[idx] = mulmod(s, negntt[(size_t)i * n + c], Q.q, Q.ratio_lo, Q.ratio_hi);
    }
}

// Products of R rotations at once, each T word read once for all of them:
//   P_r[ct][k][i] = sum_j T[i][ct][j] key'_r[j][k][ki] + corr'_r[k][i]   (S_r for i == L)
// One thread per (i, c) position and RG ciphertexts (each key word it loads serves RG
// products), with three-word sums (Acc3); the loads of digit j + 1 go out before the
// products of digit j, and in a partial last group the missing ciphertext repeats the
// last one (in bounds, not stored).  With diag, T holds no limb (i, i) for i < L and
// that term is the NTT-form input limb diag[ct * diag_stride + i * N].  Blocks of one
// tile run on one XCD (XCD = block % 8) so that the tile's key words are served from
// its L2.
constexpr int RMAX = 4;
constexpr int RG = 2;
struct RotBatch {
    const uint64_t *key[RMAX];   // prepared keys; corr' follows at corr_off words
    uint64_t *P[RMAX], *S[RMAX];
    uint32_t elt[RMAX];          // their Galois elements
};

template <int R>
__global__ __launch_bounds__(TPB) void rotation_products_kernel(const uint64_t *__restrict__ T, RotBatch rb,
                                                                size_t corr_off, const DevPrime *__restrict__ targets,
                                                                uint32_t L, uint32_t K, int logn, uint32_t count,
                                                                uint32_t tiles, const uint64_t *__restrict__ diag,
                                                                size_t diag_stride) {
    const uint32_t groups = (count + RG - 1) / RG;
    const uint32_t xcd = blockIdx.x & 7, q8 = blockIdx.x >> 3;
    const uint32_t tile = (q8 / groups) * 8 + xcd;
    if (tile >= tiles) return;
    const uint32_t ct0 = (q8 % groups) * RG;
    const int gv = (int)min<uint32_t>(RG, count - ct0);
    const size_t n = (size_t)1 << logn;
    const size_t flat = (size_t)tile * TPB + threadIdx.x;
    if (flat >= (size_t)(L + 1) << logn) return;
    const uint32_t i = (uint32_t)(flat >> logn);
    const size_t c = flat & (n - 1);
    const uint32_t ki = (i == L) ? K - 1 : i;
    const DevPrime &Q = targets[i];
    const bool short_row = diag && i < L;  // uniform over the block
    const uint32_t Lt = short_row ? L - 1 : L;
    const uint64_t *Ti = T + (((size_t)i * count * L + (size_t)ct0 * Lt) << logn) + c;
    const size_t kstride = (size_t)K << logn;
    size_t goff[RG];
    const uint64_t *dp[RG];
#pragma unroll
    for (int g = 0; g < RG; g++) {
        goff[g] = ((size_t)min(g, gv - 1) * Lt) << logn;
        dp[g] = short_row ? diag + (ct0 + min(g, gv - 1)) * diag_stride + ((size_t)i << logn) + c : nullptr;
    }
    Acc3 a0[R][RG], a1[R][RG];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int g = 0; g < RG; g++) a0[r][g] = a1[r][g] = Acc3{0, 0, 0};
    uint64_t tn[RG], k0n[R], k1n[R];
    auto fetch = [&](uint32_t j) {
        const bool dj = short_row && j == i;
        const size_t jt = short_row && j > i ? j - 1 : j;
#pragma unroll
        for (int g = 0; g < RG; g++) tn[g] = *(dj ? dp[g] : Ti + goff[g] + (jt << logn));
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint64_t *kp = rb.key[r] + (size_t)ki * n + c;
            k0n[r] = kp[(2 * j) * kstride];
            k1n[r] = kp[(2 * j + 1) * kstride];
        }
    };
    fetch(0);
    for (uint32_t j = 0; j < L; j++) {
        uint64_t t[RG], k0[R], k1[R];
#pragma unroll
        for (int g = 0; g < RG; g++) t[g] = tn[g];
#pragma unroll
        for (int r = 0; r < R; r++) k0[r] = k0n[r], k1[r] = k1n[r];
        fetch(j + 1 < L ? j + 1 : j);
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int g = 0; g < RG; g++) {
                mac(a0[r][g], t[g], k0[r]);
                mac(a1[r][g], t[g], k1[r]);
            }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint64_t *corr = rb.key[r] + corr_off;
        const uint64_t c0 = corr[(size_t)i * n + c], c1 = corr[((size_t)(L + 1) + i) * n + c];
#pragma unroll
        for (int g = 0; g < RG; g++) {
            if (g >= gv) continue;
            const size_t ct = ct0 + g;
            acc_add(a0[r][g], c0);
            acc_add(a1[r][g], c1);
            const uint64_t r0 = acc_reduce(a0[r][g], Q), r1 = acc_reduce(a1[r][g], Q);
            if (i == L) {
                rb.S[r][(ct * 2 + 0) * n + c] = r0;
                rb.S[r][(ct * 2 + 1) * n + c] = r1;
            } else {
                rb.P[r][((ct * 2 + 0) * L + i) * n + c] = r0;
                rb.P[r][((ct * 2 + 1) * L + i) * n + c] = r1;
            }
        }
    }
}

// out c0 = (first ? 0 : out c0) + s0, out c1 = first ? 0 : out c1; s0: [count][L][N].
__global__ void accumulate_c0_kernel(uint64_t *__restrict__ out, const uint64_t *__restrict__ s0,
                                     const DevPrime *__restrict__ primes, uint32_t L, int logn, bool first,
                                     size_t total) {
    const size_t lw = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t ct = idx / lw, r = idx % lw;
        uint64_t *c0 = out + ct * 2 * lw + r;
        const uint64_t v = s0[idx];
        if (first) {
            *c0 = v;
            c0[lw] = 0;
        } else {
            *c0 = addmod(*c0, v, primes[r >> logn].q);
        }
    }
}

__global__ void clear_flag_kernel(unsigned long long *flag) { *flag = 0; }

// A prepared key ends with two header words: HOISTED_TAG | L, and the Galois element.
constexpr uint64_t HOISTED_TAG = 0x484f4953ull << 32;  // "HOIS"
__global__ void write_hoisted_header_kernel(uint64_t *hdr, uint32_t L, uint32_t elt) {
    hdr[0] = HOISTED_TAG | L;
    hdr[1] = elt;
}

// flag |= 2 if a prepared key's header does not name level L and element elts[r].
constexpr int HDR_CHUNK = 16;
struct HeaderBatch {
    const uint64_t *hdr[HDR_CHUNK];
    uint32_t elt[HDR_CHUNK];
};
__global__ void check_hoisted_headers_kernel(HeaderBatch hb, uint32_t nr, uint32_t L, unsigned long long *flag) {
    const uint32_t r = threadIdx.x;
    if (r >= nr) return;
    if (hb.hdr[r][0] != (HOISTED_TAG | L) || hb.hdr[r][1] != hb.elt[r]) atomicOr(flag, 2ull);
}

// Zero digits of the batch: zc[0] counts them and entry e < zc[0] of `list` is
// (ct << 32) | (j << 24) | p; zc[2 + ct] counts ciphertext ct's (the list has room for
// every digit word, so it cannot overflow).
__global__ void zero_list_kernel(const uint64_t *__restrict__ digits, size_t dstride, uint32_t L, int logn,
                                 unsigned long long *__restrict__ zc, uint64_t *__restrict__ list, size_t total) {
    const size_t per_ct = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t ct = idx / per_ct, r = idx % per_ct;
        if (digits[ct * dstride + r] != 0) continue;
        const unsigned long long e = atomicAdd(zc, 1ull);
        atomicAdd(zc + 2 + ct, 1ull);
        list[e] = ((uint64_t)ct << 32) | ((uint64_t)(r >> logn) << 24) | (r & ((size_t(1) << logn) - 1));
    }
}

__global__ void clear_words_kernel(unsigned long long *w, size_t words) {
    for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < words; i += (size_t)gridDim.x * TPB) w[i] = 0;
}

// The zero-digit fix of R rotations' products (before their permutations): at target
// i, slot s (one thread each),
//   P_r[ct][k][i][s] += sum over zeros (ct, j, p) that sigma_r negates, j's prime not i,
//                       of (q_j mod q_i) psi_i^((2 brev(s) + 1) p) key'_r[j][k][ki][s],
// psi_i^((2 brev(s)+1) p) being slot s of NTT_i(X^p); products of a ciphertext whose
// L N digit words are all zero are set to zero (its exact products).  Exits at once
// when the batch has no zero digit.
template <int R>
__global__ __launch_bounds__(TPB) void zero_digit_fix_kernel(RotBatch rb, const unsigned long long *__restrict__ zc,
                                                             const uint64_t *__restrict__ list,
                                                             const DevPrime *__restrict__ targets,
                                                             const DevPrime *__restrict__ data, uint32_t L, uint32_t K,
                                                             int logn, uint32_t count) {
    const unsigned long long z = zc[0];
    if (z == 0) return;
    const size_t n = (size_t)1 << logn;
    const size_t flat = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (flat >= (size_t)(L + 1) << logn) return;
    const uint32_t i = (uint32_t)(flat >> logn);
    const uint32_t s = (uint32_t)(flat & (n - 1));
    const uint32_t ki = (i == L) ? K - 1 : i;
    const DevPrime &Q = targets[i];
    const uint64_t all = (uint64_t)L << logn;
    auto out = [&](int r, size_t ct, int k) -> uint64_t * {
        return i == L ? rb.S[r] + (ct * 2 + k) * n + s : rb.P[r] + ((ct * 2 + k) * L + i) * n + s;
    };
    for (uint32_t ct = 0; ct < count; ct++) {
        if (zc[2 + ct] != all) continue;
#pragma unroll
        for (int r = 0; r < R; r++) *out(r, ct, 0) = *out(r, ct, 1) = 0;
    }
    const uint64_t two_bs1 = 2 * (uint64_t)(__brev(s) >> (32 - logn)) + 1;
    const size_t kstride = (size_t)K << logn;
    for (unsigned long long e = 0; e < z; e++) {
        const uint64_t w = list[e];
        const size_t ct = w >> 32;
        if (zc[2 + ct] == all) continue;
        const uint32_t j = (uint32_t)(w >> 24) & 0xff, p = (uint32_t)w & 0xffffff;
        if (j == i) continue;  // q_j mod q_j = 0
        const uint64_t qji = barrett64(data[j].q, Q.q, Q.ratio_hi);
        const uint64_t ex = (two_bs1 * p) & (2 * n - 1);
        const uint64_t m = ex & (n - 1);
        const uint64_t root = Q.fwd[__brev((uint32_t)m) >> (32 - logn)].x;
        const uint64_t psi = ex >= n ? Q.q - root : root;
        const uint64_t f = mulmod(qji, psi, Q.q, Q.ratio_lo, Q.ratio_hi);
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (((((uint64_t)p * rb.elt[r]) >> logn) & 1) == 0) continue;  // sigma_r keeps p's sign
            const uint64_t *kp = rb.key[r] + (size_t)ki * n + s;
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const uint64_t t = mulmod(f, kp[(2 * (size_t)j + k) * kstride], Q.q, Q.ratio_lo, Q.ratio_hi);
                uint64_t *o = out(r, ct, k);
                *o = addmod(*o, t, Q.q);
            }
        }
    }
}

struct RotLayout {
    size_t gal, tmpct, P2, S2, C0, Pm, Sm, flag, zc, total;
};

RotLayout rot_layout(const Context &c, uint32_t count, uint32_t L) {
    const size_t limb = c.n * 8;
    RotLayout l;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    l.gal = take(galois_keyswitch_workspace(c, count, L));
    l.tmpct = take((size_t)count * 2 * L * limb);
    l.P2 = take((size_t)count * 2 * L * limb);
    l.S2 = take((size_t)count * 2 * limb);
    l.C0 = take((size_t)count * L * limb);
    l.Pm = take((size_t)RMAX * count * 2 * L * limb);
    l.Sm = take((size_t)RMAX * count * 2 * limb);
    l.flag = take(8);
    l.zc = take(8 * (2 + (size_t)count));
    l.total = off;
    return l;
}

void check_level(const Context &c, uint32_t L) {
    if (c.key_count() < 2) fail(ERR_LOGIC, "keyswitching is not supported by the context");
    if (L < 1 || L > c.key_count() - 1) fail(ERR_INVALID_ARGUMENT, "level out of range for key switching");
}

}  // namespace

size_t hoisted_key_words(const Context &c, uint32_t L) {
    const size_t K = c.key_count();
    return ((K - 1) * 2 * K + 2 * ((size_t)L + 1)) * c.n + 2;
}

size_t hoisted_key_workspace(const Context &c, uint32_t L) { return 3 * ((size_t)L + 1) * c.n * 8; }

void hoisted_key_prepare(const Context &c, const uint64_t *galois_key, uint32_t elt, uint32_t L, uint64_t *out,
                         void *ws, size_t ws_bytes, hipStream_t st) {
    check_level(c, L);
    const uint64_t m = 2 * c.n;
    if (!(elt & 1) || elt >= m) fail(ERR_INVALID_ARGUMENT, "Galois element is not valid");
    if (!ws || ws_bytes < hoisted_key_workspace(c, L)) fail(ERR_INVALID_ARGUMENT, "workspace too small");
    const uint32_t K = (uint32_t)c.key_count();
    const size_t n = c.n;
    const uint32_t elt_inv = (uint32_t)hm::invmod(elt, m);
    // key' = sigma^-1(key), limb by limb
    apply_galois(c, galois_key, K * n, out, K * n, (K - 1) * 2, K, elt_inv, true, st);
    // corr' = sigma^-1(NTT_i(S_sigma) * sum_j (q_j mod q_i) key_j[i])
    uint64_t *neg = static_cast<uint64_t *>(ws);
    uint64_t *C = neg + ((size_t)L + 1) * n;
    const size_t nw = ((size_t)L + 1) * n;
    negated_positions_kernel<<<blocks_for(nw), TPB, 0, st>>>(neg, elt, elt_inv, c.logn, nw);
    launch_check("negated_positions");
    const DevPrime *targets = c.keyswitch_primes(L);
    ntt_forward_grid(neg, LimbGrid{1, L + 1, L + 1, 1}, targets, c.logn, NTT_FULL, st);
    correction_kernel<<<blocks_for(2 * nw), TPB, 0, st>>>(neg, galois_key, C, targets, c.primes(), L, K, c.logn,
                                                          2 * nw);
    launch_check("correction");
    apply_galois(c, C, nw, out + (size_t)(K - 1) * 2 * K * n, nw, 2, L + 1, elt_inv, true, st);
    write_hoisted_header_kernel<<<1, 1, 0, st>>>(out + hoisted_key_words(c, L) - 2, L, elt);
    launch_check("hoisted key header");
    std::lock_guard<std::mutex> g(c.hoisted_mu);
    c.hoisted_keys[out] = {L, elt};
}

size_t rotate_sum_workspace(const Context &c, uint32_t count, uint32_t L) { return rot_layout(c, count, L).total; }

namespace {

// The hoisted rotations of `in`, summed into out (sum) or each into its own
// out + r * count * 2L N (!sum).
void rotate_hoisted(const Context &c, const uint64_t *in, uint32_t count, uint32_t L, const uint32_t *elts,
                    const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                    void *ws, size_t ws_bytes, hipStream_t st, bool sum) {
    check_level(c, L);
    if (!count) return;
    if (!rotations || !elts || !hoisted || !keys) fail(ERR_INVALID_ARGUMENT, "no rotations given");
    KernelProbe probe{nullptr, nullptr};
    const bool armed = c.take_probe(probe);  // brackets the first pass of hoisted products
    const RotLayout lay = rot_layout(c, count, L);
    if (!ws || ws_bytes < lay.total) fail(ERR_INVALID_ARGUMENT, "workspace too small for rotation sums");
    if (in == out) fail(ERR_INVALID_ARGUMENT, "rotate_sum cannot run in place");
    const uint64_t max_cts = ((uint64_t(1) << 34) >> c.logn) / ((uint64_t)L * (L + 1));
    if ((uint64_t)count > max_cts) fail(ERR_INVALID_ARGUMENT, "rotation batch too large for one call");
    const bool ckks = c.scheme == SCHEME_CKKS;
    const int logn = c.logn;
    const size_t n = c.n, lw = (size_t)L * n;
    char *base = static_cast<char *>(ws);
    uint64_t *tmpct = reinterpret_cast<uint64_t *>(base + lay.tmpct);
    uint64_t *P2 = reinterpret_cast<uint64_t *>(base + lay.P2);
    uint64_t *S2 = reinterpret_cast<uint64_t *>(base + lay.S2);
    uint64_t *C0 = reinterpret_cast<uint64_t *>(base + lay.C0);
    auto *flag = reinterpret_cast<unsigned long long *>(base + lay.flag);
    // the hoisted path uses the key-switch regions that follow the Galois scratch row
    const size_t sigma_bytes = align256((size_t)count * lw * 8);
    const KsRegions ks = keyswitch_regions(c, count, L, base + lay.gal + sigma_bytes);

    // each prepared key must be the one of this level and element: checked on the
    // host for keys prepared through this context; other keys' header words are read
    // back from the device (the only case that waits for the stream)
    const size_t hdr_off = hoisted_key_words(c, L) - 2;
    std::vector<uint32_t> unknown;
    {
        std::lock_guard<std::mutex> g(c.hoisted_mu);
        for (uint32_t r = 0; r < rotations; r++) {
            if (!hoisted[r] || !keys[r]) fail(ERR_INVALID_ARGUMENT, "null key pointer");
            auto it = c.hoisted_keys.find(hoisted[r]);
            if (it == c.hoisted_keys.end()) unknown.push_back(r);
            else if (it->second.first != L || it->second.second != elts[r])
                fail(ERR_INVALID_ARGUMENT, "a hoisted key was prepared for another level or Galois element");
        }
    }
    if (!unknown.empty()) {
        clear_flag_kernel<<<1, 1, 0, st>>>(flag);
        for (size_t r0 = 0; r0 < unknown.size(); r0 += HDR_CHUNK) {
            HeaderBatch hb{};
            const uint32_t nr = (uint32_t)std::min<size_t>(unknown.size() - r0, HDR_CHUNK);
            for (uint32_t q = 0; q < nr; q++) {
                hb.hdr[q] = hoisted[unknown[r0 + q]] + hdr_off;
                hb.elt[q] = elts[unknown[r0 + q]];
            }
            check_hoisted_headers_kernel<<<1, 64, 0, st>>>(hb, nr, L, flag);
            launch_check("hoisted key headers");
        }
        unsigned long long flags = 0;
        SHC_HIP(hipMemcpyAsync(&flags, flag, 8, hipMemcpyDeviceToHost, st));
        SHC_HIP(hipStreamSynchronize(st));
        if (flags & 2)
            fail(ERR_INVALID_ARGUMENT, "a hoisted key was prepared for another level or Galois element");
    }
    // 1. digits of c1 in coefficient form (CKKS: transformed into ks.d; c1 itself, in NTT
    // form, supplies the own-prime terms of the products)
    const uint64_t *digits = in + lw;
    size_t dstride = 2 * lw;
    if (ckks) {
        ntt_inverse_from(in + lw, LimbGrid{1, 2 * L, L, count}, ks.d, LimbGrid{1, L, L, count}, c.primes(), logn,
                         NTT_FULL, st);
        digits = ks.d;
        dstride = lw;
    }
    // 2. zero digits, listed for the products' fix (the list has room for every word)
    auto *zc = reinterpret_cast<unsigned long long *>(base + lay.zc);
    clear_words_kernel<<<1, TPB, 0, st>>>(zc, 2 + (size_t)count);
    zero_list_kernel<<<blocks_for(count * lw), TPB, 0, st>>>(digits, dstride, L, logn, zc, tmpct, count * lw);
    launch_check("zero digits");
    const size_t ct_words = (size_t)count * 2 * lw;
    // 3. one ModUp for every rotation
    const bool skipped = modup_to_T(c, digits, dstride, ks.T, count, L, st, nullptr, ckks);
    const uint64_t *diag = skipped ? in + lw : nullptr;
    const DevPrime *targets = c.keyswitch_primes(L);
    const size_t K = c.key_count();
    const size_t corr_off = (K - 1) * 2 * K * n;
    uint64_t *Pm = reinterpret_cast<uint64_t *>(base + lay.Pm);
    uint64_t *Sm = reinterpret_cast<uint64_t *>(base + lay.Sm);
    const uint32_t tiles = (uint32_t)((((size_t)(L + 1) << logn) + TPB - 1) / TPB);
    const uint64_t blocks = (uint64_t)((tiles + 7) / 8) * 8 * ((count + RG - 1) / RG);
    if (blocks * TPB >= (uint64_t(1) << 32)) fail(ERR_INVALID_ARGUMENT, "rotation batch too large for one call");
    for (uint32_t r0 = 0; r0 < rotations; r0 += RMAX) {
        const uint32_t nr = rotations - r0 < (uint32_t)RMAX ? rotations - r0 : (uint32_t)RMAX;
        // products with the prepared keys (corrections included), RMAX rotations per pass over T
        RotBatch rb{};
        for (uint32_t q = 0; q < nr; q++) {
            rb.key[q] = hoisted[r0 + q];
            rb.P[q] = Pm + (size_t)q * count * 2 * lw;
            rb.S[q] = Sm + (size_t)q * count * 2 * n;
            rb.elt[q] = elts[r0 + q];
        }
        if (armed && r0 == 0) SHC_HIP(hipEventRecord(probe.begin, st));
        switch (nr) {
            case 1: rotation_products_kernel<1><<<(uint32_t)blocks, TPB, 0, st>>>(ks.T, rb, corr_off, targets, L, (uint32_t)K, logn, count, tiles, diag, 2 * lw); break;
            case 2: rotation_products_kernel<2><<<(uint32_t)blocks, TPB, 0, st>>>(ks.T, rb, corr_off, targets, L, (uint32_t)K, logn, count, tiles, diag, 2 * lw); break;
            case 3: rotation_products_kernel<3><<<(uint32_t)blocks, TPB, 0, st>>>(ks.T, rb, corr_off, targets, L, (uint32_t)K, logn, count, tiles, diag, 2 * lw); break;
            default: rotation_products_kernel<4><<<(uint32_t)blocks, TPB, 0, st>>>(ks.T, rb, corr_off, targets, L, (uint32_t)K, logn, count, tiles, diag, 2 * lw); break;
        }
        launch_check("rotation_products");
        if (armed && r0 == 0) SHC_HIP(hipEventRecord(probe.end, st));
        const uint32_t fix_blocks = (uint32_t)((((size_t)(L + 1) << logn) + TPB - 1) / TPB);
        switch (nr) {
            case 1: zero_digit_fix_kernel<1><<<fix_blocks, TPB, 0, st>>>(rb, zc, tmpct, targets, c.primes(), L, (uint32_t)K, logn, count); break;
            case 2: zero_digit_fix_kernel<2><<<fix_blocks, TPB, 0, st>>>(rb, zc, tmpct, targets, c.primes(), L, (uint32_t)K, logn, count); break;
            case 3: zero_digit_fix_kernel<3><<<fix_blocks, TPB, 0, st>>>(rb, zc, tmpct, targets, c.primes(), L, (uint32_t)K, logn, count); break;
            default: zero_digit_fix_kernel<4><<<fix_blocks, TPB, 0, st>>>(rb, zc, tmpct, targets, c.primes(), L, (uint32_t)K, logn, count); break;
        }
        launch_check("zero_digit_fix");
        for (uint32_t q = 0; q < nr; q++) {
            const uint32_t r = r0 + q;
            uint64_t *dst = sum ? out : out + r * ct_words;
            // (sigma(c0), 0) into dst (accumulated when summing)
            apply_galois(c, in, 2 * lw, C0, lw, count, L, elts[r], ckks, st);
            accumulate_c0_kernel<<<blocks_for(count * lw), TPB, 0, st>>>(dst, C0, c.primes(), L, logn,
                                                                         !sum || r == 0, count * lw);
            launch_check("accumulate_c0");
            // the products permuted by sigma, then this rotation's ModDown added into dst
            apply_galois(c, rb.P[q], lw, P2, lw, count * 2, L, elts[r], true, st);
            apply_galois(c, rb.S[q], n, S2, n, count * 2, 1, elts[r], true, st);
            moddown_into(c, P2, S2, ks.U, dst, 2 * lw, count, L, st);
        }
    }
}

}  // namespace

void rotate_sum(const Context &c, const uint64_t *in, uint32_t count, uint32_t L, const uint32_t *elts,
                const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                void *ws, size_t ws_bytes, hipStream_t st) {
    rotate_hoisted(c, in, count, L, elts, hoisted, keys, rotations, out, ws, ws_bytes, st, true);
}

void rotate_many(const Context &c, const uint64_t *in, uint32_t count, uint32_t L, const uint32_t *elts,
                 const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                 void *ws, size_t ws_bytes, hipStream_t st) {
    rotate_hoisted(c, in, count, L, elts, hoisted, keys, rotations, out, ws, ws_bytes, st, false);
}

}  // namespace shc
