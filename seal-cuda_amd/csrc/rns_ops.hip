// rns_ops.hip -- tensor product, key switching and divide-and-round kernels.
//
// Key switching (src/evaluator.cpp:2310-2653; GPU: evaluator_cuda.cu:1012-1363) is
// restructured as a handful of wide launches over a whole batch instead of the
// reference's ~11k per-limb launches:
//
//   1. CKKS: inverse NTT of c2 (digits d_j in coefficient form)
//   2. modup_gather: T[j][i] = d_j mod q_i for every digit j < L and target prime
//      i in {q_0..q_{L-1}, special}
//   3. one batched forward NTT over all L(L+1) limbs of T
//   4. inner_product: P_k[i] = sum_j T[j][i] * key[j][k][i] with 128-bit accumulation,
//      one Barrett reduction at the end; G ciphertexts per thread share each key load
//   5. ModDown: inverse NTT of the special limbs, rounding correction, forward NTT
//      (CKKS) or inverse NTT of P (BFV), then ct_k[i] += (P_k[i] - t) * p^-1.
#include "modarith.h"
#include "ntt.h"
#include "rns_ops.h"

namespace shc {
namespace {

constexpr int TPB = 256;

// Elementwise kernels run grid-stride loops; the grid is capped so that a launch stays
// far below the 2^32 work-item limit however large the batch is.
inline uint32_t blocks_for(size_t n) {
    const size_t b = (n + TPB - 1) / TPB;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

__global__ void tensor_kernel(const uint64_t *__restrict__ a, size_t a_stride, const uint64_t *__restrict__ b,
                              size_t b_stride, uint64_t *__restrict__ out, size_t out_stride,
                              const DevPrime *__restrict__ primes, uint32_t L, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t limbwords = (size_t)L << logn;
    const size_t ct = idx / limbwords, rem = idx % limbwords;
    const DevPrime &P = primes[rem >> logn];
    const uint64_t *pa = a + ct * a_stride + rem, *pb = b + ct * b_stride + rem;
    const uint64_t a0 = pa[0], a1 = pa[limbwords], b0 = pb[0], b1 = pb[limbwords];
    uint64_t *po = out + ct * out_stride + rem;
    po[0] = mulmod(a0, b0, P.q, P.ratio_lo, P.ratio_hi);
    u128 cross = mul64x64(a0, b1);
    add128(cross, mul64x64(a1, b0));
    po[limbwords] = barrett128(cross, P.q, P.ratio_lo, P.ratio_hi);
    po[2 * limbwords] = mulmod(a1, b1, P.q, P.ratio_lo, P.ratio_hi);
    }
}

// T[ct][j][i][c] = d[ct][j][c] mod q_i, i in [0, L] (target i == L is the special prime).
__global__ void modup_gather_kernel(const uint64_t *__restrict__ d, size_t d_stride, uint64_t *__restrict__ T,
                                    const DevPrime *__restrict__ targets, const DevPrime *__restrict__ digits,
                                    uint32_t L, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    size_t r = idx >> logn;
    const uint32_t i = (uint32_t)(r % (L + 1));
    r /= (L + 1);
    const uint32_t j = (uint32_t)(r % L);
    const size_t ct = r / L;
    const uint64_t v = d[ct * d_stride + ((size_t)j << logn) + c];
    const DevPrime &Q = targets[i];
    T[idx] = digits[j].q <= Q.q ? v : barrett64(v, Q.q, Q.ratio_hi);
    }
}

// P[ct][k][i][c] = sum_j T[ct][j][i][c] * key[j][k][ki][c] mod q_i for i < L, and the
// special target (i == L) goes to S[ct][k][c].  Each thread owns one (i, c) for G
// consecutive ciphertexts, so every key word it loads is used G times.
template <int G>
__global__ void inner_product_kernel(const uint64_t *__restrict__ T, const uint64_t *__restrict__ key,
                                     uint64_t *__restrict__ P, uint64_t *__restrict__ S,
                                     const DevPrime *__restrict__ targets, uint32_t L, uint32_t K, int logn,
                                     uint32_t count) {
    const size_t n = (size_t)1 << logn;
    const size_t per_group = (size_t)(L + 1) << logn;
    const size_t groups = (count + G - 1) / G;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < groups * per_group;
         idx += (size_t)gridDim.x * TPB) {
    const size_t grp = idx / per_group;
    const size_t rem = idx % per_group;
    const uint32_t i = (uint32_t)(rem >> logn);
    const size_t c = rem & (n - 1);
    const uint32_t ki = (i == L) ? K - 1 : i;
    const DevPrime &Q = targets[i];
    const size_t ct0 = grp * G;
    const int g_valid = (int)min<size_t>(G, count - ct0);
    const size_t t_ct_stride = (size_t)L * (L + 1) << logn;
    u128 acc0[G], acc1[G];
#pragma unroll
    for (int g = 0; g < G; g++) acc0[g] = acc1[g] = u128{0, 0};
    for (uint32_t j = 0; j < L; j++) {
        const uint64_t k0 = key[(((size_t)j * 2 + 0) * K + ki) * n + c];
        const uint64_t k1 = key[(((size_t)j * 2 + 1) * K + ki) * n + c];
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (g < g_valid) {
                const uint64_t t = T[(ct0 + g) * t_ct_stride + (((size_t)j * (L + 1) + i) << logn) + c];
                add128(acc0[g], mul64x64(t, k0));
                add128(acc1[g], mul64x64(t, k1));
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g < g_valid) {
            const size_t ct = ct0 + g;
            const uint64_t r0 = barrett128(acc0[g], Q.q, Q.ratio_lo, Q.ratio_hi);
            const uint64_t r1 = barrett128(acc1[g], Q.q, Q.ratio_lo, Q.ratio_hi);
            if (i == L) {
                S[(ct * 2 + 0) * n + c] = r0;
                S[(ct * 2 + 1) * n + c] = r1;
            } else {
                P[((ct * 2 + 0) * L + i) * n + c] = r0;
                P[((ct * 2 + 1) * L + i) * n + c] = r1;
            }
        }
    }
    }
}

// U[g][i][c] = ([S[g][c] + floor(p/2)]_p mod q_i - floor(p/2) mod q_i) mod q_i
// (the rounding correction of SEAL's ModDown / divideAndRoundqLast).
__global__ void divround_prep_kernel(const uint64_t *__restrict__ S, uint64_t *__restrict__ U,
                                     const DevPrime *__restrict__ primes, const DivRoundConst *__restrict__ dc,
                                     uint64_t p, uint32_t Lout, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const size_t r = idx >> logn;
    const uint32_t i = (uint32_t)(r % Lout);
    const size_t g = r / Lout;
    const DevPrime &Q = primes[i];
    uint64_t v = S[g * n + c] + (p >> 1);
    v = v >= p ? v - p : v;
    v = barrett64(v, Q.q, Q.ratio_hi);
    U[idx] = submod(v, dc[i].half_mod, Q.q);
    }
}

// ct[ct][k][i] += (P[ct][k][i] - U[ct][k][i]) * p^-1 mod q_i
__global__ void moddown_finish_kernel(const uint64_t *__restrict__ P, const uint64_t *__restrict__ U,
                                      uint64_t *__restrict__ ct, size_t ct_stride, const DevPrime *__restrict__ primes,
                                      const DivRoundConst *__restrict__ dc, uint32_t L, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t per_ct = (size_t)2 * L << logn;
    const size_t cti = idx / per_ct, rem = idx % per_ct;
    const uint32_t i = (uint32_t)((rem >> logn) % L);
    const DevPrime &Q = primes[i];
    const uint64_t x = submod(P[idx], U[idx], Q.q);
    const uint64_t v = shoup(x, dc[i].inv, dc[i].inv_shoup, Q.q);
    uint64_t *dst = ct + cti * ct_stride + rem;
    *dst = addmod(*dst, v, Q.q);
    }
}

// out[g][i] = (in[g][i] - U[g][i]) * q_last^-1 mod q_i, in has Lin = Lout + 1 limbs per group.
__global__ void rescale_finish_kernel(const uint64_t *__restrict__ in, const uint64_t *__restrict__ U,
                                      uint64_t *__restrict__ out, const DevPrime *__restrict__ primes,
                                      const DivRoundConst *__restrict__ dc, uint32_t Lout, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const size_t r = idx >> logn;
    const uint32_t i = (uint32_t)(r % Lout);
    const size_t g = r / Lout;
    const DevPrime &Q = primes[i];
    const uint64_t x = submod(in[(g * (Lout + 1) + i) * n + c], U[idx], Q.q);
    out[idx] = shoup(x, dc[i].inv, dc[i].inv_shoup, Q.q);
    }
}

// S[g][c] = in[g][Lin-1][c]: gathers the last limb of every polynomial.
__global__ void last_limb_kernel(const uint64_t *__restrict__ in, uint64_t *__restrict__ S, uint32_t Lin, int logn,
                                 size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t g = idx >> logn, c = idx & (n - 1);
    S[idx] = in[(g * Lin + Lin - 1) * n + c];
    }
}

// dst[g][limbs][N] = src[g * src_stride ...] (strided gather of c2 polynomials)
__global__ void gather_polys_kernel(const uint64_t *__restrict__ src, size_t src_stride, uint64_t *__restrict__ dst,
                                    size_t poly_words, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t g = idx / poly_words, r = idx % poly_words;
    dst[idx] = src[g * src_stride + r];
    }
}

struct KsLayout {
    size_t d, T, P, S, U, total;
};

KsLayout ks_layout(const Context &c, uint32_t count, uint32_t L) {
    const size_t limb = c.n * 8;
    KsLayout l;
    size_t off = 0;
    auto take = [&](size_t limbs) { size_t o = off; off += (limbs * limb + 255) & ~size_t(255); return o; };
    l.d = take((size_t)count * L);
    l.T = take((size_t)count * L * (L + 1));
    l.P = take((size_t)count * 2 * L);
    l.S = take((size_t)count * 2);
    l.U = take((size_t)count * 2 * L);
    l.total = off;
    return l;
}

}  // namespace

void ckks_tensor(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count, uint32_t L,
                 hipStream_t st) {
    const size_t lw = (size_t)L * c.n;
    tensor_strided(c.primes(), a, 2 * lw, b, 2 * lw, out, 3 * lw, count, L, c.logn, st);
}

void tensor_strided(const DevPrime *primes, const uint64_t *a, size_t a_stride, const uint64_t *b, size_t b_stride,
                    uint64_t *out, size_t out_stride, uint32_t count, uint32_t L, int logn, hipStream_t st) {
    const size_t total = ((size_t)count * L) << logn;
    if (!total) return;
    tensor_kernel<<<blocks_for(total), TPB, 0, st>>>(a, a_stride, b, b_stride, out, out_stride, primes, L, logn,
                                                     total);
    launch_check("tensor");
}

size_t keyswitch_workspace(const Context &c, uint32_t count, uint32_t L) { return ks_layout(c, count, L).total; }

void switch_key(const Context &c, const uint64_t *target, size_t target_stride, uint64_t *ct, size_t ct_stride,
                const uint64_t *key, uint32_t count, uint32_t L, void *ws, size_t ws_bytes, hipStream_t st) {
    const size_t K = c.key_count();
    if (K < 2) fail(ERR_LOGIC, "keyswitching is not supported by the context");
    if (L < 1 || L > K - 1) fail(ERR_INVALID_ARGUMENT, "level out of range for key switching");
    if (!count) return;
    const KsLayout lay = ks_layout(c, count, L);
    if (!ws || ws_bytes < lay.total) fail(ERR_INVALID_ARGUMENT, "workspace too small for key switching");
    char *base = static_cast<char *>(ws);
    uint64_t *d = reinterpret_cast<uint64_t *>(base + lay.d);
    uint64_t *T = reinterpret_cast<uint64_t *>(base + lay.T);
    uint64_t *P = reinterpret_cast<uint64_t *>(base + lay.P);
    uint64_t *S = reinterpret_cast<uint64_t *>(base + lay.S);
    uint64_t *U = reinterpret_cast<uint64_t *>(base + lay.U);
    const int logn = c.logn;
    const size_t n = c.n;
    const bool ckks = c.scheme == SCHEME_CKKS;
    const DevPrime *targets = c.keyswitch_primes(L);

    // 1. digits in coefficient form
    const size_t dwords = (size_t)count * L * n;
    gather_polys_kernel<<<blocks_for(dwords), TPB, 0, st>>>(target, target_stride, d, (size_t)L * n, dwords);
    launch_check("gather");
    if (ckks) ntt_inverse(d, count * L, L, c.primes(), logn, NTT_FULL, st);
    // 2-3. ModUp
    if (!ntt_forward_modup(d, (size_t)L * n, T, count, L, targets, logn, st)) {
        const size_t twords = (size_t)count * L * (L + 1) * n;
        modup_gather_kernel<<<blocks_for(twords), TPB, 0, st>>>(d, (size_t)L * n, T, targets, c.primes(), L, logn,
                                                               twords);
        launch_check("modup_gather");
        ntt_forward(T, count * L * (L + 1), L + 1, targets, logn, NTT_FULL, st);
    }
    // 4. inner product with the key
    constexpr int G = 4;
    const size_t ip_threads = ((size_t)(count + G - 1) / G) * (L + 1) * n;
    inner_product_kernel<G><<<blocks_for(ip_threads), TPB, 0, st>>>(T, key, P, S, targets, L, (uint32_t)K, logn, count);
    launch_check("inner_product");
    // 5. ModDown by the special prime
    ntt_inverse(S, count * 2, 1, targets + L, logn, NTT_FULL, st);
    const size_t uwords = (size_t)count * 2 * L * n;
    divround_prep_kernel<<<blocks_for(uwords), TPB, 0, st>>>(S, U, c.primes(), c.ks_consts(), c.moduli[K - 1], L, logn,
                                                            uwords);
    launch_check("moddown_prep");
    if (ckks) ntt_forward(U, count * 2 * L, L, c.primes(), logn, NTT_FULL, st);
    else ntt_inverse(P, count * 2 * L, L, c.primes(), logn, NTT_FULL, st);
    moddown_finish_kernel<<<blocks_for(uwords), TPB, 0, st>>>(P, U, ct, ct_stride, c.primes(), c.ks_consts(), L, logn,
                                                             uwords);
    launch_check("moddown_finish");
}

size_t rescale_workspace(const Context &c, uint32_t count, uint32_t polys, uint32_t L) {
    const size_t g = (size_t)count * polys;
    return ((g * c.n * 8 + 255) & ~size_t(255)) + g * (L - 1) * c.n * 8;
}

void rescale_to_next(const Context &c, const uint64_t *in, uint64_t *out, uint32_t count, uint32_t polys, uint32_t L,
                     void *ws, size_t ws_bytes, hipStream_t st) {
    if (L < 2 || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "end of modulus switching chain reached");
    if (!count || !polys) return;
    if (!ws || ws_bytes < rescale_workspace(c, count, polys, L))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for rescale");
    const size_t g = (size_t)count * polys;
    const int logn = c.logn;
    uint64_t *S = static_cast<uint64_t *>(ws);
    uint64_t *U = reinterpret_cast<uint64_t *>(static_cast<char *>(ws) + ((g * c.n * 8 + 255) & ~size_t(255)));
    const bool ckks = c.scheme == SCHEME_CKKS;
    const DevPrime *primes = c.primes();
    const size_t swords = g * c.n;
    last_limb_kernel<<<blocks_for(swords), TPB, 0, st>>>(in, S, L, logn, swords);
    launch_check("last_limb");
    if (ckks) ntt_inverse(S, (uint32_t)g, 1, primes + (L - 1), logn, NTT_FULL, st);
    const size_t uwords = g * (L - 1) * c.n;
    divround_prep_kernel<<<blocks_for(uwords), TPB, 0, st>>>(S, U, primes, c.rescale_consts(L), c.moduli[L - 1], L - 1,
                                                            logn, uwords);
    launch_check("rescale_prep");
    if (ckks) ntt_forward(U, (uint32_t)(g * (L - 1)), L - 1, primes, logn, NTT_FULL, st);
    rescale_finish_kernel<<<blocks_for(uwords), TPB, 0, st>>>(in, U, out, primes, c.rescale_consts(L), L - 1, logn,
                                                             uwords);
    launch_check("rescale_finish");
}

}  // namespace shc
