// rns_ops.hip -- tensor product, key switching and divide-and-round kernels.
//
// Key switching (src/evaluator.cpp:2310-2653; GPU: evaluator_cuda.cu:1012-1363) is
// restructured as a handful of wide launches over a whole batch instead of the
// reference's ~11k per-limb launches:
//
//   1. CKKS: inverse NTT of c2 into the workspace (digits d_j in coefficient form;
//      c2 stays in NTT form)
//   2. modup: T[i][ct][j] = d_j mod q_i for every digit j < L and target prime
//      i in {q_0..q_{L-1}, special}, target-major so that consecutive limbs share a
//      prime (its twiddles stay in L2)
//   3. one batched forward NTT over the limbs of T (fused with step 2 for N >= 2^13):
//      L(L+1) of them per ciphertext, L x L for CKKS, whose own-prime limbs (j == i)
//      are c2's NTT-form limbs themselves (the reference's i == j shortcut)
//   4. inner_product: P_k[i] = sum_j T[j][i] * key[j][k][i] with 128-bit accumulation,
//      one Barrett reduction at the end; G ciphertexts per thread share each key load.
//      At N = 2^15 and 2^16 the NTT's rows pass and this product are one kernel
//      (ntt_modup_keyprod): the transformed limbs never return to HBM
//   5. ModDown: inverse NTT of the special limbs, rounding correction, forward NTT
//      (CKKS) or inverse NTT of P (BFV), then ct_k[i] += (P_k[i] - t) * p^-1.
#include "accum.h"
#include "modarith.h"
#include "ntt.h"
#include "rns_ops.h"
#include <vector>

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

// T[i][ct][j][c] = d[ct][j][c] mod q_i, i in [0, L] (target i == L is the special
// prime); left unreduced when the digit's prime is at most 8 q_i (the forward NTT
// takes inputs < 8q).
__global__ void modup_gather_kernel(const uint64_t *__restrict__ d, size_t d_stride, uint64_t *__restrict__ T,
                                    const DevPrime *__restrict__ targets, const DevPrime *__restrict__ digits,
                                    uint32_t count, uint32_t L, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const size_t r = idx >> logn;
    const size_t per_target = (size_t)count * L;
    const uint32_t i = (uint32_t)(r / per_target);
    const uint32_t j = (uint32_t)(r % L);
    const size_t ct = (r % per_target) / L;
    const uint64_t v = d[ct * d_stride + ((size_t)j << logn) + c];
    const DevPrime &Q = targets[i];
    T[idx] = digits[j].q <= 8 * Q.q ? v : barrett64(v, Q.q, Q.ratio_hi);
    }
}

// P[ct][k][i][c] = sum_j T[i][ct][j][c] * key[j][k][ki][c] mod q_i for i < L, and the
// special target (i == L) goes to S[ct][k][c].  A block owns a tile of 256 flat (i, c)
// positions for one group of G ciphertexts; a thread owns one position and keeps
// three-word sums (Acc3) for its G ciphertexts, so each key word it loads serves G
// products.  The loads of digit j + 2 are issued before the products of digit j (past
// the end, the last digit again).  In a partial last group the missing ciphertexts
// repeat the last one (loads stay in bounds; their sums are not stored).  With diag,
// T holds no limb (i, i) for i < L (ModUp's skip_diag layout) and that term is the
// NTT-form input limb diag[ct * diag_stride + i * N] itself (CKKS: NTT(INTT(x)) = x).
// Blocks are dealt to the 8 XCDs round-robin (XCD = block % 8); the mapping below puts
// every group of one tile on the same XCD, one after another, so the tile's key words
// come from HBM once and then from that XCD's L2.
template <int G>
__global__ __launch_bounds__(TPB) void inner_product_kernel(const uint64_t *__restrict__ T,
                                                            const uint64_t *__restrict__ key,
                                                            uint64_t *__restrict__ P, uint64_t *__restrict__ S,
                                                            const DevPrime *__restrict__ targets, uint32_t L,
                                                            uint32_t K, int logn, uint32_t count, uint32_t tiles,
                                                            uint32_t groups, const uint64_t *__restrict__ diag,
                                                            size_t diag_stride) {
    const uint32_t xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const uint32_t tile = (k / groups) * 8 + xcd;
    if (tile >= tiles) return;
    const uint32_t grp = k % groups;
    const size_t n = (size_t)1 << logn;
    const size_t flat = (size_t)tile * TPB + threadIdx.x;
    if (flat >= (size_t)(L + 1) << logn) return;
    const uint32_t i = (uint32_t)(flat >> logn);
    const size_t c = flat & (n - 1);
    const uint32_t ki = (i == L) ? K - 1 : i;
    const DevPrime &Q = targets[i];
    const size_t ct0 = (size_t)grp * G;
    const int g_valid = (int)min<size_t>(G, count - ct0);
    const bool short_row = diag && i < L;  // uniform over the block
    const uint32_t Lt = short_row ? L - 1 : L;
    const uint64_t *Ti = T + (((size_t)i * count * L + ct0 * Lt) << logn) + c;
    const uint64_t *kp = key + (size_t)ki * n + c;
    const size_t kstride = (size_t)K << logn;  // key[j][k] -> key[j][k+1]
    size_t goff[G];
    const uint64_t *dp[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        goff[g] = ((size_t)min(g, g_valid - 1) * Lt) << logn;
        dp[g] = short_row ? diag + (ct0 + min(g, g_valid - 1)) * diag_stride + ((size_t)i << logn) + c : nullptr;
    }
    Acc3 acc0[G], acc1[G];
#pragma unroll
    for (int g = 0; g < G; g++) acc0[g] = acc1[g] = Acc3{0, 0, 0};
    // two digits in flight: buffer b holds digit j while the loads of digit j + 2 go out
    uint64_t tb[2][G], kb[2][2];
    auto fetch = [&](int b, uint32_t j) {
        j = j < L ? j : L - 1;  // past the end: the last digit again (in bounds, unused)
        kb[b][0] = kp[(2 * j) * kstride];
        kb[b][1] = kp[(2 * j + 1) * kstride];
        const bool dj = short_row && j == i;
        const size_t jt = short_row && j > i ? j - 1 : j;
#pragma unroll
        for (int g = 0; g < G; g++) tb[b][g] = *(dj ? dp[g] : Ti + goff[g] + (jt << logn));
    };
    auto step = [&](int b, uint32_t jnext) {
        uint64_t t[G];
#pragma unroll
        for (int g = 0; g < G; g++) t[g] = tb[b][g];
        const uint64_t k0 = kb[b][0], k1 = kb[b][1];
        fetch(b, jnext);
#pragma unroll
        for (int g = 0; g < G; g++) {
            mac(acc0[g], t[g], k0);
            mac(acc1[g], t[g], k1);
        }
    };
    fetch(0, 0);
    fetch(1, 1);
    uint32_t j = 0;
    for (; j + 2 <= L; j += 2) {
        step(0, j + 2);
        step(1, j + 3);
    }
    if (j < L) step(0, L);
#pragma unroll
    for (int g = 0; g < G; g++) {
        if (g < g_valid) {
            const size_t ct = ct0 + g;
            const uint64_t r0 = acc_reduce(acc0[g], Q);
            const uint64_t r1 = acc_reduce(acc1[g], Q);
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

// U[g][i][c] = ([S[g][c] + floor(p/2)]_p mod q_i - floor(p/2) mod q_i) mod q_i
// (the rounding correction of SEAL's ModDown / divideAndRoundqLast).
__global__ void divround_prep_kernel(const uint64_t *__restrict__ S, uint64_t *__restrict__ U,
                                     const DevPrime *__restrict__ primes, const DivRoundConst *__restrict__ dc,
                                     uint64_t p, uint32_t Lout, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const uint32_t r = uniform_row(idx, logn), g = r / Lout, i = r - g * Lout;
    const DevPrime &Q = primes[i];
    uint64_t v = S[(size_t)g * n + c] + (p >> 1);
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
    const uint32_t r = uniform_row(idx, logn), g = r / Lout, i = r - g * Lout;
    const DevPrime &Q = primes[i];
    const uint64_t x = submod(in[((size_t)g * (Lout + 1) + i) * n + c], U[idx], Q.q);
    out[idx] = shoup(x, dc[i].inv, dc[i].inv_shoup, Q.q);
    }
}

// Coefficient-form divide-and-round by q_Lo: out[g][i] = (in[g][i] - u_i) q_Lo^-1 with
// u_i = ([in[g][Lo] + floor(q_Lo/2)]_{q_Lo} mod q_i) - floor(q_Lo/2) mod q_i, in with
// Lo + 1 limbs per polynomial, out with Lo.
__global__ void rescale_coeff_kernel(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                     const DevPrime *__restrict__ primes, const DivRoundConst *__restrict__ dc,
                                     uint64_t p, uint32_t Lo, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const uint32_t r = uniform_row(idx, logn), g = r / Lo, i = r - g * Lo;
    const DevPrime &Q = primes[i];
    const uint64_t *src = in + (size_t)g * (Lo + 1) * n + c;
    uint64_t v = src[(size_t)Lo * n] + (p >> 1);
    v = v >= p ? v - p : v;
    const uint64_t u = submod(barrett64(v, Q.q, Q.ratio_hi), dc[i].half_mod, Q.q);
    out[idx] = shoup(submod(src[(size_t)i * n], u, Q.q), dc[i].inv, dc[i].inv_shoup, Q.q);
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

// ---- grouped elementwise passes (one launch over several level groups) -------------
// Group g owns blocks [start[g], start[g+1]), one element per thread.
struct ElemGroup {
    const uint64_t *a, *b;
    uint64_t *out, *aux;
    size_t sa, sb, so;       // strides (words) between ciphertexts
    const DivRoundConst *dc;
    uint64_t p;
    uint32_t L, count;
    size_t total;
};
struct ElemGroups {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    ElemGroup g[MAX_GROUPS];
};
enum ElemOp { OP_TENSOR, OP_LAST_LIMB, OP_DIVROUND_PREP, OP_RESCALE_FINISH };

__device__ __forceinline__ uint32_t elem_group(const ElemGroups &eg, uint32_t b) {
    uint32_t g = 0;
#pragma unroll 1
    while (g + 1 < eg.n && b >= eg.start[g + 1]) g++;
    return g;
}

//   OP_TENSOR:         out[ct] = (a0 b0, a0 b1 + a1 b0, a1 b1) at level L (a, b: 2 polys, out: 3)
//   OP_LAST_LIMB:      out[ct][k] = a[ct][k][L-1]              (k < 2; a: stride sa)
//   OP_DIVROUND_PREP:  out[ct][k][i] = ([b[ct][k] + p/2]_p mod q_i) - (p/2 mod q_i), i < L
//   OP_RESCALE_FINISH: out[ct][k][i] = (a[ct][k][i] - aux[ct][k][i]) p^-1, i < L (a: L+1 limbs)
template <int OP>
__global__ void elem_grouped_kernel(ElemGroups eg, const DevPrime *__restrict__ primes, int logn) {
    const uint32_t gi = elem_group(eg, blockIdx.x);
    const ElemGroup &e = eg.g[gi];
    const size_t idx = (size_t)(blockIdx.x - eg.start[gi]) * TPB + threadIdx.x;
    if (idx >= e.total) return;
    const size_t n = (size_t)1 << logn;
    const size_t c = idx & (n - 1);
    const size_t r = idx >> logn;
    if constexpr (OP == OP_TENSOR) {
        const size_t ct = r / e.L;
        const uint32_t i = (uint32_t)(r % e.L);
        const size_t lw = (size_t)e.L << logn;
        const DevPrime &P = primes[i];
        const uint64_t *pa = e.a + ct * e.sa + ((size_t)i << logn) + c, *pb = e.b + ct * e.sb + ((size_t)i << logn) + c;
        const uint64_t a0 = pa[0], a1 = pa[lw], b0 = pb[0], b1 = pb[lw];
        uint64_t *po = e.out + ct * e.so + ((size_t)i << logn) + c;
        po[0] = mulmod(a0, b0, P.q, P.ratio_lo, P.ratio_hi);
        u128 cross = mul64x64(a0, b1);
        add128(cross, mul64x64(a1, b0));
        po[lw] = barrett128(cross, P.q, P.ratio_lo, P.ratio_hi);
        po[2 * lw] = mulmod(a1, b1, P.q, P.ratio_lo, P.ratio_hi);
    } else if constexpr (OP == OP_LAST_LIMB) {
        const size_t ct = r >> 1, k = r & 1;
        e.out[idx] = e.a[ct * e.sa + ((k * e.L + e.L - 1) << logn) + c];
    } else if constexpr (OP == OP_DIVROUND_PREP) {
        const uint32_t i = (uint32_t)(r % e.L);
        const size_t g = r / e.L;
        const DevPrime &Q = primes[i];
        uint64_t v = e.b[g * n + c] + (e.p >> 1);
        v = v >= e.p ? v - e.p : v;
        v = barrett64(v, Q.q, Q.ratio_hi);
        e.out[idx] = submod(v, e.dc[i].half_mod, Q.q);
    } else {
        const uint32_t i = (uint32_t)(r % e.L);
        const size_t g = r / e.L, ct = g >> 1, k = g & 1;
        const DevPrime &Q = primes[i];
        const uint64_t x = submod(e.a[ct * e.sa + ((k * (e.L + 1) + i) << logn) + c], e.aux[idx], Q.q);
        e.out[ct * e.so + ((k * e.L + i) << logn) + c] = shoup(x, e.dc[i].inv, e.dc[i].inv_shoup, Q.q);
    }
}

template <int OP>
void launch_elem_groups(const std::vector<ElemGroup> &gs, const DevPrime *primes, int logn, hipStream_t st) {
    ElemGroups eg{};
    eg.n = (uint32_t)gs.size();
    uint64_t acc = 0;
    for (size_t g = 0; g < gs.size(); g++) {
        eg.g[g] = gs[g];
        eg.start[g] = (uint32_t)acc;
        acc += (gs[g].total + TPB - 1) / TPB;
    }
    eg.start[gs.size()] = (uint32_t)acc;
    if (acc == 0) return;
    if (acc >= (uint64_t(1) << 32) / TPB) fail(ERR_INVALID_ARGUMENT, "grouped launch too large");
    elem_grouped_kernel<OP><<<(uint32_t)acc, TPB, 0, st>>>(eg, primes, logn);
    launch_check("grouped elementwise");
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

KsRegions keyswitch_regions(const Context &c, uint32_t count, uint32_t L, void *ws) {
    const KsLayout l = ks_layout(c, count, L);
    char *b = static_cast<char *>(ws);
    auto at = [b](size_t off) { return reinterpret_cast<uint64_t *>(b + off); };
    return KsRegions{at(l.d), at(l.T), at(l.P), at(l.S), at(l.U)};
}

namespace {

// One launch sequence over `count` ciphertexts (few enough that every launch stays
// within the NTT's per-launch limb bound).
void switch_key_chunk(const Context &c, const uint64_t *target, size_t target_stride, uint64_t *ct,
                      size_t ct_stride, const uint64_t *key, uint32_t count, uint32_t L, void *ws, hipStream_t st,
                      const KernelProbe *probe, const uint64_t *fold_c0) {
    const KsLayout lay = ks_layout(c, count, L);
    char *base = static_cast<char *>(ws);
    uint64_t *d = reinterpret_cast<uint64_t *>(base + lay.d);
    uint64_t *T = reinterpret_cast<uint64_t *>(base + lay.T);
    uint64_t *P = reinterpret_cast<uint64_t *>(base + lay.P);
    uint64_t *S = reinterpret_cast<uint64_t *>(base + lay.S);
    uint64_t *U = reinterpret_cast<uint64_t *>(base + lay.U);
    const int logn = c.logn;
    const size_t n = c.n;
    const bool ckks = c.scheme == SCHEME_CKKS;

    // 1. digits in coefficient form: BFV targets already are; a CKKS target is
    // transformed into d and stays in NTT form, where it supplies the key products'
    // own-prime terms (the ModUp of digit i to prime i is the input limb itself)
    const uint64_t *digits = target;
    size_t dstride = target_stride;
    const uint64_t *diag = nullptr;
    if (ckks) {
        if (target_stride % n == 0 && target_stride / n <= UINT32_MAX / count) {
            ntt_inverse_from(target, LimbGrid{1, (uint32_t)(target_stride / n), L, count}, d,
                             LimbGrid{1, L, L, count}, c.primes(), logn, NTT_FULL, st);
            diag = target;
        } else {
            const size_t dwords = (size_t)count * L * n;
            gather_polys_kernel<<<blocks_for(dwords), TPB, 0, st>>>(target, target_stride, d, (size_t)L * n, dwords);
            launch_check("gather");
            ntt_inverse(d, count * L, L, c.primes(), logn, NTT_FULL, st);
        }
        digits = d;
        dstride = (size_t)L * n;
    }
    // 2-4. ModUp and the inner product with the key: fused past the columns pass where
    // the transform shape allows (N = 2^15, 2^16), else T goes through HBM once more
    // (CKKS: the fused products also take p times the ciphertext the switch adds into, so
    // that the ModDown writes it without reading it)
    const KsShape shape = c.ks_shape(L, count);
    const KsFold fold = ckks ? KsFold{fold_c0 ? fold_c0 : ct, ct_stride, c.ks_consts(), fold_c0 != nullptr} : KsFold{};
    const bool fused = ntt_modup_keyprod(digits, dstride, T, key, P, S, diag, target_stride, count, L,
                                         (uint32_t)c.key_count(), c.keyswitch_primes(L), c.primes(), logn, st, probe,
                                         shape, fold);
    if (fold_c0 && !fused) fail(ERR_LOGIC, "a key switch into (c0, 0) needs the fused key products");
    if (!fused) {
        const bool skipped = modup_to_T(c, digits, dstride, T, count, L, st, probe, diag != nullptr);
        keyswitch_inner_product(c, T, key, P, S, count, L, st, skipped ? diag : nullptr, target_stride);
    }
    // 5. ModDown by the special prime
    moddown_into(c, P, S, U, ct, ct_stride, count, L, st, fused && fold.ct);
}

}  // namespace

bool modup_to_T(const Context &c, const uint64_t *digits, size_t dstride, uint64_t *T, uint32_t count, uint32_t L,
                hipStream_t st, const KernelProbe *probe, bool skip_diag) {
    const DevPrime *targets = c.keyswitch_primes(L);
    const int logn = c.logn;
    // T's consumers (inner products with the key, hoisted rotation products) sum L
    // products T * key (key < q < 2^60) and one more term < q in 128 bits: T may stay
    // in [0, 8q) up to L = 31 (31 * 8q^2 < 2^127.95), in [0, 4q) beyond.
    if (ntt_forward_modup(digits, dstride, T, count, L, targets, c.primes(), logn, skip_diag, L > 31, st, probe,
                          c.ks_shape(L, count)))
        return skip_diag;
    const size_t twords = ((size_t)count * L * (L + 1)) << logn;
    modup_gather_kernel<<<blocks_for(twords), TPB, 0, st>>>(digits, dstride, T, targets, c.primes(), count, L, logn,
                                                           twords);
    launch_check("modup_gather");
    if (probe) SHC_HIP(hipEventRecord(probe->begin, st));
    ntt_forward_grid(T, LimbGrid{count * L, 1, L + 1, count * L}, targets, logn, NTT_FULL, st);
    if (probe) SHC_HIP(hipEventRecord(probe->end, st));
    return false;
}

void keyswitch_inner_product(const Context &c, const uint64_t *T, const uint64_t *key, uint64_t *P, uint64_t *S,
                             uint32_t count, uint32_t L, hipStream_t st, const uint64_t *diag, size_t diag_stride) {
    constexpr int G = 4;
    const int logn = c.logn;
    const uint32_t groups = (count + G - 1) / G;
    const uint32_t tiles = (uint32_t)((((size_t)(L + 1) << logn) + TPB - 1) / TPB);
    const uint64_t ip_blocks = (uint64_t)((tiles + 7) / 8) * 8 * groups;
    if (ip_blocks * TPB >= (uint64_t(1) << 32)) fail(ERR_INVALID_ARGUMENT, "key-switch batch too large");
    inner_product_kernel<G><<<(uint32_t)ip_blocks, TPB, 0, st>>>(T, key, P, S, c.keyswitch_primes(L), L,
                                                                (uint32_t)c.key_count(), logn, count, tiles, groups,
                                                                diag, diag_stride);
    launch_check("inner_product");
}

void moddown_into(const Context &c, uint64_t *P, uint64_t *S, uint64_t *U, uint64_t *ct, size_t ct_stride,
                  uint32_t count, uint32_t L, hipStream_t st, bool folded) {
    const size_t K = c.key_count();
    const int logn = c.logn;
    const bool ckks = c.scheme == SCHEME_CKKS;
    ntt_inverse(S, count * 2, 1, c.keyswitch_primes(L) + L, logn, NTT_FULL, st);
    const size_t uwords = ((size_t)count * 2 * L) << logn;
    const ModDownArgs md{S, P, ct, ct_stride, c.ks_consts(), c.moduli[K - 1], L, folded};
    if (ntt_moddown(ckks, ckks ? U : P, md, count, c.primes(), logn, st)) return;
    if (folded) fail(ERR_LOGIC, "a folded ModDown needs the fused transforms");
    divround_prep_kernel<<<blocks_for(uwords), TPB, 0, st>>>(S, U, c.primes(), c.ks_consts(), c.moduli[K - 1], L, logn,
                                                            uwords);
    launch_check("moddown_prep");
    if (ckks) ntt_forward(U, count * 2 * L, L, c.primes(), logn, NTT_FULL, st);
    else ntt_inverse(P, count * 2 * L, L, c.primes(), logn, NTT_FULL, st);
    moddown_finish_kernel<<<blocks_for(uwords), TPB, 0, st>>>(P, U, ct, ct_stride, c.primes(), c.ks_consts(), L, logn,
                                                             uwords);
    launch_check("moddown_finish");
}

bool keyswitch_folds(const Context &c) { return c.scheme == SCHEME_CKKS && (c.logn == 15 || c.logn == 16); }

void switch_key(const Context &c, const uint64_t *target, size_t target_stride, uint64_t *ct, size_t ct_stride,
                const uint64_t *key, uint32_t count, uint32_t L, void *ws, size_t ws_bytes, hipStream_t st,
                const uint64_t *fold_c0) {
    if (fold_c0 && !keyswitch_folds(c)) fail(ERR_LOGIC, "a key switch into (c0, 0) needs the fused key products");
    const size_t K = c.key_count();
    if (K < 2) fail(ERR_LOGIC, "keyswitching is not supported by the context");
    if (L < 1 || L > K - 1) fail(ERR_INVALID_ARGUMENT, "level out of range for key switching");
    if (!count) return;
    if (!ws || ws_bytes < ks_layout(c, count, L).total)
        fail(ERR_INVALID_ARGUMENT, "workspace too small for key switching");
    // The ModUp buffer has room for L(L+1) limbs per ciphertext; one NTT launch covers at most
    // 2^34 / N limbs, so large batches run in slices of whole ciphertexts.
    const uint64_t per_ct = (uint64_t)L * (L + 1);
    const uint64_t max_cts = ((uint64_t(1) << 34) >> c.logn) / per_ct;
    if (max_cts == 0) fail(ERR_INVALID_ARGUMENT, "level too large for key switching at this degree");
    KernelProbe probe{nullptr, nullptr};
    const bool armed = c.take_probe(probe);
    for (uint32_t done = 0; done < count;) {
        const uint32_t n = (uint32_t)(count - done < max_cts ? count - done : max_cts);
        switch_key_chunk(c, target + (size_t)done * target_stride, target_stride, ct + (size_t)done * ct_stride,
                         ct_stride, key, n, L, ws, st, armed && done == 0 ? &probe : nullptr,
                         fold_c0 ? fold_c0 + (size_t)done * ct_stride : nullptr);
        done += n;
    }
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
    const uint32_t Lo = L - 1;
    if (!ckks) {
        // coefficient form: one pass, the last limb read beside every kept one
        const size_t total = g * Lo * c.n;
        rescale_coeff_kernel<<<blocks_for(total), TPB, 0, st>>>(in, out, primes, c.rescale_consts(L), c.moduli[Lo], Lo,
                                                                logn, total);
        launch_check("rescale_coeff");
        return;
    }
    if (logn >= 13 && g * Lo <= ((uint64_t(1) << 34) >> logn)) {
        // NTT form: the last limbs' inverse transforms, then the divide-and-round fused into
        // the forward transform of the correction (written once, never re-read)
        ntt_inverse_from(in + (size_t)Lo * c.n, LimbGrid{1, L, 1, (uint32_t)g}, S, LimbGrid{1, 1, 1, (uint32_t)g},
                         primes + Lo, logn, NTT_FULL, st);
        const ModDownArgs md{S, in, out, 2 * (size_t)Lo * c.n, c.rescale_consts(L), c.moduli[Lo], Lo, true,
                             2 * (size_t)L * c.n, L};
        if (ntt_moddown_polys(U, md, (uint32_t)g, primes, logn, st)) return;
    }
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

// ---- level groups in one launch per stage ------------------------------------------

namespace {
template <class F>
void in_chunks(uint32_t n, F &&f) {
    for (uint32_t g0 = 0; g0 < n; g0 += MAX_GROUPS) f(g0, n - g0 < MAX_GROUPS ? n - g0 : MAX_GROUPS);
}
void check_group_level(const Context &c, uint32_t L, uint32_t lo) {
    if (L < lo || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "limb_count out of range for a group");
}
}  // namespace

void ckks_tensor_groups(const Context &c, const TensorGroup *gs, uint32_t n, hipStream_t st) {
    if (c.scheme != SCHEME_CKKS) fail(ERR_INVALID_ARGUMENT, "grouped tensor products are CKKS (NTT form)");
    in_chunks(n, [&](uint32_t g0, uint32_t m) {
        std::vector<ElemGroup> eg;
        for (uint32_t g = g0; g < g0 + m; g++) {
            const TensorGroup &t = gs[g];
            check_group_level(c, t.L, 1);
            const size_t lw = (size_t)t.L * c.n;
            ElemGroup e{};
            e.a = t.a;
            e.b = t.b;
            e.out = t.out;
            e.sa = e.sb = 2 * lw;
            e.so = 3 * lw;
            e.L = t.L;
            e.count = t.count;
            e.total = (size_t)t.count * lw;
            eg.push_back(e);
        }
        launch_elem_groups<OP_TENSOR>(eg, c.primes(), c.logn, st);
    });
}

size_t relinearize_groups_workspace(const Context &c, const RelinGroup *gs, uint32_t n) {
    size_t total = 0;
    for (uint32_t g = 0; g < n; g++) total += ks_layout(c, gs[g].count, gs[g].L).total;
    return total;
}

void relinearize_groups(const Context &c, const RelinGroup *gs, uint32_t n, const uint64_t *key, void *ws,
                        size_t ws_bytes, hipStream_t st) {
    const size_t K = c.key_count();
    if (K < 2) fail(ERR_LOGIC, "keyswitching is not supported by the context");
    for (uint32_t g = 0; g < n; g++) {
        if (gs[g].L < 1 || gs[g].L > K - 1) fail(ERR_INVALID_ARGUMENT, "level out of range for key switching");
        if (!gs[g].ct && gs[g].count) fail(ERR_INVALID_ARGUMENT, "null ciphertext group");
    }
    if (!ws || ws_bytes < relinearize_groups_workspace(c, gs, n))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for key switching");
    const int logn = c.logn;
    const size_t nn = c.n;
    // one launch per stage needs the fused CKKS path and every group within one launch
    bool grouped = c.scheme == SCHEME_CKKS && (logn == 15 || logn == 16);
    for (uint32_t g = 0; g < n && grouped; g++)
        grouped = (uint64_t)gs[g].count * gs[g].L * (gs[g].L + 1) <= ((uint64_t(1) << 34) >> logn);
    char *base = static_cast<char *>(ws);
    if (!grouped) {
        for (uint32_t g = 0; g < n; g++) {
            const size_t poly = (size_t)gs[g].L * nn, bytes = ks_layout(c, gs[g].count, gs[g].L).total;
            switch_key(c, gs[g].ct + 2 * poly, 3 * poly, gs[g].ct, 3 * poly, key, gs[g].count, gs[g].L, base, bytes, st);
            base += bytes;
        }
        return;
    }
    in_chunks(n, [&](uint32_t g0, uint32_t m) {
        NttGroup dg[MAX_GROUPS], sg[MAX_GROUPS];
        KsGroup kg[MAX_GROUPS];
        ModDownGroup mg[MAX_GROUPS];
        uint32_t k = 0;
        for (uint32_t g = g0; g < g0 + m; g++) {
            const RelinGroup &r = gs[g];
            const uint32_t L = r.L;
            const KsLayout lay = ks_layout(c, r.count, L);
            char *const gb = base;  // this group's workspace
            auto at = [gb](size_t off) { return reinterpret_cast<uint64_t *>(gb + off); };
            base += lay.total;
            if (!r.count) continue;
            const size_t poly = (size_t)L * nn;
            uint64_t *target = r.ct + 2 * poly;
            dg[k] = NttGroup{at(lay.d), target, LimbGrid{1, L, L, r.count}, LimbGrid{1, 3 * L, L, r.count}, c.primes()};
            // the products take p times (c0, c1), and the ModDown writes them unread
            kg[k] = KsGroup{at(lay.d), poly, at(lay.T), at(lay.P), at(lay.S), target, 3 * poly, c.keyswitch_primes(L),
                            r.count, L, KsFold{r.ct, 3 * poly, c.ks_consts()}};
            sg[k] = NttGroup{at(lay.S), nullptr, LimbGrid{1, 1, 1, 2 * r.count}, LimbGrid{1, 1, 1, 2 * r.count},
                             c.keyswitch_primes(L) + L};
            mg[k] = ModDownGroup{
                at(lay.U), ModDownArgs{at(lay.S), at(lay.P), r.ct, 3 * poly, c.ks_consts(), c.moduli[K - 1], L, true},
                r.count};
            k++;
        }
        if (!k) return;
        // 1. digits: the c2 polynomials to coefficient form (they stay in NTT form for the
        // own-prime terms); 2. ModUp fused with the key products; 3. ModDown
        ntt_inverse_groups(dg, k, logn, NTT_FULL, st);
        ntt_modup_keyprod_groups(kg, k, key, (uint32_t)K, c.primes(), logn, st);
        ntt_inverse_groups(sg, k, logn, NTT_FULL, st);
        ntt_moddown_groups(mg, k, c.primes(), logn, st);
    });
}

size_t rescale_groups_workspace(const Context &c, const RescaleGroup *gs, uint32_t n) {
    size_t total = 0;
    for (uint32_t g = 0; g < n; g++) {
        if (gs[g].L < 2) continue;
        total += rescale_workspace(c, gs[g].count, 2, gs[g].L);
    }
    return total;
}

void rescale_groups(const Context &c, const RescaleGroup *gs, uint32_t n, uint32_t in_size, void *ws, size_t ws_bytes,
                    hipStream_t st) {
    if (in_size < 2) fail(ERR_INVALID_ARGUMENT, "ciphertexts of at least two polynomials are rescaled");
    for (uint32_t g = 0; g < n; g++) {
        if (gs[g].L < 2 || gs[g].L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "end of modulus switching chain reached");
        if (gs[g].count && (!gs[g].in || !gs[g].out)) fail(ERR_INVALID_ARGUMENT, "null ciphertext group");
        if (gs[g].in == gs[g].out && gs[g].count) fail(ERR_INVALID_ARGUMENT, "rescale cannot run in place");
    }
    if (!ws || ws_bytes < rescale_groups_workspace(c, gs, n)) fail(ERR_INVALID_ARGUMENT, "workspace too small for rescale");
    const bool ckks = c.scheme == SCHEME_CKKS;
    const int logn = c.logn;
    const size_t nn = c.n;
    const DevPrime *primes = c.primes();
    char *base = static_cast<char *>(ws);
    in_chunks(n, [&](uint32_t g0, uint32_t m) {
        std::vector<ElemGroup> last, prep, fin;
        NttGroup sg[MAX_GROUPS], ug[MAX_GROUPS];
        ModDownGroup mg[MAX_GROUPS];
        bool fused = ckks && logn >= 13;
        uint32_t k = 0;
        for (uint32_t g = g0; g < g0 + m; g++) {
            const RescaleGroup &r = gs[g];
            const uint32_t L = r.L, Lo = L - 1;
            const size_t bytes = rescale_workspace(c, r.count, 2, L);
            uint64_t *S = reinterpret_cast<uint64_t *>(base);
            uint64_t *U = reinterpret_cast<uint64_t *>(base + ((2 * (size_t)r.count * nn * 8 + 255) & ~size_t(255)));
            base += bytes;
            if (!r.count) continue;
            const size_t sa = (size_t)in_size * L * nn;
            ElemGroup e{};
            e.a = r.in;
            e.sa = sa;
            e.out = S;
            e.L = L;
            e.count = r.count;
            e.total = 2 * (size_t)r.count * nn;
            last.push_back(e);
            ElemGroup p{};
            p.b = S;
            p.out = U;
            p.dc = c.rescale_consts(L);
            p.p = c.moduli[L - 1];
            p.L = Lo;
            p.count = r.count;
            p.total = 2 * (size_t)r.count * Lo * nn;
            prep.push_back(p);
            ElemGroup f{};
            f.a = r.in;
            f.sa = sa;
            f.aux = U;
            f.out = r.out;
            f.so = 2 * (size_t)Lo * nn;
            f.dc = c.rescale_consts(L);
            f.L = Lo;
            f.count = r.count;
            f.total = 2 * (size_t)r.count * Lo * nn;
            fin.push_back(f);
            sg[k] = NttGroup{S, nullptr, LimbGrid{1, 1, 1, 2 * r.count}, LimbGrid{1, 1, 1, 2 * r.count}, primes + (L - 1)};
            ug[k] = NttGroup{U, nullptr, LimbGrid{1, Lo, Lo, 2 * r.count}, LimbGrid{1, Lo, Lo, 2 * r.count}, primes};
            // the divide-and-round fused into the correction's forward transform (folded:
            // out written once from in, the correction never re-read)
            mg[k] = ModDownGroup{U, ModDownArgs{S, r.in, r.out, 2 * (size_t)Lo * nn, c.rescale_consts(L),
                                                c.moduli[L - 1], Lo, true, sa, L},
                                 r.count};
            if ((uint64_t)2 * r.count * Lo > ((uint64_t(1) << 34) >> logn)) fused = false;
            k++;
        }
        if (!k) return;
        launch_elem_groups<OP_LAST_LIMB>(last, primes, logn, st);
        if (ckks) ntt_inverse_groups(sg, k, logn, NTT_FULL, st);
        if (fused && ntt_moddown_groups(mg, k, primes, logn, st)) return;
        launch_elem_groups<OP_DIVROUND_PREP>(prep, primes, logn, st);
        if (ckks) ntt_forward_groups(ug, k, logn, NTT_FULL, st);
        launch_elem_groups<OP_RESCALE_FINISH>(fin, primes, logn, st);
    });
}

}  // namespace shc
