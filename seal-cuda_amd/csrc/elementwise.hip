// elementwise.hip -- coefficient-wise evaluator operations.
//
// Evaluator::negateInplace / addInplace / subInplace (src/evaluator.cpp:149-357; GPU
// src/evaluator_cuda.cu:145-281), addPlainInplace / subPlainInplace (evaluator.cpp:
// 1603-1764, GPU :1544-1690; BFV through src/utils/scalingvariant.cpp:53-131),
// multiplyPlainNtt (evaluator.cpp:1932, GPU :1811) and the GPU multiplyPlainNormal
// (src/evaluator_cuda.cu:1726-1807: the plaintext is always lifted, m >= ceil(t/2) ->
// m + q_i - t, then NTT, dyadic product, inverse NTT).  The CPU multiplyPlainNormal
// multiplies by a monomial plaintext without lifting it (evaluator.cpp:1818-1866), so
// for a monomial with m >= ceil(t/2) the two reference paths give different words;
// this build follows the GPU path it replaces.
#include "elementwise.h"
#include "accum.h"
#include "behz.h"
#include "crypto.h"
#include "modarith.h"
#include "ntt.h"

namespace shc {
namespace {

constexpr int TPB = 256;

inline uint32_t blocks_for(size_t n) {
    const size_t b = (n + TPB - 1) / TPB;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

enum BinOp { BIN_ADD, BIN_SUB, BIN_MUL_PLAIN };

template <int OP>
__global__ void binary_kernel(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b, uint64_t *out,
                              const DevPrime *__restrict__ primes, uint32_t L, int logn, size_t total) {
    const size_t limb_words = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t r = idx % limb_words;
        const DevPrime &P = primes[r >> logn];
        const uint64_t x = a[idx];
        uint64_t o;
        if (OP == BIN_ADD) o = addmod(x, b[idx], P.q);
        else if (OP == BIN_SUB) o = submod(x, b[idx], P.q);
        else o = mulmod(x, b[r], P.q, P.ratio_lo, P.ratio_hi);
        out[idx] = o;
    }
}

// The same two words per thread with 16-byte accesses (all pointers 16-byte aligned;
// N is even, so a pair never straddles two limbs).  From N = 128 a wave's 128 words lie
// in one limb, and the limb index is wave-uniform.
__device__ __forceinline__ uint32_t pair_limb(size_t w, int logn, uint32_t L) {
    uint32_t r = (uint32_t)(w >> logn);
    if (logn >= 7) r = __builtin_amdgcn_readfirstlane(r);
    return r % L;
}

template <int OP>
__global__ void binary_pairs_kernel(const ulonglong2 *__restrict__ a, const uint64_t *__restrict__ b, ulonglong2 *out,
                                    const DevPrime *__restrict__ primes, uint32_t L, int logn, size_t pairs) {
    const size_t n = (size_t)1 << logn;
    for (size_t p = (size_t)blockIdx.x * TPB + threadIdx.x; p < pairs; p += (size_t)gridDim.x * TPB) {
        const size_t w = 2 * p;
        const uint32_t l = pair_limb(w, logn, L);
        const DevPrime &P = primes[l];
        const ulonglong2 x = a[p];
        const ulonglong2 y = OP == BIN_MUL_PLAIN ? *reinterpret_cast<const ulonglong2 *>(b + ((size_t)l << logn) + (w & (n - 1)))
                                                 : reinterpret_cast<const ulonglong2 *>(b)[p];
        ulonglong2 o;
        if (OP == BIN_ADD) {
            o.x = addmod(x.x, y.x, P.q);
            o.y = addmod(x.y, y.y, P.q);
        } else if (OP == BIN_SUB) {
            o.x = submod(x.x, y.x, P.q);
            o.y = submod(x.y, y.y, P.q);
        } else {
            o.x = mulmod(x.x, y.x, P.q, P.ratio_lo, P.ratio_hi);
            o.y = mulmod(x.y, y.y, P.q, P.ratio_lo, P.ratio_hi);
        }
        out[p] = o;
    }
}

__global__ void negate_pairs_kernel(const ulonglong2 *__restrict__ a, ulonglong2 *out,
                                    const DevPrime *__restrict__ primes, uint32_t L, int logn, size_t pairs) {
    for (size_t p = (size_t)blockIdx.x * TPB + threadIdx.x; p < pairs; p += (size_t)gridDim.x * TPB) {
        const uint64_t q = primes[pair_limb(2 * p, logn, L)].q;
        const ulonglong2 x = a[p];
        out[p] = make_ulonglong2(negmod(x.x, q), negmod(x.y, q));
    }
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <int OP>
void launch_binary(const uint64_t *a, const uint64_t *b, uint64_t *out, const DevPrime *primes, uint32_t L, int logn,
                   size_t total, hipStream_t st) {
    if (logn >= 1 && aligned16(a) && aligned16(b) && aligned16(out)) {
        const size_t pairs = total / 2;
        binary_pairs_kernel<OP><<<blocks_for(pairs), TPB, 0, st>>>(reinterpret_cast<const ulonglong2 *>(a), b,
                                                                    reinterpret_cast<ulonglong2 *>(out), primes, L,
                                                                    logn, pairs);
    } else {
        binary_kernel<OP><<<blocks_for(total), TPB, 0, st>>>(a, b, out, primes, L, logn, total);
    }
}

__global__ void negate_kernel(const uint64_t *__restrict__ a, uint64_t *out, const DevPrime *__restrict__ primes,
                              uint32_t L, int logn, size_t total) {
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB)
        out[idx] = negmod(a[idx], primes[(idx >> logn) % L].q);
}

__global__ void add_plain_poly0_kernel(const uint64_t *ct, const uint64_t *__restrict__ plain, uint64_t *out,
                                       const DevPrime *__restrict__ primes, uint32_t size, uint32_t L, int logn,
                                       bool subtract, size_t total) {
    const size_t limb_words = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t r = idx % limb_words;
        const uint64_t x = ct[idx];
        if ((idx / limb_words) % size == 0) {
            const uint64_t q = primes[r >> logn].q;
            out[idx] = subtract ? submod(x, plain[r], q) : addmod(x, plain[r], q);
        } else {
            out[idx] = x;
        }
    }
}

// lifted[l][c] = (m - t [m >= ceil(t/2)]) mod q_l for m = plain[c] < t.
__global__ void plain_lift_kernel(const uint64_t *__restrict__ plain, uint64_t *__restrict__ lifted, uint64_t t,
                                  const DevPrime *__restrict__ primes, uint32_t L, int logn, size_t total) {
    const uint64_t thr = (t + 1) >> 1;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t c = idx & (((size_t)1 << logn) - 1);
        const DevPrime &P = primes[idx >> logn];
        const uint64_t m = plain[c];
        const uint64_t mq = m >= P.q ? barrett64(m, P.q, P.ratio_hi) : m;
        lifted[idx] = m >= thr ? submod(mq, t >= P.q ? barrett64(t, P.q, P.ratio_hi) : t, P.q) : mq;
    }
}

// out[h][ct][p][l][c] = sum_j plains[h][j][l][c] * cts[j][ct][p][l][c] mod q_l for the
// H outputs h0 .. h0 + H - 1 (hn of them valid; the rest repeat the last, unstored).
// A thread owns one word position (l, c) of G consecutive polynomials (ct, p): each
// ciphertext word it loads serves H outputs (the diagonal matrix-vector products' giant
// steps) and each plaintext word G polynomials, so the plaintexts -- shared by every
// ciphertext of the batch -- are loaded polys / G times instead of polys times (and the
// block order below serves the repeats from L2).  Sums in three words (Acc3), folded mod
// q every 64 terms.  With G = 2 and H = 8 the sums fit three waves per SIMD (the launch
// bound), which measured faster than G = 4 at two waves.
//
// xcd: the grid is exactly groups x (lw / TPB) blocks with lw / TPB a multiple of 8, and
// block B (dispatched to XCD B % 8) takes position block 8 (B / 8 / groups) + B % 8 of
// polynomial group (B / 8) % groups: consecutive blocks of one XCD share their plaintext
// words, which then come from that XCD's L2 after the first.
template <int H, int G>
__global__ __launch_bounds__(TPB, 3) void plain_dot_many_kernel(const uint64_t *__restrict__ cts, const uint64_t *__restrict__ plains,
                                      uint64_t *__restrict__ out, const DevPrime *__restrict__ primes, uint32_t terms,
                                      uint32_t h0, uint32_t hn, uint32_t polys, uint32_t L, int logn, size_t total,
                                      bool xcd) {
    const size_t lw = (size_t)L << logn, pw = (size_t)polys * lw;
    size_t first = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (xcd) {
        const uint32_t groups = (polys + G - 1) / G, k = blockIdx.x >> 3;
        first = (size_t)(k % groups) * lw + ((size_t)(k / groups) * 8 + (blockIdx.x & 7)) * TPB + threadIdx.x;
    }
    for (size_t idx = first; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t row = uniform_row(idx, logn), grp = row / L;  // idx = grp * lw + r
        const size_t r = idx - (size_t)grp * lw;
        const uint32_t p0 = grp * G, gn = min<uint32_t>(G, polys - p0);
        const DevPrime &P = primes[row - grp * L];
        const uint64_t *pp[H];
#pragma unroll
        for (int h = 0; h < H; h++) pp[h] = plains + (size_t)(h0 + min<uint32_t>(h, hn - 1)) * terms * lw + r;
        const uint64_t *cp = cts + (size_t)p0 * lw + r;
        Acc3 acc[G][H];
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int h = 0; h < H; h++) acc[g][h] = Acc3{0, 0, 0};
        for (uint32_t j = 0; j < terms; j++) {
            uint64_t pv[H];
#pragma unroll
            for (int h = 0; h < H; h++) pv[h] = pp[h][(size_t)j * lw];
#pragma unroll
            for (int g = 0; g < G; g++) {
                if ((uint32_t)g < gn) {  // uniform: a wave shares its polynomial group
                    const uint64_t t = cp[(size_t)j * pw + (size_t)g * lw];
#pragma unroll
                    for (int h = 0; h < H; h++) mac(acc[g][h], t, pv[h]);
                }
            }
            if ((j & 63) == 63) {
#pragma unroll
                for (int g = 0; g < G; g++)
#pragma unroll
                    for (int h = 0; h < H; h++) acc_fold(acc[g][h], P);
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
            if ((uint32_t)g < gn)
#pragma unroll
                for (int h = 0; h < H; h++)
                    if ((uint32_t)h < hn) out[(size_t)(h0 + h) * pw + (size_t)(p0 + g) * lw + r] = acc_reduce(acc[g][h], P);
    }
}

template <int H>
void launch_plain_dot(uint32_t G, const uint64_t *cts, const uint64_t *plains, uint64_t *out, const DevPrime *primes,
                      uint32_t terms, uint32_t h0, uint32_t hn, uint32_t polys, uint32_t L, int logn, hipStream_t st) {
    const size_t groups = (polys + G - 1) / G, lw = (size_t)L << logn, total = groups * lw;
    const size_t rblocks = lw / TPB;
    // the XCD-aware order where the grid is one block per (group, position block)
    const bool xcd = groups > 1 && lw % TPB == 0 && rblocks % 8 == 0 && groups * rblocks < (size_t(1) << 20);
    if (G == 2)
        plain_dot_many_kernel<H, 2><<<blocks_for(total), TPB, 0, st>>>(cts, plains, out, primes, terms, h0, hn, polys, L,
                                                                        logn, total, xcd);
    else
        plain_dot_many_kernel<H, 1><<<blocks_for(total), TPB, 0, st>>>(cts, plains, out, primes, terms, h0, hn, polys, L,
                                                                        logn, total, xcd);
}

void check_level(const Context &c, uint32_t L) {
    if (L < 1 || L > c.key_count()) fail(ERR_INVALID_ARGUMENT, "limb count out of range");
}

}  // namespace

void poly_add(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys, uint32_t L,
              hipStream_t st) {
    check_level(c, L);
    const size_t total = ((size_t)polys * L) << c.logn;
    if (!total) return;
    launch_binary<BIN_ADD>(a, b, out, c.primes(), L, c.logn, total, st);
    launch_check("add");
}

void poly_sub(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys, uint32_t L,
              hipStream_t st) {
    check_level(c, L);
    const size_t total = ((size_t)polys * L) << c.logn;
    if (!total) return;
    launch_binary<BIN_SUB>(a, b, out, c.primes(), L, c.logn, total, st);
    launch_check("sub");
}

void poly_negate(const Context &c, const uint64_t *a, uint64_t *out, uint32_t polys, uint32_t L, hipStream_t st) {
    check_level(c, L);
    const size_t total = ((size_t)polys * L) << c.logn;
    if (!total) return;
    if (c.logn >= 1 && aligned16(a) && aligned16(out))
        negate_pairs_kernel<<<blocks_for(total / 2), TPB, 0, st>>>(reinterpret_cast<const ulonglong2 *>(a),
                                                                   reinterpret_cast<ulonglong2 *>(out), c.primes(), L,
                                                                   c.logn, total / 2);
    else
        negate_kernel<<<blocks_for(total), TPB, 0, st>>>(a, out, c.primes(), L, c.logn, total);
    launch_check("negate");
}

void multiply_plain_ntt(const Context &c, const uint64_t *a, const uint64_t *plain, uint64_t *out, uint32_t polys,
                        uint32_t L, hipStream_t st) {
    check_level(c, L);
    const size_t total = ((size_t)polys * L) << c.logn;
    if (!total) return;
    launch_binary<BIN_MUL_PLAIN>(a, plain, out, c.primes(), L, c.logn, total, st);
    launch_check("multiply_plain");
}

void add_plain_poly0(const Context &c, const uint64_t *ct, const uint64_t *plain, uint64_t *out, uint32_t count,
                     uint32_t size, uint32_t L, bool subtract, hipStream_t st) {
    check_level(c, L);
    if (size < 1) fail(ERR_INVALID_ARGUMENT, "ciphertext size must be at least 1");
    const size_t total = (((size_t)count * size) * L) << c.logn;
    if (!total) return;
    add_plain_poly0_kernel<<<blocks_for(total), TPB, 0, st>>>(ct, plain, out, c.primes(), size, L, c.logn, subtract,
                                                              total);
    launch_check("add_plain");
}

void bfv_add_plain(const Context &c, const uint64_t *ct, const uint64_t *plain, uint64_t *out, uint32_t count,
                   uint32_t size, uint32_t L, bool subtract, hipStream_t st) {
    if (c.scheme != SCHEME_BFV) fail(ERR_LOGIC, "scaled plaintext addition is for BFV");
    if (L < 1 || L >= c.key_count()) fail(ERR_INVALID_ARGUMENT, "level out of range");
    if (size < 1) fail(ERR_INVALID_ARGUMENT, "ciphertext size must be at least 1");
    const size_t ct_words = ((size_t)size * L) << c.logn;
    if (out != ct) SHC_HIP(hipMemcpyAsync(out, ct, (size_t)count * ct_words * 8, hipMemcpyDeviceToDevice, st));
    for (uint32_t i = 0; i < count; i++) bfv_scale_add(c, out + i * ct_words, plain, L, subtract, st);
}

void bfv_plain_to_ntt(const Context &c, const uint64_t *plain, uint32_t L, uint64_t *out, hipStream_t st) {
    if (c.scheme != SCHEME_BFV || !c.plain) fail(ERR_LOGIC, "plaintext lift is for BFV");
    check_level(c, L);
    const size_t lw = (size_t)L << c.logn;
    plain_lift_kernel<<<blocks_for(lw), TPB, 0, st>>>(plain, out, c.plain, c.primes(), L, c.logn, lw);
    launch_check("plain_lift");
    ntt_forward(out, L, L, c.primes(), c.logn, NTT_FULL, st);
}

void multiply_plain_accumulate(const Context &c, const uint64_t *cts, const uint64_t *plains, uint32_t terms,
                               uint32_t polys, uint32_t L, uint64_t *out, hipStream_t st) {
    multiply_plain_accumulate_many(c, cts, plains, terms, 1, polys, L, out, st);
}

void multiply_plain_accumulate_many(const Context &c, const uint64_t *cts, const uint64_t *plains, uint32_t terms,
                                    uint32_t outputs, uint32_t polys, uint32_t L, uint64_t *out, hipStream_t st) {
    check_level(c, L);
    if (!terms) fail(ERR_INVALID_ARGUMENT, "no terms to accumulate");
    if (!outputs) fail(ERR_INVALID_ARGUMENT, "no outputs to accumulate");
    if (!polys) return;
    // two polynomials per thread (the plaintexts' reuse) -- a ciphertext's pair -- one
    // for a lone polynomial
    const uint32_t G = polys >= 2 ? 2 : 1;
    const DevPrime *primes = c.primes();
    const int logn = c.logn;
    for (uint32_t h0 = 0; h0 < outputs;) {
        const uint32_t left = outputs - h0;
        const uint32_t hn = left >= 8 ? 8 : left;
        if (hn > 4) launch_plain_dot<8>(G, cts, plains, out, primes, terms, h0, hn, polys, L, logn, st);
        else if (hn > 2) launch_plain_dot<4>(G, cts, plains, out, primes, terms, h0, hn, polys, L, logn, st);
        else if (hn == 2) launch_plain_dot<2>(G, cts, plains, out, primes, terms, h0, hn, polys, L, logn, st);
        else launch_plain_dot<1>(G, cts, plains, out, primes, terms, h0, hn, polys, L, logn, st);
        launch_check("plain_dot");
        h0 += hn;
    }
}

size_t bfv_multiply_plain_workspace(const Context &c, uint32_t L) { return ((size_t)L << c.logn) * 8; }

void bfv_multiply_plain(const Context &c, const uint64_t *a, const uint64_t *plain, uint64_t *out, uint32_t polys,
                        uint32_t L, void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme != SCHEME_BFV || !c.plain) fail(ERR_LOGIC, "coefficient-form plain multiply is for BFV");
    check_level(c, L);
    if (!ws || ws_bytes < bfv_multiply_plain_workspace(c, L))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for plain multiply");
    const size_t words = ((size_t)polys * L) << c.logn;
    if (!words) return;
    uint64_t *lifted = static_cast<uint64_t *>(ws);
    bfv_plain_to_ntt(c, plain, L, lifted, st);
    if (out != a) SHC_HIP(hipMemcpyAsync(out, a, words * 8, hipMemcpyDeviceToDevice, st));
    ntt_forward(out, polys * L, L, c.primes(), c.logn, NTT_FULL, st);
    multiply_plain_ntt(c, out, lifted, out, polys, L, st);
    ntt_inverse(out, polys * L, L, c.primes(), c.logn, NTT_FULL, st);
}

}  // namespace shc
