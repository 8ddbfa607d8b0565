// galois.hip -- Galois automorphisms X -> X^elt on RNS polynomials.
//
// Same maps as GaloisTool::applyGalois / applyGaloisNtt (src/utils/galois.cpp:13-30,
// 143-175; GPU twins src/utils/galois_cuda.cu:139-196), written as gathers so every
// store is coalesced:
//   coefficient form: out[k] = (-1)^s in[i] with i = k elt^-1 mod N and
//                     s = bit logN of i*elt (the reference scatters out[i elt mod N]);
//   NTT form:         out[i] = in[bitrev(((elt * bitrev(N + i)) >> 1) mod N)],
//                     the permutation table the reference builds on the host and
//                     uploads per element, computed here in-register from i.
#include "galois.h"
#include "modarith.h"

namespace shc {
namespace {

constexpr int TPB = 256;

inline uint32_t blocks_for(size_t n) {
    const size_t b = (n + TPB - 1) / TPB;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

// in/out: `polys` polynomials of L limbs; polynomial g of in starts at in + g * in_stride.
__global__ void galois_ntt_kernel(const uint64_t *__restrict__ in, size_t in_stride, uint64_t *__restrict__ out,
                                  size_t out_stride, uint32_t L, int logn, uint32_t elt, size_t total) {
    const uint32_t n = 1u << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t i = (uint32_t)(idx & (n - 1));
        const uint32_t r = uniform_row(idx, logn), g32 = r / L;
        const size_t g = g32, l = r - g32 * L;
        const uint32_t rev = rev_bits(n + i, logn + 1);
        const uint32_t raw = (uint32_t)((((uint64_t)elt * rev) >> 1) & (n - 1));
        const uint32_t src = rev_bits(raw, logn);
        out[g * out_stride + (l << logn) + i] = in[g * in_stride + (l << logn) + src];
    }
}

__global__ void galois_coeff_kernel(const uint64_t *__restrict__ in, size_t in_stride, uint64_t *__restrict__ out,
                                    size_t out_stride, const DevPrime *__restrict__ primes, uint32_t L, int logn,
                                    uint32_t elt, uint32_t elt_inv, size_t total) {
    const uint32_t n = 1u << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const uint32_t k = (uint32_t)(idx & (n - 1));
        const uint32_t r = uniform_row(idx, logn), g32 = r / L;
        const size_t g = g32, l = r - g32 * L;
        const uint32_t i = (uint32_t)(((uint64_t)k * elt_inv) & (n - 1));
        const uint64_t raw = (uint64_t)i * elt;
        uint64_t v = in[g * in_stride + (l << logn) + i];
        if ((raw >> logn) & 1) v = negmod(v, primes[l].q);
        out[g * out_stride + (l << logn) + k] = v;
    }
}

// c0 <- c1, c1 <- 0 for each ciphertext (a kernel rather than memcpy/memset so the
// sequence stays capturable in a HIP graph).
__global__ void shift_zero_kernel(uint64_t *__restrict__ ct, uint32_t L, int logn, size_t total) {
    const size_t poly = (size_t)L << logn;
    for (size_t idx = (size_t)blockIdx.x * TPB + threadIdx.x; idx < total; idx += (size_t)gridDim.x * TPB) {
        const size_t g = uniform_row(idx, logn) / L, r = idx - g * poly;
        uint64_t *base = ct + g * 2 * poly;
        base[r] = base[poly + r];
        base[poly + r] = 0;
    }
}

}  // namespace

void apply_galois(const Context &c, const uint64_t *in, size_t in_stride, uint64_t *out, size_t out_stride,
                  uint32_t polys, uint32_t L, uint32_t elt, bool ntt_form, hipStream_t st) {
    const uint64_t m = 2 * c.n;
    if (!(elt & 1) || elt >= m) fail(ERR_INVALID_ARGUMENT, "Galois element is not valid");
    if (in == out) fail(ERR_INVALID_ARGUMENT, "apply_galois cannot run in place");
    const size_t total = ((size_t)polys * L) << c.logn;
    if (!total) return;
    if (ntt_form) {
        galois_ntt_kernel<<<blocks_for(total), TPB, 0, st>>>(in, in_stride, out, out_stride, L, c.logn, elt, total);
    } else {
        const uint32_t elt_inv = (uint32_t)hm::invmod(elt, m);
        galois_coeff_kernel<<<blocks_for(total), TPB, 0, st>>>(in, in_stride, out, out_stride, c.primes(), L, c.logn,
                                                               elt, elt_inv, total);
    }
    launch_check("apply_galois");
}

size_t galois_keyswitch_workspace(const Context &c, uint32_t count, uint32_t L) {
    const size_t sigma = ((size_t)count * L * c.n * 8 + 255) & ~size_t(255);
    return sigma + keyswitch_workspace(c, count, L);
}

void apply_galois_keyswitch(const Context &c, uint64_t *ct, uint32_t count, uint32_t L, uint32_t elt,
                            const uint64_t *key, void *ws, size_t ws_bytes, hipStream_t st) {
    // Evaluator::applyGaloisInplace (src/evaluator.cpp:2210-2250): c0 <- sigma(c0),
    // target <- sigma(c1), c1 <- 0, then switch the target with the Galois key.
    if (!count) return;
    if (!ws || ws_bytes < galois_keyswitch_workspace(c, count, L))
        fail(ERR_INVALID_ARGUMENT, "workspace too small for Galois key switching");
    const bool ntt = c.scheme == SCHEME_CKKS;
    const size_t poly = (size_t)L * c.n;
    uint64_t *sigma = static_cast<uint64_t *>(ws);
    const size_t sigma_bytes = ((size_t)count * poly * 8 + 255) & ~size_t(255);
    void *ks_ws = static_cast<char *>(ws) + sigma_bytes;
    // sigma(c1) first (into scratch), then sigma(c0) into the row of c1
    apply_galois(c, ct + poly, 2 * poly, sigma, poly, count, L, elt, ntt, st);
    apply_galois(c, ct, 2 * poly, ct + poly, 2 * poly, count, L, elt, ntt, st);
    if (keyswitch_folds(c)) {
        // the fused key products read sigma(c0) where it lies and take c1 = 0; the ModDown
        // writes both polynomials
        switch_key(c, sigma, poly, ct, 2 * poly, key, count, L, ks_ws, ws_bytes - sigma_bytes, st, ct + poly);
        return;
    }
    const size_t words = (size_t)count * poly;
    shift_zero_kernel<<<blocks_for(words), TPB, 0, st>>>(ct, L, c.logn, words);
    launch_check("shift_zero");
    switch_key(c, sigma, poly, ct, 2 * poly, key, count, L, ks_ws, ws_bytes - sigma_bytes, st);
}

}  // namespace shc
