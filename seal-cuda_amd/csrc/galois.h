// galois.h -- Galois automorphisms and the rotation key switch.
#pragma once
#include "context.h"
#include "rns_ops.h"

namespace shc {

// out[g] = sigma_elt(in[g]) for `polys` polynomials of L limbs; strides in words.
// ntt_form selects the NTT-domain permutation (CKKS) or the signed coefficient
// permutation (BFV).  in and out must not overlap.
void apply_galois(const Context &c, const uint64_t *in, size_t in_stride, uint64_t *out, size_t out_stride,
                  uint32_t polys, uint32_t L, uint32_t elt, bool ntt_form, hipStream_t st);

size_t galois_keyswitch_workspace(const Context &c, uint32_t count, uint32_t L);

// ct: [count][2][L][N] in place: (sigma(c0), 0) + switch_key(sigma(c1), key).
void apply_galois_keyswitch(const Context &c, uint64_t *ct, uint32_t count, uint32_t L, uint32_t elt,
                            const uint64_t *key, void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace shc
