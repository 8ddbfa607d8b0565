// rotations.h -- hoisted rotations: one ModUp of c1 shared by many Galois elements.
#pragma once
#include "context.h"

namespace shc {

// A Galois key prepared for hoisting at level L: the key permuted by sigma^-1 in the
// NTT domain ([K-1][2][K][N]) followed by the sign-pattern correction corr'[2][L+1][N]
// and two header words naming L and the Galois element, which rotate_sum / rotate_many
// check (a key prepared for another level or element is rejected).
size_t hoisted_key_words(const Context &c, uint32_t L);
size_t hoisted_key_workspace(const Context &c, uint32_t L);
void hoisted_key_prepare(const Context &c, const uint64_t *galois_key, uint32_t elt, uint32_t L, uint64_t *out,
                         void *ws, size_t ws_bytes, hipStream_t st);

// out[ct] = sum_r rotate_{elt_r}(in[ct]) for ciphertexts [count][2][L][N], with the same
// words as summing `rotations` separate apply_galois_keyswitch results.  hoisted[r] is
// the prepared key of elts[r] at level L, keys[r] its Galois key (used when a digit of
// c1 is zero, where hoisting is not exact, for the whole call).  Synchronises the
// stream once (the zero-digit test), so it cannot be captured in a graph.
size_t rotate_sum_workspace(const Context &c, uint32_t count, uint32_t L);
void rotate_sum(const Context &c, const uint64_t *in, uint32_t count, uint32_t L, const uint32_t *elts,
                const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                void *ws, size_t ws_bytes, hipStream_t st);

// out[r][ct] = rotate_{elt_r}(in[ct]) (out: [rotations][count][2][L][N]), the words of
// separate apply_galois_keyswitch calls, with one ModUp for all rotations.
void rotate_many(const Context &c, const uint64_t *in, uint32_t count, uint32_t L, const uint32_t *elts,
                 const uint64_t *const *hoisted, const uint64_t *const *keys, uint32_t rotations, uint64_t *out,
                 void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace shc
