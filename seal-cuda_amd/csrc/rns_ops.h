// rns_ops.h -- elementwise RNS kernels and the key-switching pipeline.
#pragma once
#include "common.h"
#include "context.h"

namespace shc {

// out_k = sum_{i+j=k} a_i * b_j (dyadic, NTT form) for `count` pairs of size-2
// ciphertexts with L limbs; a, b: [count][2][L][N], out: [count][3][L][N].
void ckks_tensor(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count, uint32_t L,
                 hipStream_t st);

// The same tensor product with explicit per-ciphertext strides (in words) and an
// arbitrary prime list (limb l uses primes[l]).
void tensor_strided(const DevPrime *primes, const uint64_t *a, size_t a_stride, const uint64_t *b, size_t b_stride,
                    uint64_t *out, size_t out_stride, uint32_t count, uint32_t L, int logn, hipStream_t st);

// Bytes of scratch device memory switch_key needs for `count` ciphertexts at level L.
size_t keyswitch_workspace(const Context &c, uint32_t count, uint32_t L);

// ct_k += ModDown(sum_j NTT(ModUp(d_j)) * key[j][k]) for k = 0, 1 (src/evaluator.cpp:2310-2653).
//   target: c2 of each ciphertext, element stride `target_stride` words between ciphertexts
//           (CKKS: NTT form; BFV: coefficient form), L limbs each.
//   ct:     c0,c1 of each ciphertext, stride `ct_stride` words between ciphertexts.
//   key:    [K-1][2][K][N] key-switching key at key level (NTT form).
// The target is read only.  fold_c0 (only where keyswitch_folds(c)): the switch adds
// into (c0, 0) with c0 at fold_c0 + ct * ct_stride, and ct_0, ct_1 are written without
// being read -- a Galois switch, whose permuted c0 needs no copy and whose c1 no zeros.
bool keyswitch_folds(const Context &c);
void switch_key(const Context &c, const uint64_t *target, size_t target_stride, uint64_t *ct, size_t ct_stride,
                const uint64_t *key, uint32_t count, uint32_t L, void *ws, size_t ws_bytes, hipStream_t st,
                const uint64_t *fold_c0 = nullptr);

// Stages of a key switch at level L over `count` ciphertexts, for callers that compose
// them differently (hoisted rotations):
//   modup_to_T: T[i][ct][j] = NTT_i(d_j mod q_i), digits d (coefficient form) at
//               digits + ct * dstride + j * N; targets i <= L (special last); values
//               in [0, 8 q_i) (L <= 31) or [0, 4 q_i).  skip_diag (CKKS, whose inputs
//               are NTT-form): the limbs (i, i) are left out where N >= 2^13 and the
//               targets i < L hold L - 1 limbs per ciphertext; returns whether it did.
//   keyswitch_inner_product: P[ct][k][i] = sum_j T[i][ct][j] key[j][k][ki] (i < L),
//               S[ct][k] the special target's; with diag (after a skipped diagonal)
//               the term j == i reads the NTT-form input limb diag + ct * diag_stride
//               + i * N.
//   moddown_into: ct_k[i] += ModDown(P_k, S_k)[i] (folded: P holds p ct_k, and ct_k = ModDown; S is consumed; U is scratch
//               [count][2][L][N]; P in NTT form).
// The regions of a key-switch workspace of `count` ciphertexts at level L.
struct KsRegions {
    uint64_t *d, *T, *P, *S, *U;
};
KsRegions keyswitch_regions(const Context &c, uint32_t count, uint32_t L, void *ws);
bool modup_to_T(const Context &c, const uint64_t *digits, size_t dstride, uint64_t *T, uint32_t count, uint32_t L,
                hipStream_t st, const KernelProbe *probe = nullptr, bool skip_diag = false);
void keyswitch_inner_product(const Context &c, const uint64_t *T, const uint64_t *key, uint64_t *P, uint64_t *S,
                             uint32_t count, uint32_t L, hipStream_t st, const uint64_t *diag = nullptr,
                             size_t diag_stride = 0);
void moddown_into(const Context &c, uint64_t *P, uint64_t *S, uint64_t *U, uint64_t *ct, size_t ct_stride,
                  uint32_t count, uint32_t L, hipStream_t st, bool folded = false);

// Divide-and-round by the last prime of level L, dropping it (CKKS rescale in NTT form,
// BFV modulus switch in coefficient form); in: [count][polys][L][N], out: [count][polys][L-1][N].
size_t rescale_workspace(const Context &c, uint32_t count, uint32_t polys, uint32_t L);
void rescale_to_next(const Context &c, const uint64_t *in, uint64_t *out, uint32_t count, uint32_t polys, uint32_t L,
                     void *ws, size_t ws_bytes, hipStream_t st);

// Level groups: each group is a batch at one level with its own buffers; a grouped call
// gives every group the words of the one-level call, with one launch per stage for all
// groups (CKKS, N = 2^15 and 2^16 for the key switch; other shapes run group by group).
struct TensorGroup {
    const uint64_t *a, *b;  // [count][2][L][N]
    uint64_t *out;          // [count][3][L][N]
    uint32_t count, L;
};
void ckks_tensor_groups(const Context &c, const TensorGroup *groups, uint32_t n, hipStream_t st);

struct RelinGroup {
    uint64_t *ct;  // [count][3][L][N], relinearized in place (polys 0, 1)
    uint32_t count, L;
};
size_t relinearize_groups_workspace(const Context &c, const RelinGroup *groups, uint32_t n);
void relinearize_groups(const Context &c, const RelinGroup *groups, uint32_t n, const uint64_t *key, void *ws,
                        size_t ws_bytes, hipStream_t st);

struct RescaleGroup {
    const uint64_t *in;  // [count][in_size][L][N]: polys 0 and 1 are rescaled
    uint64_t *out;       // [count][2][L-1][N]
    uint32_t count, L;
};
size_t rescale_groups_workspace(const Context &c, const RescaleGroup *groups, uint32_t n);
void rescale_groups(const Context &c, const RescaleGroup *groups, uint32_t n, uint32_t in_size, void *ws,
                    size_t ws_bytes, hipStream_t st);

}  // namespace shc
