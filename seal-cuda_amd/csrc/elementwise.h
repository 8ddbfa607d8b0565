// elementwise.h -- the evaluator's coefficient-wise operations: add, sub, negate,
// plaintext add / multiply, over batches of RNS polynomials.
#pragma once
#include "context.h"

namespace shc {

// a, b, out: [polys][L][N], limb l reduced by the context's prime l; out may alias.
void poly_add(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys, uint32_t L,
              hipStream_t st);
void poly_sub(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t polys, uint32_t L,
              hipStream_t st);
void poly_negate(const Context &c, const uint64_t *a, uint64_t *out, uint32_t polys, uint32_t L, hipStream_t st);
// out[p] = a[p] * plain (dyadic, both NTT form); plain: [L][N].
void multiply_plain_ntt(const Context &c, const uint64_t *a, const uint64_t *plain, uint64_t *out, uint32_t polys,
                        uint32_t L, hipStream_t st);
// Ciphertexts [count][size][L][N]: out = ct with poly 0 +- plain ([L][N], same form).
void add_plain_poly0(const Context &c, const uint64_t *ct, const uint64_t *plain, uint64_t *out, uint32_t count,
                     uint32_t size, uint32_t L, bool subtract, hipStream_t st);
// out[p] = sum_j plains[j] * cts[j][p] (dyadic, all NTT form): cts [terms][polys][L][N],
// plains [terms][L][N], out [polys][L][N] (not overlapping cts).
void multiply_plain_accumulate(const Context &c, const uint64_t *cts, const uint64_t *plains, uint32_t terms,
                               uint32_t polys, uint32_t L, uint64_t *out, hipStream_t st);
// `outputs` such sums over the same cts: plains [outputs][terms][L][N], out
// [outputs][polys][L][N]; each ct word is read once for up to 8 outputs.
void multiply_plain_accumulate_many(const Context &c, const uint64_t *cts, const uint64_t *plains, uint32_t terms,
                                    uint32_t outputs, uint32_t polys, uint32_t L, uint64_t *out, hipStream_t st);
// BFV: out = ct with c0 +- Delta m (SEAL's scaling variant), plain m: [N] mod t.
void bfv_add_plain(const Context &c, const uint64_t *ct, const uint64_t *plain, uint64_t *out, uint32_t count,
                   uint32_t size, uint32_t L, bool subtract, hipStream_t st);
// BFV plain multiply of coefficient-form polynomials [polys][L][N] by m ([N] mod t),
// lifted to [L][N] as m - t for m >= ceil(t/2), through the NTT.
size_t bfv_multiply_plain_workspace(const Context &c, uint32_t L);
// The lifted plaintext in NTT form at L limbs (transformToNttInplace(PlaintextCuda&)).
void bfv_plain_to_ntt(const Context &c, const uint64_t *plain, uint32_t L, uint64_t *out, hipStream_t st);
void bfv_multiply_plain(const Context &c, const uint64_t *a, const uint64_t *plain, uint64_t *out, uint32_t polys,
                        uint32_t L, void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace shc
