// ntt.h -- host launchers for the batched negacyclic NTT / inverse NTT.
#pragma once
#include "common.h"

namespace shc {

// Output range of a transform.
enum NttOut { NTT_FULL = 0,   // fully reduced, [0, q)
              NTT_LAZY = 1 }; // forward: [0, 4q); inverse: [0, 2q)

// In-place transform of `count` limbs stored back to back, each of length 2^logn.
// Limb k (0 <= k < count) is reduced modulo primes[k % limbs_per_poly], i.e. the data is
// a [poly][limb][coeff] array with `limbs_per_poly` limbs per polynomial.
//   forward: natural order in (values < 4q), bit-reversed order out
//   inverse: bit-reversed order in (values < 2q), natural order out, times N^-1
void ntt_forward(uint64_t *data, uint32_t count, uint32_t limbs_per_poly, const DevPrime *primes, int logn,
                 NttOut out, hipStream_t stream);
void ntt_inverse(uint64_t *data, uint32_t count, uint32_t limbs_per_poly, const DevPrime *primes, int logn,
                 NttOut out, hipStream_t stream);

// Key-switching ModUp fused with the forward NTT: T[ct][j][i] = NTT_i(d[ct][j] mod t_i)
// for digits j < L and targets i <= L (targets[i]), d[ct] at d + ct * d_ct_stride with
// L limbs.  Output fully reduced.  Returns false (and does nothing) for N < 2^13, where
// the caller gathers first and runs the plain transform.
bool ntt_forward_modup(const uint64_t *d, size_t d_ct_stride, uint64_t *T, uint32_t count, uint32_t L,
                       const DevPrime *targets, int logn, hipStream_t stream);

}  // namespace shc
