// scratch.h -- the fixed device scratch array behind shc_relinearize_noalloc.
#pragma once
#include "context.h"

namespace shc {

// Size of the per-device scratch array (128 MiB): the key-switch workspace of, e.g.,
// 40 ciphertexts at N=8192 with 4 data primes, or one at N=2^15 with 15.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;

// The current device's scratch address (looked up once per device and kept).
void *noalloc_scratch_setup();

// Relinearize with the key-switch scratch in the fixed array: no allocation; fails
// with ERR_OUT_OF_RANGE when the key switch needs more than NOALLOC_SCRATCH_BYTES.
void relinearize_noalloc(const Context &c, uint64_t *ct, const uint64_t *key, uint32_t count, uint32_t L,
                         hipStream_t st);

}  // namespace shc
