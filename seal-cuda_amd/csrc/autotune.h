// autotune.h -- launch-shape autotuner for the key switch (autotune.cpp).
#pragma once
#include "context.h"
#include <string>
#include <vector>

namespace shc {

struct TuneResult {
    KsShape best;
    float best_ms = 0, builtin_ms = 0;  // per relinearization of the batch
    uint32_t tried = 0;
};

// The shapes the tuner tries at level L: the built-in choice, then every
// (targets per columns block, ciphertexts per rows block) pair.
std::vector<KsShape> keyswitch_candidates(const Context &c, uint32_t L);

// Times `reps` relinearizations of ct [count][3][L][N] (used as scratch: its words are
// overwritten) per candidate shape and keeps the fastest in c's table for (L, count).
TuneResult autotune_relinearize(const Context &c, uint64_t *ct, const uint64_t *key, uint32_t count, uint32_t L,
                                void *ws, size_t ws_bytes, hipStream_t st, int reps);

// The table: set (shape == nullptr or all zero: back to the built-in choice), and the
// whole table as text lines "L count cols_targets rows_group" (export / import).
void keyswitch_shape_set(const Context &c, uint32_t L, uint32_t count, const KsShape *shape);
std::string keyswitch_shapes_export(const Context &c);
void keyswitch_shapes_import(const Context &c, const std::string &text);

}  // namespace shc
