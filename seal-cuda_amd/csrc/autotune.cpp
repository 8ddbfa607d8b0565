// autotune.cpp -- launch-shape autotuner for the key switch's ModUp kernels.
//
// The reference launches every kernel with one fixed shape, <<<ceil(n/256), 256>>>
// (src/kernelutils.cu:3).  Here the two kernels that dominate a key switch have shape
// choices whose best value depends on the level, the batch and the degree:
//   - targets per columns-pass block (a block keeps a 16-column tile of a digit in
//     registers and transforms it for that many target primes in turn: fewer targets
//     per block give more, shorter blocks, which small batches need to fill the GPU);
//   - ciphertexts per fused rows-pass block (4, 2 or 1: more ciphertexts share each key
//     word loaded, at the price of registers and occupancy).
// The tuner times a relinearization with every candidate shape on the caller's buffers
// and keeps the fastest in the context's table, which later key switches of that
// (level, batch) read.  The candidates include the built-in choice, and another shape
// replaces it only when it is more than 1% faster on the tuning run.
#include "autotune.h"
#include "rns_ops.h"
#include <algorithm>
#include <cstdio>
#include <sstream>

namespace shc {

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        SHC_HIP(hipEventCreate(&a));
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            fail(ERR_HIP, "hipEventCreate failed");
        }
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

void set_shape(const Context &c, uint32_t L, uint32_t count, const KsShape *s) {
    std::lock_guard<std::mutex> g(c.tune_mu);
    if (s && (s->cols_targets || s->rows_group)) c.ks_shapes[{L, count}] = *s;
    else c.ks_shapes.erase({L, count});
}

}  // namespace

std::vector<KsShape> keyswitch_candidates(const Context &c, uint32_t L) {
    std::vector<uint32_t> targets;
    for (uint32_t chunks : {1u, 2u, 3u, 4u, 6u, 8u, 12u, 16u, L + 1}) {
        if (chunks > L + 1) continue;
        const uint32_t tc = (L + chunks) / chunks;  // ceil((L + 1) / chunks)
        if (std::find(targets.begin(), targets.end(), tc) == targets.end()) targets.push_back(tc);
    }
    const bool fused = c.logn == 15 || c.logn == 16;
    std::vector<KsShape> out{KsShape{}};  // the built-in choice first
    for (uint32_t g : fused ? std::vector<uint32_t>{4, 2, 1} : std::vector<uint32_t>{0})
        for (uint32_t tc : targets) out.push_back(KsShape{tc, g});
    return out;
}

TuneResult autotune_relinearize(const Context &c, uint64_t *ct, const uint64_t *key, uint32_t count, uint32_t L,
                                void *ws, size_t ws_bytes, hipStream_t st, int reps) {
    if (count == 0) fail(ERR_INVALID_ARGUMENT, "nothing to tune");
    if (reps < 1) reps = 1;
    const size_t poly = (size_t)L * c.n;
    const KsShape before = c.ks_shape(L, count);
    TuneResult r;
    EventPair ev;
    auto time_shape = [&](const KsShape &s) {
        set_shape(c, L, count, &s);
        switch_key(c, ct + 2 * poly, 3 * poly, ct, 3 * poly, key, count, L, ws, ws_bytes, st);  // warm
        SHC_HIP(hipEventRecord(ev.a, st));
        for (int i = 0; i < reps; i++)
            switch_key(c, ct + 2 * poly, 3 * poly, ct, 3 * poly, key, count, L, ws, ws_bytes, st);
        SHC_HIP(hipEventRecord(ev.b, st));
        SHC_HIP(hipEventSynchronize(ev.b));
        float ms = 0;
        SHC_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        return ms / reps;
    };
    try {
        for (const KsShape &s : keyswitch_candidates(c, L)) {
            const float ms = time_shape(s);
            r.tried++;
            if (!s.cols_targets && !s.rows_group) r.builtin_ms = ms;
            if (r.tried == 1 || ms < r.best_ms) {
                r.best_ms = ms;
                r.best = s;
            }
        }
    } catch (...) {
        set_shape(c, L, count, &before);
        throw;
    }
    // a shape has to beat the built-in one by more than run-to-run noise (1%) to be kept
    if (r.best_ms > 0.99f * r.builtin_ms) {
        r.best = KsShape{};
        r.best_ms = r.builtin_ms;
    }
    set_shape(c, L, count, &r.best);
    return r;
}

void keyswitch_shape_set(const Context &c, uint32_t L, uint32_t count, const KsShape *shape) {
    if (shape && shape->rows_group && shape->rows_group != 1 && shape->rows_group != 2 && shape->rows_group != 4)
        fail(ERR_INVALID_ARGUMENT, "rows_group must be 0, 1, 2 or 4");
    if (shape && shape->cols_targets > L + 1) fail(ERR_INVALID_ARGUMENT, "cols_targets must be at most L + 1");
    set_shape(c, L, count, shape);
}

std::string keyswitch_shapes_export(const Context &c) {
    std::lock_guard<std::mutex> g(c.tune_mu);
    std::ostringstream os;
    for (const auto &e : c.ks_shapes)
        os << e.first.first << ' ' << e.first.second << ' ' << e.second.cols_targets << ' ' << e.second.rows_group
           << '\n';
    return os.str();
}

void keyswitch_shapes_import(const Context &c, const std::string &text) {
    std::istringstream is(text);
    std::vector<std::pair<std::pair<uint32_t, uint32_t>, KsShape>> rows;
    std::string line;
    while (std::getline(is, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ls(line);
        long long L, count, tc, g;
        if (!(ls >> L >> count >> tc >> g) || L < 1 || count < 1 || tc < 0 || g < 0 || L > 0xffff ||
            count > 0xffffffffll || tc > L + 1 || (g != 0 && g != 1 && g != 2 && g != 4))
            fail(ERR_INVALID_ARGUMENT, "bad launch-shape line: " + line);
        rows.push_back({{(uint32_t)L, (uint32_t)count}, KsShape{(uint32_t)tc, (uint32_t)g}});
    }
    for (const auto &r : rows) set_shape(c, r.first.first, r.first.second, &r.second);
}

}  // namespace shc
