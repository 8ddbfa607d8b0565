// context.h -- device-resident tables for one parameter set (the analogue of the
// reference's SEALContextCuda / ContextDataCuda, src/context_cuda.cuh:11-238).
//
// Primes are addressed by one flat device array of DevPrime descriptors:
//   [0, K)            the coefficient modulus at key level (last = special prime)
//   [K, K + |Bsk|)    BFV only: the Bsk base used by BEHZ multiplication
//   plain_index       BFV only: the plain modulus t when it supports batching
// Per data level L (1 <= L <= K-1) a second array `keyswitch_primes(L)` holds
// {q_0 .. q_{L-1}, q_special}, the target primes of a key switch at that level.
#pragma once
#include "common.h"
#include "hostmath.h"
#include <memory>
#include <vector>

namespace shc {

enum Scheme { SCHEME_NONE = 0, SCHEME_BFV = 1, SCHEME_CKKS = 2, SCHEME_BGV = 3 };

// RAII device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b);
    ~DevBuf();
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            this->~DevBuf();
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct Context {
    Scheme scheme = SCHEME_NONE;
    int logn = 0;
    uint64_t n = 0;
    std::vector<uint64_t> moduli;      // key level, special prime last
    uint64_t plain = 0;

    std::vector<hm::NttTables> host_tables;  // per entry of the flat prime array
    std::vector<DevPrime> host_primes;
    std::vector<DevBuf> table_bufs;
    DevBuf d_primes;                   // DevPrime[total]
    DevBuf d_keyswitch_primes;         // DevPrime[(K-1) * K]: row L-1 = level L's targets
    DevBuf d_ks_consts;                // DivRoundConst[K]: division by the special prime, entry i < K-1
    DevBuf d_rescale_consts;           // DivRoundConst[(K+1) * K]: row L = division by q_{L-1}, entries i < L-1

    int device = 0;

    const DivRoundConst *ks_consts() const { return d_ks_consts.as<DivRoundConst>(); }
    const DivRoundConst *rescale_consts(size_t level) const {
        return d_rescale_consts.as<DivRoundConst>() + level * moduli.size();
    }

    size_t key_count() const { return moduli.size(); }
    const DevPrime *primes() const { return d_primes.as<DevPrime>(); }
    const DevPrime *keyswitch_primes(size_t level) const {
        return d_keyswitch_primes.as<DevPrime>() + (level - 1) * moduli.size();
    }
};

std::unique_ptr<Context> make_context(Scheme scheme, int logn, const std::vector<uint64_t> &moduli, uint64_t plain);

}  // namespace shc
