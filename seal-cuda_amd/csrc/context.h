// context.h -- device-resident tables for one parameter set (the analogue of the
// reference's SEALContextCuda / ContextDataCuda, src/context_cuda.cuh:11-238).
//
// Primes are addressed by device arrays of DevPrime descriptors:
//   primes()              the coefficient modulus at key level (last = special prime)
//   keyswitch_primes(L)   {q_0 .. q_{L-1}, q_special}: the targets of a key switch at level L
//   behz(L)               BFV only: BEHZ tables of level L, including the Bsk primes
//   plain_prime()         BFV only: the plain modulus t when it supports batching
#pragma once
#include "behz.h"
#include "common.h"
#include "hostmath.h"
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

namespace shc {

enum Scheme { SCHEME_NONE = 0, SCHEME_BFV = 1, SCHEME_CKKS = 2, SCHEME_BGV = 3 };

struct CkksTables;

struct Context {
    Scheme scheme = SCHEME_NONE;
    int logn = 0;
    uint64_t n = 0;
    std::vector<uint64_t> moduli;      // key level, special prime last
    uint64_t plain = 0;

    std::vector<DevPrime> host_primes;
    std::vector<DevBuf> table_bufs;
    DevBuf d_primes;                   // DevPrime[K]
    DevBuf d_keyswitch_primes;         // DevPrime[(K-1) * K]: row L-1 = level L's targets
    DevBuf d_ks_consts;                // DivRoundConst[K]: division by the special prime, entry i < K-1
    DevBuf d_rescale_consts;           // DivRoundConst[(K+1) * K]: row L = division by q_{L-1}, entries i < L-1

    std::map<uint64_t, DevPrime> aux;  // NTT-ready auxiliary primes (Bsk, t) by value
    std::vector<std::unique_ptr<BehzLevel>> behz_levels;  // index L (BFV)
    DevBuf d_plain_prime;              // DevPrime[1] for t (BFV, batching)
    std::shared_ptr<CkksTables> ckks;  // CKKS encoder tables

    int device = 0;

    // Sampler: with sampler_keyed, word i of the stream of seed s is taken from the
    // ChaCha20 keystream under this 256-bit key (nonce s, block counter i / 8);
    // otherwise from splitmix64 (deterministic streams for explicitly seeded calls).
    bool sampler_keyed = false;
    uint32_t sampler_key[8] = {};

    // (P mod q_j, its Shoup quotient) for j < K-1, P the special prime: the factors a
    // key-switching key folds the new key in with (key generation).
    DevBuf d_kswitch_factors;

    // Instrumentation: when armed, the next key switch records these events around its
    // ModUp rows-pass launch (the whole ModUp forward NTT for N < 2^13; the first
    // slice's, for a batch run in slices) and disarms.  Key switches that run while it
    // is not armed only read probe_armed; arming and taking the probe hold probe_mu, so
    // concurrent key switches on one context (different streams) take it at most once.
    mutable std::mutex probe_mu;
    mutable KernelProbe probe{nullptr, nullptr};
    mutable std::atomic<bool> probe_armed{false};

    void arm_probe(KernelProbe p) const {
        std::lock_guard<std::mutex> g(probe_mu);
        probe = p;
        probe_armed.store(true, std::memory_order_release);
    }
    // The armed probe, disarming it; false when none is armed.
    bool take_probe(KernelProbe &out) const {
        if (!probe_armed.load(std::memory_order_acquire)) return false;
        std::lock_guard<std::mutex> g(probe_mu);
        if (!probe_armed.load(std::memory_order_relaxed)) return false;
        out = probe;
        probe_armed.store(false, std::memory_order_relaxed);
        return true;
    }

    // Hoisted keys prepared through this context (device address -> level, Galois
    // element), so that rotation sums check a prepared key on the host, without reading
    // its header words back from the device.
    mutable std::mutex hoisted_mu;
    mutable std::map<const void *, std::pair<uint32_t, uint32_t>> hoisted_keys;

    // Key-switch launch shapes by (level, ciphertexts per launch), written by the
    // autotuner (or set by the caller); shapes not in the table take the built-in choice.
    mutable std::mutex tune_mu;
    mutable std::map<std::pair<uint32_t, uint32_t>, KsShape> ks_shapes;
    KsShape ks_shape(uint32_t L, uint32_t count) const {
        std::lock_guard<std::mutex> g(tune_mu);
        auto it = ks_shapes.find({L, count});
        return it == ks_shapes.end() ? KsShape{} : it->second;
    }

    size_t key_count() const { return moduli.size(); }
    const DevPrime *primes() const { return d_primes.as<DevPrime>(); }
    const DevPrime *keyswitch_primes(size_t level) const {
        return d_keyswitch_primes.as<DevPrime>() + (level - 1) * moduli.size();
    }
    const DivRoundConst *ks_consts() const { return d_ks_consts.as<DivRoundConst>(); }
    const DivRoundConst *rescale_consts(size_t level) const {
        return d_rescale_consts.as<DivRoundConst>() + level * moduli.size();
    }
    const BehzLevel &behz(size_t level) const;
    const DevPrime *plain_prime() const { return d_plain_prime.as<DevPrime>(); }

    // Host-side descriptor (NTT tables uploaded) of an auxiliary prime.
    DevPrime aux_prime(uint64_t q);
};

std::unique_ptr<Context> make_context(Scheme scheme, int logn, const std::vector<uint64_t> &moduli, uint64_t plain);

// Uploads NTT tables of q; throws invalid_argument when q has no 2N-th root of unity.
DevPrime upload_prime(Context &c, uint64_t q);

}  // namespace shc
