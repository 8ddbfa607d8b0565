// common.h -- shared types of the HIP library: device-resident prime descriptors
// and error plumbing between kernels, the context and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdexcept>
#include <string>

namespace shc {

// One RNS prime as the kernels see it.  `fwd[i]` = (root_powers[i], Shoup quotient),
// bit-reversed powers of the minimal primitive 2N-th root; `inv[i]` the "scrambled"
// inverse powers (layouts of NTTTables, src/utils/ntt.cpp:36-54).
struct alignas(64) DevPrime {
    uint64_t q;
    uint64_t two_q;
    uint64_t ratio_lo, ratio_hi;     // floor(2^128 / q)
    uint64_t inv_n, inv_n_shoup;     // N^-1 mod q as a Shoup operand
    const ulonglong2 *fwd;           // [N]  (w, floor(w 2^64 / q))
    const ulonglong2 *inv;           // [N]
    // 2^64 - 4q and 2^64 - 8q: the NTT's conditional subtractions as one 64-bit add
    // (a constant the compiler cannot see as a negation stays an add)
    uint64_t neg_four_q, neg_eight_q;
};

// Constants of a divide-and-round by a prime p into prime q_i (p = the special prime
// for key switching, q_last for rescale / modulus switching).
struct alignas(32) DivRoundConst {
    uint64_t inv, inv_shoup;   // p^-1 mod q_i as a Shoup operand
    uint64_t half_mod;         // floor(p/2) mod q_i
    uint64_t p_mod;            // p mod q_i
};

// Launch shape of a key switch's two ModUp kernels (0 = the built-in choice):
//   cols_targets  target primes one columns-pass block transforms a digit tile for
//                 (1 .. L+1; the built-in choice aims at ~16k blocks);
//   rows_group    ciphertexts one fused rows-pass block takes: 4 (then 2 and 1 for the
//                 rest of the batch), 2 (then 1), or 1 (two rows of one ciphertext).
struct KsShape {
    uint32_t cols_targets = 0;
    uint32_t rows_group = 0;
};

// Instrumentation: a pair of caller-created events recorded around one kernel launch.
struct KernelProbe {
    hipEvent_t begin, end;
};

// Row idx >> logn of a flat index over rows of 2^logn words.  In the grid-stride loops
// here a wave's 64 lanes hold 64 consecutive indices starting at a multiple of 64, so
// for 2^logn >= 64 every lane is in the same row: the row is made wave-uniform and the
// divisions that split it run on the scalar unit.  Rows are below 2^32.
__device__ __forceinline__ uint32_t uniform_row(size_t idx, int logn) {
    const uint32_t r = (uint32_t)(idx >> logn);
    return logn >= 6 ? __builtin_amdgcn_readfirstlane(r) : r;
}

// Error raised by host code; carries the exception family of the reference
// (std::invalid_argument / logic_error / out_of_range / runtime_error).
enum ErrKind { ERR_OK = 0, ERR_INVALID_ARGUMENT = -1, ERR_LOGIC = -2, ERR_OUT_OF_RANGE = -3, ERR_RUNTIME = -4,
               ERR_HIP = -5 };

struct Error : std::runtime_error {
    ErrKind kind;
    Error(ErrKind k, const std::string &m) : std::runtime_error(m), kind(k) {}
};

[[noreturn]] inline void fail(ErrKind k, const std::string &m) { throw Error(k, m); }

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define SHC_HIP(call) ::shc::hip_check((call), #call)

inline void launch_check(const char *what) { hip_check(hipGetLastError(), what); }

// RAII device buffer (hipMalloc / hipFree).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) : bytes(b) {
        if (b) SHC_HIP(hipMalloc(&ptr, b));
    }
    ~DevBuf() {
        if (ptr) (void)hipFree(ptr);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            if (ptr) (void)hipFree(ptr);
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

}  // namespace shc
