// troyhip.hpp -- the troy_cuda C++ class API over the MI355X C ABI (sealhip.h).
//
// Header only.  A program written against the reference's troyn:: aliases
// (src/troy_cuda.cuh:20-42) switches by using namespace troyhip instead: the same class
// names, method names, argument meanings and exception types.  Device buffers are
// owned by RAII arrays (the reference's DeviceArray, src/utils/devicearray.cuh:106-200)
// whose blocks come from a device memory pool (the reference's MemoryPoolCuda,
// src/utils/memorypool_cuda.cuh:8-73), so steady-state operations do not call hipMalloc
// or hipFree.  Every operation runs on the default stream, as the reference's do, and
// returns when its work is queued (host-visible results -- decode, toHost --
// synchronise).
//
// A ParmsID is the reference's: the 4-word BLAKE2b hash of a level's parameters
// (src/encryptionparams.h:36).  Keys and encryptions are drawn from ChaCha20 under a
// 256-bit key that every SEALContext takes from std::random_device.  The CPU-side
// containers the reference converts to and from (cpu(), toHost(), the conversion
// constructors) are troyhip::host::Plaintext and troyhip::host::Ciphertext.
//
// Link with libsealhip.so (and the HIP runtime).
#pragma once
#include "sealhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <istream>
#include <map>
#include <memory>
#include <mutex>
#include <ostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace troyhip {

// ---------------------------------------------------------------------------- errors

// Maps a C ABI status to the reference's exception family (std::invalid_argument,
// std::logic_error, std::out_of_range, std::runtime_error).
inline void check(int status) {
    if (status == SHC_OK) return;
    const std::string msg = shc_last_error();
    switch (status) {
        case SHC_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case SHC_LOGIC_ERROR: throw std::logic_error(msg);
        case SHC_OUT_OF_RANGE: throw std::out_of_range(msg);
        default: throw std::runtime_error(msg);
    }
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace detail {
// BLAKE2b (RFC 7693), unkeyed, outlen <= 64 bytes: the hash the reference names
// parameter sets with (src/utils/hash.h:25-32 over src/encryptionparams.cpp:118-156).
inline void blake2b(const uint8_t *in, size_t inlen, uint8_t *out, size_t outlen) {
    static const uint64_t IV[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                                   0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                                   0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};
    static const uint8_t SIGMA[12][16] = {
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
        {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
        {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
        {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
        {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0},
        {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3}};
    auto rotr = [](uint64_t x, int r) { return (x >> r) | (x << (64 - r)); };
    uint64_t h[8];
    for (int i = 0; i < 8; i++) h[i] = IV[i];
    h[0] ^= 0x01010000ull ^ (uint64_t)outlen;
    uint64_t t = 0;
    auto compress = [&](const uint8_t *block, bool last) {
        uint64_t m[16], v[16];
        for (int i = 0; i < 16; i++) {
            m[i] = 0;
            for (int b = 0; b < 8; b++) m[i] |= (uint64_t)block[8 * i + b] << (8 * b);
        }
        for (int i = 0; i < 8; i++) v[i] = h[i], v[i + 8] = IV[i];
        v[12] ^= t;
        if (last) v[14] = ~v[14];
        auto G = [&](int a, int b, int c, int d, uint64_t x, uint64_t y) {
            v[a] = v[a] + v[b] + x;
            v[d] = rotr(v[d] ^ v[a], 32);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 24);
            v[a] = v[a] + v[b] + y;
            v[d] = rotr(v[d] ^ v[a], 16);
            v[c] = v[c] + v[d];
            v[b] = rotr(v[b] ^ v[c], 63);
        };
        for (int r = 0; r < 12; r++) {
            const uint8_t *s = SIGMA[r];
            G(0, 4, 8, 12, m[s[0]], m[s[1]]);
            G(1, 5, 9, 13, m[s[2]], m[s[3]]);
            G(2, 6, 10, 14, m[s[4]], m[s[5]]);
            G(3, 7, 11, 15, m[s[6]], m[s[7]]);
            G(0, 5, 10, 15, m[s[8]], m[s[9]]);
            G(1, 6, 11, 12, m[s[10]], m[s[11]]);
            G(2, 7, 8, 13, m[s[12]], m[s[13]]);
            G(3, 4, 9, 14, m[s[14]], m[s[15]]);
        }
        for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
    };
    uint8_t block[128];
    size_t off = 0;
    while (inlen - off > 128) {
        t += 128;
        compress(in + off, false);
        off += 128;
    }
    const size_t rest = inlen - off;
    std::memset(block, 0, 128);
    if (rest) std::memcpy(block, in + off, rest);
    t += rest;
    compress(block, true);
    for (size_t i = 0; i < outlen; i++) out[i] = (uint8_t)(h[i / 8] >> (8 * (i % 8)));
}


// parms_id of a parameter set: BLAKE2b-256 of the words (scheme, N, q_0 .. q_{L-1},
// plain modulus), little endian (src/encryptionparams.cpp:118-156).
inline std::array<uint64_t, 4> parms_hash(uint64_t scheme, uint64_t n, const std::vector<uint64_t> &q, uint64_t t) {
    std::vector<uint64_t> words{scheme, n};
    words.insert(words.end(), q.begin(), q.end());
    words.push_back(t);
    std::vector<uint8_t> bytes(words.size() * 8);
    for (size_t i = 0; i < words.size(); i++)
        for (int b = 0; b < 8; b++) bytes[8 * i + b] = (uint8_t)(words[i] >> (8 * b));
    uint8_t out[32];
    blake2b(bytes.data(), bytes.size(), out, 32);
    std::array<uint64_t, 4> h{};
    for (int i = 0; i < 32; i++) h[i / 8] |= (uint64_t)out[i] << (8 * (i % 8));
    return h;
}
}  // namespace detail

// KernelProvider::initialize (src/kernelprovider.cuh:18-21): must run before any object
// that touches the device is created.
class KernelProvider {
public:
    static void initialize(int device = 0) {
        check(shc_init(device));
        flag() = true;
    }
    static void checkInitialized() {
        if (!flag()) throw std::logic_error("KernelProvider not initialized");
    }

private:
    static bool &flag() {
        static bool f = false;
        return f;
    }
};

// ----------------------------------------------------------------------- memory pool

// MemoryPoolCuda (src/utils/memorypool_cuda.cuh:8-73): device blocks that arrays free
// are kept per device and size and handed to the next request for at least that size
// and at most twice it, so repeated operations stop calling hipMalloc / hipFree.  A
// freed block is reused by later work on the default stream, which the stream orders
// after every earlier use of the block.  When free device memory runs below a reserve,
// or hipMalloc fails, the cached blocks are released and the allocation retried.
class MemoryPool {
public:
    // A block of at least `bytes` (granted: its size); bytes > 0.
    static void *get(size_t bytes, size_t &granted) {
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        const size_t want = round(bytes);
        {
            State &s = state();
            std::lock_guard<std::mutex> g(s.mu);
            auto &m = s.free[dev];
            auto it = m.lower_bound(want);
            if (it != m.end() && it->first <= 2 * want) {
                void *p = it->second;
                granted = it->first;
                s.cached -= it->first;
                m.erase(it);
                return p;
            }
        }
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b < want + RESERVE) release();
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            release();
            hip_check(hipMalloc(&p, want), "hipMalloc");
        }
        {
            State &s = state();
            std::lock_guard<std::mutex> g(s.mu);
            s.mallocs++;
        }
        granted = want;
        return p;
    }
    static void put(void *p, size_t granted) {
        if (!p) return;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        State &s = state();
        std::lock_guard<std::mutex> g(s.mu);
        s.free[dev].emplace(granted, p);
        s.cached += granted;
    }
    // Frees every cached block (all devices); returns the bytes freed.
    static size_t release() {
        State &s = state();
        std::lock_guard<std::mutex> g(s.mu);
        int cur = 0;
        (void)hipGetDevice(&cur);
        size_t freed = 0;
        for (auto &dm : s.free) {
            if (dm.second.empty()) continue;
            (void)hipSetDevice(dm.first);
            (void)hipDeviceSynchronize();  // blocks may still be read by queued work
            for (auto &e : dm.second) {
                (void)hipFree(e.second);
                freed += e.first;
            }
            dm.second.clear();
        }
        (void)hipSetDevice(cur);
        s.cached = 0;
        return freed;
    }
    static size_t cachedBytes() {
        State &s = state();
        std::lock_guard<std::mutex> g(s.mu);
        return s.cached;
    }
    // hipMalloc calls the pool has made so far.
    static size_t deviceAllocations() {
        State &s = state();
        std::lock_guard<std::mutex> g(s.mu);
        return s.mallocs;
    }

private:
    static constexpr size_t RESERVE = size_t(32) << 20;
    static size_t round(size_t b) { return (b + 4095) & ~size_t(4095); }
    struct State {
        std::mutex mu;
        std::map<int, std::multimap<size_t, void *>> free;
        size_t cached = 0;
        size_t mallocs = 0;
    };
    // Never destroyed: arrays with static storage may return blocks during exit.
    static State &state() {
        static State *s = new State;
        return *s;
    }
};

// ---------------------------------------------------------------------- device array

template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray &o) { *this = o; }
    DeviceArray(DeviceArray &&o) noexcept : ptr_(o.ptr_), size_(o.size_), cap_(o.cap_) {
        o.ptr_ = nullptr;
        o.size_ = o.cap_ = 0;
    }
    DeviceArray &operator=(const DeviceArray &o) {
        if (this != &o) {
            resize(o.size_);
            if (size_) hip_check(hipMemcpy(ptr_, o.ptr_, size_ * sizeof(T), hipMemcpyDeviceToDevice), "copy");
        }
        return *this;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) {
            release();
            ptr_ = o.ptr_;
            size_ = o.size_;
            cap_ = o.cap_;
            o.ptr_ = nullptr;
            o.size_ = o.cap_ = 0;
        }
        return *this;
    }
    ~DeviceArray() { release(); }

    // The block is kept while n fits its capacity (as DeviceDynamicArray::resize,
    // src/utils/devicearray.cuh:282-285, so a reserve lasts); a larger n takes a new
    // block from the pool and, unlike the reference's, does not copy the contents.
    void resize(size_t n) {
        if (n == size_) return;
        const size_t bytes = n * sizeof(T);
        if (ptr_ && bytes <= cap_) {
            size_ = n;
            return;
        }
        release();
        if (n) ptr_ = static_cast<T *>(MemoryPool::get(bytes, cap_));
        size_ = n;
    }
    // Resizes keeping the first min(size, n) elements and zeroing any new ones (the
    // reference's DeviceDynamicArray::resize, src/utils/devicearray.cuh:208-216, 289-292).
    void resizeKeep(size_t n) {
        if (n == size_) return;
        if (ptr_ && n * sizeof(T) <= cap_) {
            if (n > size_) hip_check(hipMemset(ptr_ + size_, 0, (n - size_) * sizeof(T)), "memset");
            size_ = n;
            return;
        }
        DeviceArray t(n);
        const size_t keep = std::min(n, size_);
        if (keep) hip_check(hipMemcpy(t.ptr_, ptr_, keep * sizeof(T), hipMemcpyDeviceToDevice), "copy");
        if (n > keep) hip_check(hipMemset(t.ptr_ + keep, 0, (n - keep) * sizeof(T)), "memset");
        *this = std::move(t);
    }
    // Grows to at least n elements, keeping nothing.
    void ensure(size_t n) {
        if (n > size_) resize(n);
    }
    // Capacity for n elements, keeping the contents (DeviceDynamicArray::reserve,
    // src/utils/devicearray.cuh); shrinkToFit moves them to a block of their own size.
    void reserve(size_t n) {
        if (n * sizeof(T) <= cap_) return;
        moveTo(n);
    }
    void shrinkToFit() {
        if (cap_ > size_ * sizeof(T)) moveTo(size_);
    }
    void fromHost(const T *src, size_t n) {
        resize(n);
        if (n) hip_check(hipMemcpy(ptr_, src, n * sizeof(T), hipMemcpyHostToDevice), "copy to device");
    }
    std::vector<T> toHost() const {
        std::vector<T> v(size_);
        if (size_) hip_check(hipMemcpy(v.data(), ptr_, size_ * sizeof(T), hipMemcpyDeviceToHost), "copy to host");
        return v;
    }
    void setZero() {
        if (size_) hip_check(hipMemset(ptr_, 0, size_ * sizeof(T)), "memset");
    }
    void release() {
        if (ptr_) MemoryPool::put(ptr_, cap_);
        ptr_ = nullptr;
        size_ = cap_ = 0;
    }
    T *get() { return ptr_; }
    const T *get() const { return ptr_; }
    size_t size() const { return size_; }
    size_t capacity() const { return cap_ / sizeof(T); }

private:
    // A block for `elements` (none when zero) holding the first min(size, elements) ones.
    void moveTo(size_t elements) {
        const size_t keep = std::min(size_, elements);
        T *p = nullptr;
        size_t cap = 0;
        if (elements) p = static_cast<T *>(MemoryPool::get(elements * sizeof(T), cap));
        if (keep) hip_check(hipMemcpy(p, ptr_, keep * sizeof(T), hipMemcpyDeviceToDevice), "copy");
        if (ptr_) MemoryPool::put(ptr_, cap_);
        ptr_ = p;
        cap_ = cap;
        size_ = keep;
    }
    T *ptr_ = nullptr;
    size_t size_ = 0;
    size_t cap_ = 0;  // bytes of the pool block
};

// ------------------------------------------------------------------ parameter types

enum class SchemeType : uint8_t { none = 0, bfv = 1, ckks = 2, bgv = 3 };
enum class SecurityLevel : int { none = 0, tc128 = 128, tc192 = 192, tc256 = 256 };

// Modulus (src/modulus.h:16-24): the value and its bit count.
class Modulus {
public:
    Modulus(uint64_t value = 0) : value_(value) {}
    uint64_t value() const { return value_; }
    int bitCount() const {
        int b = 0;
        for (uint64_t v = value_; v; v >>= 1) b++;
        return b;
    }
    bool isZero() const { return value_ == 0; }
    bool operator==(const Modulus &o) const { return value_ == o.value_; }
    bool operator!=(const Modulus &o) const { return value_ != o.value_; }

private:
    uint64_t value_;
};

// CoeffModulus (src/modulus.h:400-480, src/modulus.cpp:80-165).
struct CoeffModulus {
    // The HE-standard maximum total bit count for a ternary secret
    // (src/utils/hestdparams.h); 0 for unsupported degrees.
    static int MaxBitCount(size_t n, SecurityLevel sec = SecurityLevel::tc128) {
        static const size_t degs[6] = {1024, 2048, 4096, 8192, 16384, 32768};
        static const int t128[6] = {27, 54, 109, 218, 438, 881};
        static const int t192[6] = {19, 37, 75, 152, 305, 611};
        static const int t256[6] = {14, 29, 58, 118, 237, 476};
        if (sec == SecurityLevel::none) return 0x7fffffff;
        for (int i = 0; i < 6; i++)
            if (degs[i] == n)
                return sec == SecurityLevel::tc128 ? t128[i] : sec == SecurityLevel::tc192 ? t192[i] : t256[i];
        return 0;
    }
    static std::vector<Modulus> Create(size_t n, const std::vector<int> &bit_sizes) {
        return create(n, bit_sizes, 0);
    }
    // CoeffModulus::BFVDefault (src/modulus.cpp:53-78): the HE-standard default primes for
    // N = 1024 ... 32768 at 128-, 192- and 256-bit security -- a published table, kept
    // word for word (src/utils/globals.cpp:10-200), one row per degree.
    static std::vector<Modulus> BFVDefault(size_t n, SecurityLevel sec = SecurityLevel::tc128) {
        if (!MaxBitCount(n, sec)) throw std::invalid_argument("non-standard poly_modulus_degree");
        if (sec == SecurityLevel::none) throw std::invalid_argument("invalid security level");
        static const std::vector<uint64_t> t128[6] = {
            {0x7e00001},
            {0x3fffffff000001},
            {0xffffee001, 0xffffc4001, 0x1ffffe0001},
            {0x7fffffd8001, 0x7fffffc8001, 0xfffffffc001, 0xffffff6c001, 0xfffffebc001},
            {0xfffffffd8001, 0xfffffffa0001, 0xfffffff00001, 0x1fffffff68001, 0x1fffffff50001, 0x1ffffffee8001,
             0x1ffffffea0001, 0x1ffffffe88001, 0x1ffffffe48001},
            {0x7fffffffe90001, 0x7fffffffbf0001, 0x7fffffffbd0001, 0x7fffffffba0001, 0x7fffffffaa0001,
             0x7fffffffa50001, 0x7fffffff9f0001, 0x7fffffff7e0001, 0x7fffffff770001, 0x7fffffff380001,
             0x7fffffff330001, 0x7fffffff2d0001, 0x7fffffff170001, 0x7fffffff150001, 0x7ffffffef00001,
             0xfffffffff70001}};
        static const std::vector<uint64_t> t192[6] = {
            {0x7f001},
            {0x1ffffc0001},
            {0x1ffc001, 0x1fce001, 0x1fc0001},
            {0x3ffffac001, 0x3ffff54001, 0x3ffff48001, 0x3ffff28001},
            {0x3ffffffdf0001, 0x3ffffffd48001, 0x3ffffffd20001, 0x3ffffffd18001, 0x3ffffffcd0001, 0x3ffffffc70001},
            {0x3fffffffd60001, 0x3fffffffca0001, 0x3fffffff6d0001, 0x3fffffff5d0001, 0x3fffffff550001,
             0x7fffffffe90001, 0x7fffffffbf0001, 0x7fffffffbd0001, 0x7fffffffba0001, 0x7fffffffaa0001,
             0x7fffffffa50001}};
        static const std::vector<uint64_t> t256[6] = {
            {0x3001},
            {0x1ffc0001},
            {0x3ffffffff040001},
            {0x7ffffec001, 0x7ffffb0001, 0xfffffdc001},
            {0x7ffffffc8001, 0x7ffffff00001, 0x7fffffe70001, 0xfffffffd8001, 0xfffffffa0001},
            {0xffffffff00001, 0x1fffffffe30001, 0x1fffffffd80001, 0x1fffffffd10001, 0x1fffffffc50001,
             0x1fffffffbf0001, 0x1fffffffb90001, 0x1fffffffb60001, 0x1fffffffa50001}};
        int row = 0;
        while ((size_t(1024) << row) != n) row++;
        const std::vector<uint64_t> &v = sec == SecurityLevel::tc128 ? t128[row] : sec == SecurityLevel::tc192 ? t192[row] : t256[row];
        return std::vector<Modulus>(v.begin(), v.end());
    }
    // Primes that are also 1 mod t (CoeffModulus::Create(n, plain_modulus, bit_sizes)).
    static std::vector<Modulus> Create(size_t n, const Modulus &plain_modulus, const std::vector<int> &bit_sizes) {
        return create(n, bit_sizes, plain_modulus.value());
    }

private:
    static std::vector<Modulus> create(size_t n, const std::vector<int> &bits, uint64_t t) {
        std::vector<uint64_t> out(bits.size());
        check(shc_coeff_modulus_create((uint32_t)n, bits.data(), (uint32_t)bits.size(), t, out.data()));
        return std::vector<Modulus>(out.begin(), out.end());
    }
};

// PlainModulus::Batching (src/modulus.h:486-520): a prime 1 mod 2N of the given size.
struct PlainModulus {
    static Modulus Batching(size_t n, int bit_size) { return CoeffModulus::Create(n, {bit_size})[0]; }
    static std::vector<Modulus> Batching(size_t n, const std::vector<int> &bit_sizes) {
        return CoeffModulus::Create(n, bit_sizes);
    }
};

// EncryptionParametersCuda (src/encryptionparams_cuda.cuh:63-204).
class EncryptionParameters {
public:
    explicit EncryptionParameters(SchemeType scheme = SchemeType::none) : scheme_(scheme) {}
    void setPolyModulusDegree(size_t n) { n_ = n; }
    void setCoeffModulus(const std::vector<Modulus> &q) { q_ = q; }
    void setPlainModulus(const Modulus &t) { t_ = t; }
    void setPlainModulus(uint64_t t) { t_ = Modulus(t); }
    SchemeType scheme() const { return scheme_; }
    size_t polyModulusDegree() const { return n_; }
    const std::vector<Modulus> &coeffModulus() const { return q_; }
    const Modulus &plainModulus() const { return t_; }

private:
    SchemeType scheme_;
    size_t n_ = 0;
    std::vector<Modulus> q_;
    Modulus t_;
};

// ParmsID (src/encryptionparams.h:36): the BLAKE2b-256 hash block of a level's
// parameters -- scheme, N, that level's primes, plain modulus -- as 4 little-endian
// words (EncryptionParameters::computeParmsID, src/encryptionparams.cpp:118-156).
using ParmsID = std::array<uint64_t, 4>;
static const ParmsID parmsIDZero{0, 0, 0, 0};

class SEALContext;

// EncryptionParameterQualifiers (src/context.h:22-215): what the parameters of a level
// support (every level of a context that was created is valid here).
struct EncryptionParameterQualifiers {
    enum class ErrorType : int { none = -1, success = 0 };
    ErrorType parameter_error = ErrorType::success;
    bool using_fft = true;
    bool using_ntt = true;
    bool using_batching = false;
    bool using_fast_plain_lift = false;
    bool using_descending_modulus_chain = false;
    SecurityLevel sec_level = SecurityLevel::tc128;
    bool parametersSet() const { return parameter_error == ErrorType::success; }
};

namespace detail {
// Deterministic Miller-Rabin for 64-bit values (the bases 2..37 decide every n < 2^64).
inline bool is_prime64(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull})
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) {
        d >>= 1;
        r++;
    }
    auto mulm = [n](uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % n); };
    for (uint64_t a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
        uint64_t x = 1, base = a, e = d;
        while (e) {
            if (e & 1) x = mulm(x, base);
            base = mulm(base, base);
            e >>= 1;
        }
        if (x == 1 || x == n - 1) continue;
        bool composite = true;
        for (int k = 1; k < r && composite; k++) {
            x = mulm(x, x);
            if (x == n - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}
}  // namespace detail

// ContextDataCuda (src/context_cuda.cuh:11-170): one level of the chain.
class ContextData {
public:
    ContextData(std::shared_ptr<const SEALContext> ctx, ParmsID id, uint32_t level, uint32_t last_level,
                std::vector<Modulus> q, SchemeType scheme, size_t n, Modulus t, ParmsID next, SecurityLevel sec)
        : ctx_(std::move(ctx)), id_(id), level_(level), last_(last_level), parms_(scheme), next_(next) {
        parms_.setPolyModulusDegree(n);
        parms_.setCoeffModulus(q);
        parms_.setPlainModulus(t);
        // the reference's qualifier rules (src/context.cpp:205-420)
        qual_.sec_level = totalCoeffModulusBitCount() > CoeffModulus::MaxBitCount(n, sec) ? SecurityLevel::none : sec;
        if (scheme == SchemeType::bfv) {
            qual_.using_batching = detail::is_prime64(t.value()) && t.value() % (2 * n) == 1;
            qual_.using_fast_plain_lift = true;
            for (const Modulus &m : q) qual_.using_fast_plain_lift &= m.value() > t.value();
        } else {
            qual_.using_batching = true;
        }
        qual_.using_descending_modulus_chain = true;
        for (size_t i = 0; i + 1 < q.size(); i++) qual_.using_descending_modulus_chain &= q[i].value() > q[i + 1].value();
    }
    ParmsID parmsID() const { return id_; }
    // Position in the chain: 0 for the last level, growing towards the key level.
    size_t chainIndex() const { return level_ - last_; }
    const EncryptionParameters &parms() const { return parms_; }
    const EncryptionParameterQualifiers &qualifiers() const { return qual_; }
    const std::vector<Modulus> &coeffModulus() const { return parms_.coeffModulus(); }
    SchemeType scheme() const { return parms_.scheme(); }
    size_t polyModulusDegree() const { return parms_.polyModulusDegree(); }
    const Modulus &plainModulus() const { return parms_.plainModulus(); }
    int totalCoeffModulusBitCount() const {
        int b = 0;
        for (auto &m : parms_.coeffModulus()) b += m.bitCount();
        return b;
    }
    // The product of the level's primes as little-endian 64-bit words, on the host (the
    // reference keeps the same words on the device).
    std::vector<uint64_t> totalCoeffModulus() const {
        std::vector<uint64_t> w{1};
        for (const Modulus &m : parms_.coeffModulus()) {
            unsigned __int128 carry = 0;
            for (uint64_t &x : w) {
                carry += (unsigned __int128)x * m.value();
                x = (uint64_t)carry;
                carry >>= 64;
            }
            if (carry) w.push_back((uint64_t)carry);
        }
        w.resize(parms_.coeffModulus().size(), 0);
        return w;
    }
    // (t + 1) / 2: plaintext coefficients at or above it stand for negative values.
    uint64_t plainUpperHalfThreshold() const { return (parms_.plainModulus().value() + 1) >> 1; }
    // The product of the primes modulo t (BFV).
    uint64_t coeffModulusModPlainModulus() const {
        const uint64_t t = parms_.plainModulus().value();
        if (!t) return 0;
        uint64_t r = 1 % t;
        for (const Modulus &m : parms_.coeffModulus()) r = (uint64_t)((unsigned __int128)r * (m.value() % t) % t);
        return r;
    }
    ParmsID nextParmsID() const { return next_; }
    // The neighbouring levels (src/context_cuda.cuh:88-97): next drops a prime, prev adds
    // one (the first level's prev is the key level); null past either end.
    std::shared_ptr<const ContextData> nextContextData() const;
    std::shared_ptr<const ContextData> prevContextData() const;

private:
    std::shared_ptr<const SEALContext> ctx_;
    ParmsID id_;
    uint32_t level_, last_;
    EncryptionParameters parms_;
    EncryptionParameterQualifiers qual_;
    ParmsID next_;
};

namespace detail {
// 64 bits from the operating system's generator (std::random_device).
inline uint64_t os_random64() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ rd();
}
}  // namespace detail

// SEALContextCuda (src/context_cuda.cuh:172-238).  Copies share the device tables.
// Levels are named by their parms_id; the chain holds the key level (all K primes)
// and the data levels K-1 .. 1 (only K-1 when expand_mod_chain is false).  Every
// context samples keys and encryptions from ChaCha20 under a fresh 256-bit key from
// std::random_device (shc_context_set_sampler_key).
class SEALContext {
public:
    explicit SEALContext(const EncryptionParameters &parms, bool expand_mod_chain = true,
                         SecurityLevel sec = SecurityLevel::tc128)
        : s_(std::make_shared<State>()) {
        KernelProvider::checkInitialized();
        s_->parms = parms;
        s_->expand = expand_mod_chain;
        s_->sec = sec;
        const auto &q = parms.coeffModulus();
        int bits = 0;
        for (auto &m : q) bits += m.bitCount();
        if (parms.scheme() != SchemeType::bfv && parms.scheme() != SchemeType::ckks) {
            s_->error = "scheme not supported";
            return;
        }
        if (q.empty()) {
            s_->error = "coeff_modulus is empty";
            return;
        }
        if (bits > CoeffModulus::MaxBitCount(parms.polyModulusDegree(), sec)) {
            s_->error = "parameters are not compliant with HomomorphicEncryption.org security standard";
            return;
        }
        std::vector<uint64_t> v;
        for (auto &m : q) v.push_back(m.value());
        shc_context *h = nullptr;
        const int rc = shc_context_create((int)parms.scheme(), (uint32_t)parms.polyModulusDegree(), v.data(),
                                          (uint32_t)v.size(), parms.plainModulus().value(), &h);
        if (rc != SHC_OK) {
            s_->error = shc_last_error();
            return;
        }
        s_->handle = h;
        uint8_t key[32];
        {
            std::random_device rd;
            for (int i = 0; i < 8; i++) {
                const uint32_t w = rd();
                for (int b = 0; b < 4; b++) key[4 * i + b] = (uint8_t)(w >> (8 * b));
            }
        }
        check(shc_context_set_sampler_key(h, key));
        std::memset(key, 0, sizeof key);
        s_->ids.resize(q.size() + 1, parmsIDZero);
        for (uint32_t l = 1; l <= q.size(); l++) {
            std::vector<uint64_t> ql(v.begin(), v.begin() + l);
            s_->ids[l] = detail::parms_hash((uint64_t)parms.scheme(), parms.polyModulusDegree(), ql,
                                            parms.plainModulus().value());
            s_->level_of[s_->ids[l]] = l;
        }
    }

    bool parametersSet() const { return s_->handle != nullptr; }
    const std::string &parameterErrorMessage() const { return s_->error; }
    void checkSet() const {
        if (!parametersSet()) throw std::invalid_argument("encryption parameters are not set correctly: " + s_->error);
    }

    SchemeType scheme() const { return s_->parms.scheme(); }
    size_t polyModulusDegree() const { return s_->parms.polyModulusDegree(); }
    const EncryptionParameters &parms() const { return s_->parms; }
    size_t keyCount() const { return s_->parms.coeffModulus().size(); }

    ParmsID keyParmsID() const { return idOf((uint32_t)keyCount()); }
    ParmsID firstParmsID() const { return idOf(firstLevel()); }
    ParmsID lastParmsID() const { return idOf(s_->expand ? 1 : firstLevel()); }
    bool usingKeyswitching() const { return keyCount() > 1; }
    bool using_keyswitching() const { return usingKeyswitching(); }  // the reference's spelling

    std::shared_ptr<const ContextData> getContextData(ParmsID id) const {
        const uint32_t l = levelOf(id);
        if (!l || !inChain(l)) return nullptr;
        const auto &q = s_->parms.coeffModulus();
        const uint32_t last = s_->expand ? 1 : firstLevel();
        // the key level's next is the first level (src/context.cpp:464-500)
        const ParmsID next = l == keyCount() ? (keyCount() > 1 ? firstParmsID() : parmsIDZero)
                                             : (l > last ? idOf(l - 1) : parmsIDZero);
        return std::make_shared<ContextData>(std::make_shared<const SEALContext>(*this), id, l, last,
                                             std::vector<Modulus>(q.begin(), q.begin() + l), scheme(),
                                             polyModulusDegree(), s_->parms.plainModulus(), next, s_->sec);
    }
    std::shared_ptr<const ContextData> firstContextData() const { return getContextData(firstParmsID()); }
    std::shared_ptr<const ContextData> keyContextData() const { return getContextData(keyParmsID()); }
    std::shared_ptr<const ContextData> lastContextData() const { return getContextData(lastParmsID()); }

    // ---- used by the other classes
    shc_context *handle() const {
        checkSet();
        return s_->handle;
    }
    // The number of primes of the level named `id` (0 when no level of this context has
    // that parms_id).
    uint32_t levelOf(const ParmsID &id) const {
        auto it = s_->level_of.find(id);
        return it == s_->level_of.end() ? 0 : it->second;
    }
    // The same, throwing std::invalid_argument for an unknown parms_id.
    uint32_t levelOrThrow(const ParmsID &id, const char *what) const {
        const uint32_t l = levelOf(id);
        if (!l || !inChain(l)) throw std::invalid_argument(std::string(what) + " is not valid for encryption parameters");
        return l;
    }
    // The parms_id of the level with `level` primes.
    ParmsID idOf(uint32_t level) const {
        checkSet();
        if (level == 0 || level > keyCount()) throw std::invalid_argument("no such level");
        return s_->ids[level];
    }
    // Scratch device memory shared by the operations of this context (one stream; the
    // reference's evaluators are not thread safe either).
    void *workspace(size_t bytes) const {
        s_->ws.ensure((bytes + 7) / 8);
        return s_->ws.get();
    }
    uint64_t prime(size_t i) const { return s_->parms.coeffModulus()[i].value(); }
    int levelBits(uint32_t level) const {
        int b = 0;
        for (uint32_t i = 0; i < level; i++) b += s_->parms.coeffModulus()[i].bitCount();
        return b;
    }

private:
    uint32_t firstLevel() const { return (uint32_t)(keyCount() > 1 ? keyCount() - 1 : 1); }
    bool inChain(uint32_t l) const {
        return l == keyCount() || (l <= firstLevel() && l >= (s_->expand ? 1u : firstLevel()));
    }
    struct State {
        EncryptionParameters parms;
        bool expand = true;
        SecurityLevel sec = SecurityLevel::tc128;
        std::string error;
        shc_context *handle = nullptr;
        std::vector<ParmsID> ids;               // by level (prime count)
        std::map<ParmsID, uint32_t> level_of;   // parms_id -> level
        DeviceArray<uint64_t> ws;
        ~State() {
            if (handle) shc_context_destroy(handle);
        }
    };
    std::shared_ptr<State> s_;
};

inline std::shared_ptr<const ContextData> ContextData::nextContextData() const {
    return next_ == parmsIDZero ? nullptr : ctx_->getContextData(next_);
}
inline std::shared_ptr<const ContextData> ContextData::prevContextData() const {
    // every level between a chain level and the key level is in the chain
    return level_ == ctx_->keyCount() ? nullptr : ctx_->getContextData(ctx_->idOf(level_ + 1));
}

namespace detail {
template <class T>
inline void put(std::ostream &os, const T &v) {
    os.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <class T>
inline T get(std::istream &is, const char *what = "stream ended early") {
    T v{};
    is.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!is) throw std::invalid_argument(what);
    return v;
}

// The fields that open a saved ciphertext (src/ciphertext_cuda.cu:8-17): parms_id (32 B),
// is_ntt_form (1 B), size, N, coefficient modulus count (8 B each), scale (double),
// correction factor, seed, terms flag (1 B); the word count and the words follow.
struct CtHeader {
    ParmsID id{};
    bool ntt = false;
    uint64_t size = 0, n = 0, L = 0;
    double scale = 1.0;
    uint64_t correction = 1, seed = 0;
    bool terms = false;
};
inline void put_header(std::ostream &os, const CtHeader &h) {
    put(os, h.id);
    put(os, h.ntt);
    put(os, h.size);
    put(os, h.n);
    put(os, h.L);
    put(os, h.scale);
    put(os, h.correction);
    put(os, h.seed);
    put(os, h.terms);
}
inline CtHeader get_header(std::istream &is) {
    const char *what = "stream ended before the ciphertext did";
    CtHeader h;
    h.id = get<ParmsID>(is, what);
    h.ntt = get<bool>(is, what);
    h.size = get<uint64_t>(is, what);
    h.n = get<uint64_t>(is, what);
    h.L = get<uint64_t>(is, what);
    h.scale = get<double>(is, what);
    h.correction = get<uint64_t>(is, what);
    h.seed = get<uint64_t>(is, what);
    h.terms = get<bool>(is, what);
    return h;
}
// `words` words into h (the count was read and checked by the caller).
inline void get_words(std::istream &is, std::vector<uint64_t> &h, size_t words) {
    h.resize(words);
    is.read(reinterpret_cast<char *>(h.data()), (std::streamsize)(words * 8));
    if (!is) throw std::invalid_argument("stream ended before the ciphertext did");
}

// Device words to and from a stream through one pinned staging buffer (hipHostMalloc,
// 8 MiB, allocated on first use and kept for the process; calls take it in turn): an
// object moves in chunks at pinned-copy speed instead of through a pageable host copy of
// the whole object.
struct PinnedStage {
    static constexpr size_t BYTES = size_t(8) << 20;
    std::mutex mu;
    char *buf = nullptr;
    static PinnedStage &get() {
        static PinnedStage *s = new PinnedStage;  // never destroyed: outlives the runtime's teardown
        return *s;
    }
    char *take() {
        if (!buf) hip_check(hipHostMalloc(reinterpret_cast<void **>(&buf), BYTES), "pinned staging buffer");
        return buf;
    }
};

inline void write_device_words(std::ostream &os, const uint64_t *d, size_t words) {
    PinnedStage &s = PinnedStage::get();
    std::lock_guard<std::mutex> g(s.mu);
    char *b = s.take();
    const size_t per = PinnedStage::BYTES / 8;
    for (size_t done = 0; done < words; done += per) {
        const size_t k = std::min(per, words - done);
        hip_check(hipMemcpy(b, d + done, k * 8, hipMemcpyDeviceToHost), "copy to host");
        os.write(b, (std::streamsize)(k * 8));
    }
}

inline void read_device_words(std::istream &is, uint64_t *d, size_t words, const char *what) {
    PinnedStage &s = PinnedStage::get();
    std::lock_guard<std::mutex> g(s.mu);
    char *b = s.take();
    const size_t per = PinnedStage::BYTES / 8;
    for (size_t done = 0; done < words; done += per) {
        const size_t k = std::min(per, words - done);
        is.read(b, (std::streamsize)(k * 8));
        if (!is) throw std::invalid_argument(what);
        hip_check(hipMemcpy(d + done, b, k * 8, hipMemcpyHostToDevice), "copy to device");
    }
}

// Plaintext::operator=(const std::string &) (src/plaintext.cpp:91-200): "7FFx^3 + 1x^1
// + 3" -- hexadecimal coefficients, decimal exponents strictly decreasing, "x^1" for
// the linear term, a bare constant term, terms joined by exactly " + ".
inline std::vector<uint64_t> parse_hex_poly(const std::string &s) {
    auto is_hex = [](char c) { return std::isxdigit((unsigned char)c) != 0; };
    auto hex_val = [](char c) { return c <= '9' ? c - '0' : (c | 0x20) - 'a' + 10; };
    std::vector<std::pair<uint64_t, uint64_t>> terms;  // (power, coefficient)
    size_t pos = 0;
    uint64_t last = UINT64_MAX;
    while (pos < s.size()) {
        size_t len = 0;
        uint64_t coeff = 0;
        int bits = 0;
        while (pos + len < s.size() && is_hex(s[pos + len])) {
            const int v = hex_val(s[pos + len]);
            if (bits || v) bits = bits ? bits + 4 : (v >= 8 ? 4 : v >= 4 ? 3 : v >= 2 ? 2 : 1);
            if (bits > 64) throw std::invalid_argument("hex_poly has too large coefficients");
            coeff = (coeff << 4) | (uint64_t)v;
            len++;
        }
        if (!len) throw std::invalid_argument("unable to parse hex_poly");
        pos += len;
        uint64_t power = 0;
        if (pos < s.size() && s[pos] != ' ') {
            if (s[pos] != 'x' || pos + 1 >= s.size() || s[pos + 1] != '^')
                throw std::invalid_argument("unable to parse hex_poly");
            pos += 2;
            size_t digits = 0;
            while (pos < s.size() && s[pos] >= '0' && s[pos] <= '9') {
                power = power * 10 + (uint64_t)(s[pos] - '0');
                if (power > (uint64_t)0x7fffffff) throw std::invalid_argument("unable to parse hex_poly");
                pos++;
                digits++;
            }
            if (!digits) throw std::invalid_argument("unable to parse hex_poly");
        }
        if (power >= last) throw std::invalid_argument("unable to parse hex_poly");
        last = power;
        terms.emplace_back(power, coeff);
        if (pos < s.size()) {
            if (s.compare(pos, 3, " + ") != 0) throw std::invalid_argument("unable to parse hex_poly");
            pos += 3;  // a trailing " + " ends the string, as the reference allows
        }
    }
    std::vector<uint64_t> out(terms.empty() ? 0 : (size_t)terms.front().first + 1, 0);
    bool any = false;
    for (auto &t : terms) {
        out[t.first] = t.second;
        any = any || t.second;
    }
    if (!any) out.clear();  // all-zero: the empty polynomial (setZero of nothing)
    return out;
}

// Plaintext::to_string (src/plaintext.h:491-498 over util::polyToHexString): nonzero
// terms from the highest power down, upper-case hexadecimal, "0" for the zero polynomial.
inline std::string hex_poly_string(const uint64_t *c, size_t count) {
    std::ostringstream os;
    bool empty = true;
    for (size_t i = count; i-- > 0;) {
        if (!c[i]) continue;
        if (!empty) os << " + ";
        char buf[24];
        std::snprintf(buf, sizeof buf, "%llX", (unsigned long long)c[i]);
        os << buf;
        if (i) os << "x^" << i;
        empty = false;
    }
    if (empty) os << "0";
    return os.str();
}
}  // namespace detail

// ------------------------------------------------------------------ data containers

// The CPU-side containers that the reference's cpu() / toHost() return and its
// conversion constructors take (troy::Plaintext, troy::Ciphertext, src/plaintext.h,
// src/ciphertext.h): the same fields in host memory.
namespace host {

class Plaintext {
public:
    Plaintext() = default;
    explicit Plaintext(size_t coeff_count) : data_(coeff_count, 0) {}
    Plaintext(const std::string &hex_poly) { *this = hex_poly; }
    Plaintext &operator=(const std::string &hex_poly) {
        if (isNttForm()) throw std::logic_error("cannot set an NTT transformed Plaintext");
        data_ = detail::parse_hex_poly(hex_poly);
        return *this;
    }
    Plaintext &operator=(uint64_t const_coeff) {
        data_.assign(1, const_coeff);
        parms_id_ = parmsIDZero;
        return *this;
    }
    void resize(size_t coeff_count) {
        if (isNttForm()) throw std::logic_error("cannot resize an NTT transformed Plaintext");
        data_.resize(coeff_count, 0);
    }
    size_t coeffCount() const { return data_.size(); }
    size_t significantCoeffCount() const {
        size_t k = data_.size();
        while (k && !data_[k - 1]) k--;
        return k;
    }
    size_t nonzeroCoeffCount() const { return (size_t)std::count_if(data_.begin(), data_.end(), [](uint64_t x) { return x; }); }
    bool isZero() const { return significantCoeffCount() == 0; }
    bool isNttForm() const { return parms_id_ != parmsIDZero; }
    std::string to_string() const {
        if (isNttForm()) throw std::invalid_argument("cannot convert NTT transformed plaintext to string");
        return detail::hex_poly_string(data_.data(), data_.size());
    }
    uint64_t operator[](size_t i) const { return data_.at(i); }
    uint64_t &operator[](size_t i) { return data_.at(i); }
    // Plaintext::operator== (src/plaintext.h:396-410): the same significant
    // coefficients, the same NTT level (or neither in NTT form), close scales.
    bool operator==(const Plaintext &o) const {
        const size_t a = significantCoeffCount(), b = o.significantCoeffCount();
        const bool ids = (isNttForm() && o.isNttForm() && parms_id_ == o.parms_id_) || (!isNttForm() && !o.isNttForm());
        const double d = std::fabs(scale_ - o.scale_);
        return ids && a == b && std::equal(data_.begin(), data_.begin() + a, o.data_.begin()) &&
               d <= std::max(std::fabs(scale_), std::fabs(o.scale_)) * 1e-9;
    }
    bool operator!=(const Plaintext &o) const { return !(*this == o); }
    ParmsID &parmsID() { return parms_id_; }
    const ParmsID &parmsID() const { return parms_id_; }
    double &scale() { return scale_; }
    double scale() const { return scale_; }
    std::vector<uint64_t> &data() { return data_; }
    const std::vector<uint64_t> &data() const { return data_; }

private:
    ParmsID parms_id_ = parmsIDZero;
    double scale_ = 1.0;
    std::vector<uint64_t> data_;
};

class Ciphertext {
public:
    ParmsID &parmsID() { return parms_id_; }
    const ParmsID &parmsID() const { return parms_id_; }
    bool &isNttForm() { return ntt_form_; }
    bool isNttForm() const { return ntt_form_; }
    size_t &size() { return size_; }
    size_t size() const { return size_; }
    size_t &polyModulusDegree() { return n_; }
    size_t polyModulusDegree() const { return n_; }
    size_t &coeffModulusSize() { return L_; }
    size_t coeffModulusSize() const { return L_; }
    double &scale() { return scale_; }
    double scale() const { return scale_; }
    uint64_t &correctionFactor() { return correction_; }
    uint64_t correctionFactor() const { return correction_; }
    std::vector<uint64_t> &data() { return data_; }
    const std::vector<uint64_t> &data() const { return data_; }
    // All-zero c1.. (an unencrypted value, Ciphertext::isTransparent, src/ciphertext.h).
    bool isTransparent() const {
        if (size_ < 2) return true;
        return std::all_of(data_.begin() + (std::ptrdiff_t)(n_ * L_), data_.end(), [](uint64_t x) { return !x; });
    }
    bool operator==(const Ciphertext &o) const {
        return parms_id_ == o.parms_id_ && ntt_form_ == o.ntt_form_ && size_ == o.size_ && n_ == o.n_ &&
               L_ == o.L_ && scale_ == o.scale_ && correction_ == o.correction_ && data_ == o.data_;
    }
    bool operator!=(const Ciphertext &o) const { return !(*this == o); }

private:
    ParmsID parms_id_ = parmsIDZero;
    bool ntt_form_ = false;
    size_t size_ = 0, n_ = 0, L_ = 0;
    double scale_ = 1.0;
    uint64_t correction_ = 1;
    std::vector<uint64_t> data_;
};

}  // namespace host

// PlaintextCuda (src/plaintext_cuda.cuh): BFV plaintexts are coefficients mod t
// (coefficient form, parms_id zero) or, after transformToNtt, [L][N] in NTT form; CKKS
// plaintexts are [L][N] in NTT form with that level's parms_id.
class Plaintext {
public:
    using pt_coeff_type = uint64_t;
    Plaintext() = default;
    explicit Plaintext(size_t coeff_count) { resize(coeff_count); }
    Plaintext(const host::Plaintext &h) : parms_id_(h.parmsID()), scale_(h.scale()) {
        data_.fromHost(h.data().data(), h.data().size());
    }
    Plaintext(const std::string &hex_poly) { *this = hex_poly; }
    Plaintext(const Plaintext &) = default;
    Plaintext(Plaintext &&) = default;
    Plaintext &operator=(const Plaintext &) = default;
    Plaintext &operator=(Plaintext &&) = default;
    Plaintext &operator=(const std::string &hex_poly) {
        if (isNttForm()) throw std::logic_error("cannot set an NTT transformed Plaintext");
        host::Plaintext h(hex_poly);
        data_.fromHost(h.data().data(), h.data().size());
        return *this;
    }
    Plaintext &operator=(uint64_t const_coeff) {
        data_.fromHost(&const_coeff, 1);
        parms_id_ = parmsIDZero;
        return *this;
    }

    host::Plaintext cpu() const {
        host::Plaintext h;
        h.parmsID() = parms_id_;
        h.scale() = scale_;
        h.data() = data_.toHost();
        return h;
    }
    host::Plaintext toHost() const { return cpu(); }
    std::string to_string() const { return cpu().to_string(); }
    bool isZero() const { return cpu().isZero(); }
    bool operator==(const Plaintext &o) const { return cpu() == o.cpu(); }
    bool operator!=(const Plaintext &o) const { return !(*this == o); }
    uint64_t operator[](size_t i) const {
        if (i >= coeffCount()) throw std::out_of_range("coeff_index must be within [0, coeff_count)");
        uint64_t v = 0;
        hip_check(hipMemcpy(&v, data_.get() + i, 8, hipMemcpyDeviceToHost), "copy to host");
        return v;
    }

    void resize(size_t coeff_count) {
        if (isNttForm()) throw std::logic_error("cannot resize an NTT transformed Plaintext");
        data_.resizeKeep(coeff_count);
    }
    void release() {
        parms_id_ = parmsIDZero;
        scale_ = 1.0;
        data_.release();
    }
    void setZero() { data_.setZero(); }
    // PlaintextCuda::reserve / shrinkToFit (src/plaintext_cuda.cuh:59-72): capacity for
    // `capacity` coefficients, keeping the current ones.
    void reserve(size_t capacity) {
        if (isNttForm()) throw std::logic_error("cannot reserve for an NTT transformed Plaintext");
        data_.reserve(capacity);
    }
    void shrinkToFit() { data_.shrinkToFit(); }
    void setZero(size_t start_coeff) {
        if (start_coeff >= coeffCount()) throw std::out_of_range("start_coeff must be within [0, coeff_count)");
        hip_check(hipMemset(data_.get() + start_coeff, 0, (coeffCount() - start_coeff) * 8), "memset");
    }

    // PlaintextCuda::save / load (src/plaintext_cuda.cu:6-27): parms_id (32 B), coeff
    // count, scale, word count, words.
    void save(std::ostream &os) const {
        detail::put(os, parms_id_);
        detail::put(os, (uint64_t)coeffCount());
        detail::put(os, scale_);
        detail::put(os, (uint64_t)data_.size());
        detail::write_device_words(os, data_.get(), data_.size());
    }
    void load(std::istream &is) {
        const char *what = "stream ended before the plaintext did";
        const ParmsID id = detail::get<ParmsID>(is, what);
        const uint64_t count = detail::get<uint64_t>(is, what);
        const double scale = detail::get<double>(is, what);
        const uint64_t words = detail::get<uint64_t>(is, what);
        if (words != count) throw std::invalid_argument("plaintext word count does not match its coefficient count");
        DeviceArray<uint64_t> d(words);  // the plaintext changes only once the words are in
        detail::read_device_words(is, d.get(), words, what);
        data_ = std::move(d);
        parms_id_ = id;
        scale_ = scale;
    }

    ParmsID &parmsID() { return parms_id_; }
    const ParmsID &parmsID() const { return parms_id_; }
    double &scale() { return scale_; }
    double scale() const { return scale_; }
    bool isNttForm() const { return parms_id_ != parmsIDZero; }
    size_t coeffCount() const { return data_.size(); }
    size_t capacity() const { return data_.capacity(); }
    uint64_t *data() { return data_.get(); }
    const uint64_t *data() const { return data_.get(); }
    DeviceArray<uint64_t> &dynArray() { return data_; }
    const DeviceArray<uint64_t> &dynArray() const { return data_; }

private:
    DeviceArray<uint64_t> data_;
    ParmsID parms_id_ = parmsIDZero;
    double scale_ = 1.0;
};

class Evaluator;

// CiphertextCuda (src/ciphertext_cuda.cuh:12-246): [size][L][N] words on the device.
class Ciphertext {
public:
    Ciphertext() = default;
    // Empty ciphertexts of the first level (or of `id`), as CiphertextCuda(context[, id]).
    explicit Ciphertext(const SEALContext &ctx) { setLevel(ctx, ctx.firstParmsID()); }
    Ciphertext(const SEALContext &ctx, ParmsID id) { setLevel(ctx, id); }
    Ciphertext(const SEALContext &ctx, ParmsID id, size_t size) { resize(ctx, id, size); }
    Ciphertext(const host::Ciphertext &h)
        : parms_id_(h.parmsID()), size_(h.size()), n_(h.polyModulusDegree()), L_(h.coeffModulusSize()),
          ntt_form_(h.isNttForm()), scale_(h.scale()), correction_(h.correctionFactor()) {
        data_.fromHost(h.data().data(), h.data().size());
    }
    Ciphertext(const Ciphertext &) = default;
    Ciphertext(Ciphertext &&) = default;
    Ciphertext &operator=(const Ciphertext &) = default;
    Ciphertext &operator=(Ciphertext &&) = default;

    void resize(const SEALContext &ctx, ParmsID id, size_t size) {
        setLevel(ctx, id);
        resize(size);
    }
    void resize(const SEALContext &ctx, size_t size) { resize(ctx, ctx.firstParmsID(), size); }
    void resize(size_t size) {
        size_ = size;
        data_.resize(size * L_ * n_);
    }
    // CiphertextCuda::reserve / sizeCapacity (src/ciphertext_cuda.cuh:125-215): room for
    // size_capacity polynomials of the level; the words that fit are kept and the size is
    // capped at size_capacity, as the reference's reserveInternal.
    void reserve(const SEALContext &ctx, ParmsID id, size_t size_capacity) {
        const uint32_t L = ctx.levelOrThrow(id, "parms_id");
        parms_id_ = id;
        reserveWords(size_capacity, ctx.polyModulusDegree(), L);
    }
    void reserve(const SEALContext &ctx, size_t size_capacity) { reserve(ctx, ctx.firstParmsID(), size_capacity); }
    void reserve(size_t size_capacity) { reserveWords(size_capacity, n_, L_); }
    size_t sizeCapacity() const { return L_ * n_ ? data_.capacity() / (L_ * n_) : 0; }
    // Ciphertexts here are never seeded (c1 is always stored).
    uint64_t seed() const { return 0; }
    void release() {
        parms_id_ = parmsIDZero;
        ntt_form_ = false;
        size_ = n_ = L_ = 0;
        scale_ = 1.0;
        correction_ = 1;
        data_.release();
    }

    host::Ciphertext cpu() const {
        host::Ciphertext h;
        h.parmsID() = parms_id_;
        h.isNttForm() = ntt_form_;
        h.size() = size_;
        h.polyModulusDegree() = n_;
        h.coeffModulusSize() = L_;
        h.scale() = scale_;
        h.correctionFactor() = correction_;
        h.data() = data_.toHost();
        h.data().resize(size_ * L_ * n_);
        return h;
    }
    host::Ciphertext toHost() const { return cpu(); }
    bool isTransparent() const { return cpu().isTransparent(); }

    // CiphertextCuda::save / load (src/ciphertext_cuda.cu:8-35, 138-183): parms_id
    // (32 B), is_ntt_form (1 B), size, N, L (8 B each), scale (double), correction
    // factor, seed (0), terms flag (false), word count, words -- the reference's raw
    // layout, so files move between the two.  Seeded ciphertexts (c1 expanded from a
    // curand stream) are not loadable here.
    void save(std::ostream &os) const {
        saveHeader(os, false);
        const uint64_t words = (uint64_t)(size_ * polyWords());
        detail::put(os, words);
        detail::write_device_words(os, data_.get(), words);
    }
    void load(std::istream &is, const SEALContext &ctx) {
        loadHeader(is, ctx, false);
        loadWords(is);
    }
    // CiphertextCuda::load(stream) (src/ciphertext_cuda.cu:74-96): the shape comes from
    // the header alone, with no context to check it against (keys are stored this way).
    void load(std::istream &is) {
        const detail::CtHeader h = detail::get_header(is);
        if (h.terms) throw std::invalid_argument("Trying to load a termed ciphertext, but indices is not specified");
        if (h.seed != 0) throw std::invalid_argument("seed is not zero.");
        if (h.n == 0 || (h.n & (h.n - 1)) || h.n > (1u << 17) || h.L == 0 || h.L > 64 || h.size > 16)
            throw std::invalid_argument("ciphertext header is not valid");
        parms_id_ = h.id;
        ntt_form_ = h.ntt;
        n_ = (size_t)h.n;
        L_ = (size_t)h.L;
        scale_ = h.scale;
        correction_ = h.correction;
        resize((size_t)h.size);
        loadWords(is);
    }
    // saveTerms / loadTerms (src/ciphertext_cuda.cu:37-136): only the listed coefficients
    // of c0 (coefficient form), then every other polynomial in full.
    void saveTerms(std::ostream &os, const Evaluator &ev, const std::vector<size_t> &term_ids) const;
    void loadTerms(std::istream &is, const Evaluator &ev, const SEALContext &ctx, const std::vector<size_t> &term_ids);

    ParmsID &parmsID() { return parms_id_; }
    const ParmsID &parmsID() const { return parms_id_; }
    size_t size() const { return size_; }
    size_t polyModulusDegree() const { return n_; }
    size_t coeffModulusSize() const { return L_; }
    bool isNttForm() const { return ntt_form_; }
    bool &isNttForm() { return ntt_form_; }
    double scale() const { return scale_; }
    double &scale() { return scale_; }
    uint64_t correctionFactor() const { return correction_; }
    uint64_t &correctionFactor() { return correction_; }
    uint64_t *data() { return data_.get(); }
    const uint64_t *data() const { return data_.get(); }
    uint64_t *data(size_t poly) { return data_.get() + poly * L_ * n_; }
    const uint64_t *data(size_t poly) const { return data_.get() + poly * L_ * n_; }
    size_t polyWords() const { return L_ * n_; }
    size_t polyCoeffSize() const { return L_ * n_; }
    // Keeps the first `size` polynomials (relinearization drops the third).
    void shrink(size_t size) { size_ = size; }
    DeviceArray<uint64_t> &dynArray() { return data_; }
    const DeviceArray<uint64_t> &dynArray() const { return data_; }

private:
    void setLevel(const SEALContext &ctx, ParmsID id) {
        L_ = ctx.levelOrThrow(id, "parms_id");
        parms_id_ = id;
        n_ = ctx.polyModulusDegree();
    }
    void saveHeader(std::ostream &os, bool terms) const {
        detail::CtHeader h;
        h.id = parms_id_;
        h.ntt = ntt_form_;
        h.size = size_;
        h.n = n_;
        h.L = L_;
        h.scale = scale_;
        h.correction = correction_;
        h.terms = terms;
        detail::put_header(os, h);
    }
    void loadHeader(std::istream &is, const SEALContext &ctx, bool terms) {
        const detail::CtHeader h = detail::get_header(is);
        if (h.seed != 0) throw std::invalid_argument("seeded ciphertexts are not supported");
        if (h.terms != terms)
            throw std::invalid_argument(terms ? "Trying to load a normal ciphertext, but term indices is specified"
                                              : "Trying to load a termed ciphertext, but indices is not specified");
        if (ctx.levelOf(h.id) == 0 || ctx.levelOf(h.id) != h.L || h.n != ctx.polyModulusDegree())
            throw std::invalid_argument("ciphertext is not valid for encryption parameters");
        resize(ctx, h.id, h.size);
        ntt_form_ = h.ntt;
        scale_ = h.scale;
        correction_ = h.correction;
    }
    void reserveWords(size_t size_capacity, size_t n, size_t L) {
        if (size_capacity < 2 || size_capacity > 16) throw std::invalid_argument("invalid size_capacity");
        const size_t words = size_capacity * n * L;
        data_.reserve(words);
        data_.resizeKeep(std::min(words, data_.size()));
        size_ = std::min(size_capacity, size_);
        n_ = n;
        L_ = L;
    }
    void loadWords(std::istream &is) {
        const uint64_t words = detail::get<uint64_t>(is, "stream ended before the ciphertext did");
        if (words != size_ * polyWords()) throw std::invalid_argument("ciphertext word count does not match its shape");
        data_.resize(words);
        detail::read_device_words(is, data_.get(), words, "stream ended before the ciphertext did");
    }

    DeviceArray<uint64_t> data_;
    ParmsID parms_id_ = parmsIDZero;
    size_t size_ = 0, n_ = 0, L_ = 0;
    bool ntt_form_ = false;
    double scale_ = 1.0;
    uint64_t correction_ = 1;
};

// SecretKeyCuda: a key-level plaintext [K][N] in NTT form.
class SecretKey {
public:
    Plaintext &data() { return data_; }
    const Plaintext &data() const { return data_; }
    ParmsID &parmsID() { return data_.parmsID(); }
    const ParmsID &parmsID() const { return data_.parmsID(); }
    // SecretKeyCuda::save / load (src/secretkey_cuda.cuh:292-297): the plaintext's layout.
    void save(std::ostream &os) const { data_.save(os); }
    void load(std::istream &is) { data_.load(is); }

private:
    Plaintext data_;
};

// PublicKeyCuda: a size-2 key-level ciphertext in NTT form.
class PublicKey {
public:
    Ciphertext &data() { return data_; }
    const Ciphertext &data() const { return data_; }
    ParmsID &parmsID() { return data_.parmsID(); }
    const ParmsID &parmsID() const { return data_.parmsID(); }
    // PublicKeyCuda::save / load (src/publickey_cuda.cuh:252-257): the ciphertext's layout.
    void save(std::ostream &os) const { data_.save(os); }
    void load(std::istream &is) { data_.load(is); }

private:
    Ciphertext data_;
};

// KSwitchKeysCuda (src/kswitchkeys_cuda.cuh:31-170): a list of key-switching keys, each
// [K-1][2][K][N] in NTT form (the words of the reference's K-1 public keys, back to
// back); empty entries are absent keys.  The base of RelinKeys and GaloisKeys.
class KSwitchKeys {
public:
    // The number of keys present.
    size_t size() const {
        return (size_t)std::count_if(keys_.begin(), keys_.end(), [](const DeviceArray<uint64_t> &k) { return k.size(); });
    }
    std::vector<DeviceArray<uint64_t>> &data() { return keys_; }
    const std::vector<DeviceArray<uint64_t>> &data() const { return keys_; }
    DeviceArray<uint64_t> &data(size_t index) {
        if (index >= keys_.size() || !keys_[index].size()) throw std::invalid_argument("keyswitching key does not exist");
        return keys_[index];
    }
    const DeviceArray<uint64_t> &data(size_t index) const {
        if (index >= keys_.size() || !keys_[index].size()) throw std::invalid_argument("keyswitching key does not exist");
        return keys_[index];
    }
    ParmsID &parmsID() { return parms_id_; }
    const ParmsID &parmsID() const { return parms_id_; }
    // The shape of every key: degree N and key-level modulus count K (set by KeyGenerator
    // and by load; save needs it).
    void setShape(size_t n, size_t key_moduli) {
        n_ = n;
        K_ = key_moduli;
    }
    size_t polyModulusDegree() const { return n_; }
    size_t keyModulusSize() const { return K_; }

    // KSwitchKeysCuda::save / load (src/kswitchkeys_cuda.cuh:330-355): parms_id (32 B), the
    // number of key slots (8 B), then per slot the number of its public keys (8 B; 0 for
    // an absent key) and each public key in the ciphertext layout (size 2, K limbs, NTT
    // form, scale 1, correction 1; its parms_id is the key level's).  Public key j of a
    // slot is part j of this build's [K-1][2][K][N] key.
    void save(std::ostream &os) const {
        detail::put(os, parms_id_);
        detail::put(os, (uint64_t)keys_.size());
        const size_t pk_words = 2 * K_ * n_;
        for (const DeviceArray<uint64_t> &k : keys_) {
            if (!k.size()) {
                detail::put(os, (uint64_t)0);
                continue;
            }
            if (K_ < 2 || k.size() != (K_ - 1) * pk_words)
                throw std::logic_error("key shape is not set or does not match the key");
            detail::put(os, (uint64_t)(K_ - 1));
            detail::CtHeader hd;
            hd.id = parms_id_;
            hd.ntt = true;
            hd.size = 2;
            hd.n = n_;
            hd.L = K_;
            for (size_t j = 0; j + 1 < K_; j++) {
                detail::put_header(os, hd);
                detail::put(os, (uint64_t)pk_words);
                detail::write_device_words(os, k.get() + j * pk_words, pk_words);
            }
        }
    }
    void load(std::istream &is) {
        const char *what = "stream ended before the keys did";
        const ParmsID id = detail::get<ParmsID>(is, what);
        const uint64_t slots = detail::get<uint64_t>(is, what);
        if (slots > (1u << 18)) throw std::invalid_argument("key slot count is not valid");
        std::vector<DeviceArray<uint64_t>> keys(slots);
        size_t n = 0, K = 0;
        for (uint64_t s = 0; s < slots; s++) {
            const uint64_t m = detail::get<uint64_t>(is, what);
            if (!m) continue;
            for (uint64_t j = 0; j < m; j++) {
                const detail::CtHeader hd = detail::get_header(is);
                if (hd.seed != 0 || hd.terms || hd.size != 2 || !hd.ntt || hd.id != id)
                    throw std::invalid_argument("key part is not a size-2 NTT-form ciphertext of the key level");
                if (!n) {
                    n = (size_t)hd.n;
                    K = (size_t)hd.L;
                    if (!n || (n & (n - 1)) || n > (1u << 17) || K < 2 || K > 64)
                        throw std::invalid_argument("key part header is not valid");
                }
                if (hd.n != n || hd.L != K || m != K - 1)
                    throw std::invalid_argument("key parts differ in shape or count");
                const uint64_t words = detail::get<uint64_t>(is, what);
                if (words != 2 * K * n) throw std::invalid_argument("key part word count does not match its shape");
                if (j == 0) keys[s].resize(m * words);
                detail::read_device_words(is, keys[s].get() + j * words, words, what);
            }
        }
        keys_ = std::move(keys);
        parms_id_ = id;
        n_ = n;
        K_ = K;
    }

protected:
    std::vector<DeviceArray<uint64_t>> keys_;
    ParmsID parms_id_ = parmsIDZero;
    size_t n_ = 0, K_ = 0;
};

// RelinKeysCuda (src/relinkeys_cuda.cuh:11-60): key for s^p at index p - 2.
class RelinKeys : public KSwitchKeys {
public:
    static size_t getIndex(size_t key_power) {
        if (key_power < 2) throw std::invalid_argument("key_power cannot be less than 2");
        return key_power - 2;
    }
    bool hasKey(size_t key_power) const {
        const size_t i = getIndex(key_power);
        return keys_.size() > i && keys_[i].size();
    }
    const DeviceArray<uint64_t> &key(size_t key_power) const { return data(getIndex(key_power)); }
};

// GaloisKeysCuda (src/galoiskeys_cuda.cuh:34-80): key of Galois element e at index (e - 1) / 2.
class GaloisKeys : public KSwitchKeys {
public:
    static size_t getIndex(uint32_t galois_elt) {
        if (!(galois_elt & 1)) throw std::invalid_argument("galois_elt is not valid");
        return (galois_elt - 1) >> 1;
    }
    bool hasKey(uint32_t galois_elt) const {
        const size_t i = getIndex(galois_elt);
        return keys_.size() > i && keys_[i].size();
    }
    const DeviceArray<uint64_t> &key(uint32_t galois_elt) const {
        if (!hasKey(galois_elt)) throw std::invalid_argument("Galois key not present");
        return keys_[getIndex(galois_elt)];
    }
};

namespace detail {
inline uint32_t elt_from_step(size_t n, int step) {
    uint32_t elt = 0;
    check(shc_galois_elt_from_step((uint32_t)n, step, &elt));
    return elt;
}
}  // namespace detail

// ---------------------------------------------------------------------- key generator

// KeyGeneratorCuda (src/keygenerator_cuda.cuh:30-300).  Keys are generated on the
// device from the context's ChaCha20 sampler stream (a 256-bit key from
// std::random_device); every key takes a fresh 64-bit nonce from std::random_device.
class KeyGenerator {
public:
    explicit KeyGenerator(const SEALContext &ctx) : ctx_(ctx) {
        ctx_.checkSet();
        Plaintext &sk = sk_.data();
        sk.dynArray().resize(ctx_.keyCount() * ctx_.polyModulusDegree());
        check(shc_keygen_secret(ctx_.handle(), detail::os_random64(), sk.data(), nullptr));
        sk.parmsID() = ctx_.keyParmsID();
    }
    KeyGenerator(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk) {
        ctx_.checkSet();
        if (sk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("secret key is not valid for encryption parameters");
    }
    const SecretKey &secretKey() const { return sk_; }

    void createPublicKey(PublicKey &pk) const {
        pk.data().resize(ctx_, ctx_.keyParmsID(), 2);
        pk.data().isNttForm() = true;
        check(shc_keygen_public(ctx_.handle(), sk_.data().data(), detail::os_random64(), pk.data().data(), nullptr));
    }
    PublicKey createPublicKey() const {
        PublicKey pk;
        createPublicKey(pk);
        return pk;
    }
    void createRelinKeys(RelinKeys &rk) const {
        requireKeyswitching();
        rk.data().resize(1);
        rk.data()[0].resize(keyWords());
        size_t ws = 0;
        check(shc_keygen_workspace_size(ctx_.handle(), &ws));
        check(shc_keygen_relin(ctx_.handle(), sk_.data().data(), detail::os_random64(), rk.data()[0].get(),
                               ctx_.workspace(ws), ws, nullptr));
        rk.parmsID() = ctx_.keyParmsID();
        rk.setShape(ctx_.polyModulusDegree(), ctx_.keyCount());
    }
    RelinKeys createRelinKeys() const {
        RelinKeys rk;
        createRelinKeys(rk);
        return rk;
    }
    void createGaloisKeys(const std::vector<uint32_t> &elts, GaloisKeys &gk) const {
        requireKeyswitching();
        size_t ws = 0;
        check(shc_keygen_workspace_size(ctx_.handle(), &ws));
        if (gk.data().size() < ctx_.polyModulusDegree()) gk.data().resize(ctx_.polyModulusDegree());
        for (uint32_t e : elts) {
            auto &k = gk.data()[GaloisKeys::getIndex(e)];
            k.resize(keyWords());
            check(shc_keygen_galois(ctx_.handle(), sk_.data().data(), e, detail::os_random64(), k.get(),
                                    ctx_.workspace(ws), ws, nullptr));
        }
        gk.parmsID() = ctx_.keyParmsID();
        gk.setShape(ctx_.polyModulusDegree(), ctx_.keyCount());
    }
    GaloisKeys createGaloisKeys(const std::vector<uint32_t> &elts) const {
        GaloisKeys gk;
        createGaloisKeys(elts, gk);
        return gk;
    }
    // Keys for the given rotation steps (0 means the column swap / conjugation).
    void createGaloisKeys(const std::vector<int> &steps, GaloisKeys &gk) const {
        std::vector<uint32_t> elts;
        for (int s : steps) elts.push_back(detail::elt_from_step(ctx_.polyModulusDegree(), s));
        createGaloisKeys(elts, gk);
    }
    GaloisKeys createGaloisKeys(const std::vector<int> &steps) const {
        GaloisKeys gk;
        createGaloisKeys(steps, gk);
        return gk;
    }
    // GaloisTool::getEltsAll (src/utils/galois.cpp:101-127): 2N-1 and 3^(+-2^i).
    void createGaloisKeys(GaloisKeys &gk) const { createGaloisKeys(eltsAll(), gk); }
    GaloisKeys createGaloisKeys() const {
        GaloisKeys gk;
        createGaloisKeys(gk);
        return gk;
    }

    std::vector<uint32_t> eltsAll() const {
        const uint64_t m = 2 * ctx_.polyModulusDegree();
        std::vector<uint32_t> elts{(uint32_t)(m - 1)};
        uint64_t pos = 3, neg = 0;
        for (uint64_t x = 1; x < m; x += 2)
            if ((3 * x) % m == 1) neg = x;
        int logn = 0;
        while ((size_t(1) << logn) < ctx_.polyModulusDegree()) logn++;
        for (int i = 0; i + 1 < logn; i++) {
            elts.push_back((uint32_t)pos);
            elts.push_back((uint32_t)neg);
            pos = pos * pos % m;
            neg = neg * neg % m;
        }
        return elts;
    }

private:
    size_t keyWords() const {
        return (ctx_.keyCount() - 1) * 2 * ctx_.keyCount() * ctx_.polyModulusDegree();
    }
    void requireKeyswitching() const {
        if (!ctx_.usingKeyswitching()) throw std::logic_error("keyswitching is not supported by the context");
    }
    SEALContext ctx_;
    SecretKey sk_;
};

// ------------------------------------------------------------------------- encoders

// BatchEncoderCuda (src/batchencoder_cuda.cuh:55-95).
class BatchEncoder {
public:
    explicit BatchEncoder(const SEALContext &ctx) : ctx_(ctx) {
        ctx_.checkSet();
        if (ctx_.scheme() != SchemeType::bfv) throw std::invalid_argument("unsupported scheme");
    }
    size_t slotCount() const { return ctx_.polyModulusDegree(); }
    void encode(const std::vector<uint64_t> &values, Plaintext &plain) const {
        const size_t n = slotCount();
        if (values.size() > n) throw std::invalid_argument("values has size larger than slot_count");
        const uint64_t t = ctx_.parms().plainModulus().value();
        for (uint64_t v : values)
            if (v >= t) throw std::invalid_argument("input value is larger than plain_modulus");
        DeviceArray<uint64_t> dv;
        dv.fromHost(values.data(), values.size());
        plain.parmsID() = parmsIDZero;
        plain.resize(n);
        check(shc_batch_encode(ctx_.handle(), dv.get(), (uint32_t)values.size(), plain.data(), nullptr));
    }
    void encode(const std::vector<int64_t> &values, Plaintext &plain) const {
        const int64_t t = (int64_t)ctx_.parms().plainModulus().value();
        std::vector<uint64_t> u(values.size());
        for (size_t i = 0; i < values.size(); i++) {
            const int64_t v = values[i];
            if (v >= (t + 1) / 2 || v < -(t - 1) / 2) throw std::invalid_argument("input value out of range");
            u[i] = v < 0 ? (uint64_t)(v + t) : (uint64_t)v;
        }
        encode(u, plain);
    }
    void decode(const Plaintext &plain, std::vector<uint64_t> &out) const {
        if (plain.isNttForm()) throw std::invalid_argument("plain cannot be in NTT form");
        const size_t n = slotCount();
        DeviceArray<uint64_t> full, dv(n), scratch(n);
        const uint64_t *p = plain.data();
        if (plain.coeffCount() != n) {
            if (plain.coeffCount() > n) throw std::invalid_argument("plain is not valid for encryption parameters");
            full.resize(n);
            full.setZero();
            if (plain.coeffCount())
                hip_check(hipMemcpy(full.get(), plain.data(), plain.coeffCount() * 8, hipMemcpyDeviceToDevice), "copy");
            p = full.get();
        }
        check(shc_batch_decode(ctx_.handle(), p, dv.get(), scratch.get(), nullptr));
        out = dv.toHost();
    }
    void decode(const Plaintext &plain, std::vector<int64_t> &out) const {
        std::vector<uint64_t> u;
        decode(plain, u);
        const uint64_t t = ctx_.parms().plainModulus().value();
        out.resize(u.size());
        for (size_t i = 0; i < u.size(); i++)
            out[i] = u[i] >= (t + 1) / 2 ? (int64_t)u[i] - (int64_t)t : (int64_t)u[i];
    }
    // BatchEncoderCuda::encodePolynomial / decodePolynomial (src/batchencoder_cuda.cu:124-170,
    // 267-285): the values are the plaintext's coefficients, reduced mod t, with no slot
    // permutation.  As in the reference, the unsigned form keeps values.size()
    // coefficients and the signed form pads to N, a negative v becoming t - (|v| mod t).
    void encodePolynomial(const std::vector<uint64_t> &values, Plaintext &plain) const {
        if (values.size() > slotCount()) throw std::invalid_argument("values_matrix size is too large");
        const uint64_t t = ctx_.parms().plainModulus().value();
        std::vector<uint64_t> v(values.size());
        for (size_t i = 0; i < v.size(); i++) v[i] = values[i] % t;
        plain.parmsID() = parmsIDZero;
        plain.dynArray().fromHost(v.data(), v.size());
    }
    void encodePolynomial(const std::vector<int64_t> &values, Plaintext &plain) const {
        if (values.size() > slotCount()) throw std::invalid_argument("values_matrix size is too large");
        const uint64_t t = ctx_.parms().plainModulus().value();
        std::vector<uint64_t> v(slotCount(), 0);
        for (size_t i = 0; i < values.size(); i++) {
            const int64_t x = values[i];
            // |x| without overflow at INT64_MIN
            v[i] = x < 0 ? t - ((uint64_t)(-(x + 1)) + 1) % t : (uint64_t)x % t;
        }
        plain.parmsID() = parmsIDZero;
        plain.dynArray().fromHost(v.data(), v.size());
    }
    // The first N coefficients (zero past a shorter plaintext).
    void decodePolynomial(const Plaintext &plain, std::vector<uint64_t> &out) const {
        out = plain.dynArray().toHost();
        out.resize(slotCount(), 0);
    }
    // Centred: a coefficient above t / 2 (rounded down) becomes negative.
    void decodePolynomial(const Plaintext &plain, std::vector<int64_t> &out) const {
        std::vector<uint64_t> u;
        decodePolynomial(plain, u);
        const uint64_t t = ctx_.parms().plainModulus().value();
        out.resize(u.size());
        for (size_t i = 0; i < u.size(); i++) out[i] = u[i] > (t >> 1) ? (int64_t)(u[i] - t) : (int64_t)u[i];
    }

private:
    SEALContext ctx_;
};

// CKKSEncoderCuda (src/ckks_cuda.cuh:15-110).
class CKKSEncoder {
public:
    explicit CKKSEncoder(const SEALContext &ctx) : ctx_(ctx) {
        ctx_.checkSet();
        if (ctx_.scheme() != SchemeType::ckks) throw std::invalid_argument("unsupported scheme");
    }
    size_t slotCount() const { return ctx_.polyModulusDegree() / 2; }
    void encode(const std::vector<std::complex<double>> &values, ParmsID id, double scale, Plaintext &plain) const {
        if (values.size() > slotCount()) throw std::invalid_argument("values has invalid size");
        const uint32_t L = ctx_.levelOrThrow(id, "parms_id");
        if (scale <= 0 || (int)std::log2(scale) >= ctx_.levelBits(L)) throw std::invalid_argument("scale out of bounds");
        std::vector<double> flat(2 * values.size());
        for (size_t i = 0; i < values.size(); i++) {
            flat[2 * i] = values[i].real();
            flat[2 * i + 1] = values[i].imag();
        }
        DeviceArray<double> dv;
        dv.fromHost(flat.data(), flat.size());
        plain.parmsID() = parmsIDZero;
        plain.resize(L * ctx_.polyModulusDegree());
        size_t ws = 0;
        check(shc_ckks_encode_workspace_size(ctx_.handle(), &ws));
        check(shc_ckks_encode(ctx_.handle(), dv.get(), (uint32_t)values.size(), scale, L, plain.data(),
                              ctx_.workspace(ws), ws, nullptr));
        plain.parmsID() = id;
        plain.scale() = scale;
    }
    void encode(const std::vector<std::complex<double>> &values, double scale, Plaintext &plain) const {
        encode(values, ctx_.firstParmsID(), scale, plain);
    }
    void encode(const std::vector<double> &values, ParmsID id, double scale, Plaintext &plain) const {
        std::vector<std::complex<double>> c(values.begin(), values.end());
        encode(c, id, scale, plain);
    }
    void encode(const std::vector<double> &values, double scale, Plaintext &plain) const {
        encode(values, ctx_.firstParmsID(), scale, plain);
    }
    // A constant in every slot.
    void encode(std::complex<double> value, ParmsID id, double scale, Plaintext &plain) const {
        encode(std::vector<std::complex<double>>(slotCount(), value), id, scale, plain);
    }
    void encode(double value, ParmsID id, double scale, Plaintext &plain) const {
        encode(std::complex<double>(value, 0.0), id, scale, plain);
    }
    void encode(double value, double scale, Plaintext &plain) const {
        encode(value, ctx_.firstParmsID(), scale, plain);
    }
    void decode(const Plaintext &plain, std::vector<std::complex<double>> &out) const {
        if (!plain.isNttForm()) throw std::invalid_argument("plain is not in NTT form");
        const uint32_t L = ctx_.levelOrThrow(plain.parmsID(), "plain");
        const size_t n = ctx_.polyModulusDegree();
        DeviceArray<double> dv(n);
        size_t ws = 0;
        check(shc_ckks_decode_workspace_size(ctx_.handle(), L, &ws));
        check(shc_ckks_decode(ctx_.handle(), plain.data(), L, plain.scale(), dv.get(), ctx_.workspace(ws), ws,
                              nullptr));
        const std::vector<double> h = dv.toHost();
        out.resize(slotCount());
        for (size_t i = 0; i < out.size(); i++) out[i] = {h[2 * i], h[2 * i + 1]};
    }
    // CKKSEncoderCuda::encodePolynomial / decodePolynomial (src/ckks_cuda.cuh:34-40, 84-87):
    // the values are the coefficients themselves (times scale, rounded half away from
    // zero), no FFT; decode returns all N coefficients over the plaintext's scale.
    void encodePolynomial(const std::vector<double> &values, ParmsID id, double scale, Plaintext &plain) const {
        const uint32_t L = ctx_.levelOrThrow(id, "parms_id");
        if (values.size() > ctx_.polyModulusDegree()) throw std::invalid_argument("values_size is too large");
        DeviceArray<double> dv;
        dv.fromHost(values.data(), values.size());
        plain.parmsID() = parmsIDZero;
        plain.resize(L * ctx_.polyModulusDegree());
        size_t ws = 0;
        check(shc_ckks_encode_workspace_size(ctx_.handle(), &ws));
        check(shc_ckks_encode_polynomial(ctx_.handle(), dv.get(), (uint32_t)values.size(), scale, L, plain.data(),
                                         ctx_.workspace(ws), ws, nullptr));
        plain.parmsID() = id;
        plain.scale() = scale;
    }
    void encodePolynomial(const std::vector<double> &values, double scale, Plaintext &plain) const {
        encodePolynomial(values, ctx_.firstParmsID(), scale, plain);
    }
    void decodePolynomial(const Plaintext &plain, std::vector<double> &out) const {
        if (!plain.isNttForm()) throw std::invalid_argument("plain is not in NTT form");
        const uint32_t L = ctx_.levelOrThrow(plain.parmsID(), "plain");
        DeviceArray<double> dv(ctx_.polyModulusDegree());
        size_t ws = 0;
        check(shc_ckks_decode_workspace_size(ctx_.handle(), L, &ws));
        check(shc_ckks_decode_polynomial(ctx_.handle(), plain.data(), L, plain.scale(), dv.get(), ctx_.workspace(ws),
                                         ws, nullptr));
        out = dv.toHost();
    }

private:
    SEALContext ctx_;
};

namespace detail {
// A BFV plaintext of fewer than N coefficients, zero-padded to N into `tmp` (the C ABI
// takes N coefficients); the plaintext's own words when it already has N.
inline const uint64_t *padded_plain(const SEALContext &ctx, const Plaintext &p, DeviceArray<uint64_t> &tmp) {
    const size_t n = ctx.polyModulusDegree();
    if (p.coeffCount() == n) return p.data();
    if (p.coeffCount() > n) throw std::invalid_argument("plain is not valid for encryption parameters");
    tmp.resize(n);
    tmp.setZero();
    if (p.coeffCount()) hip_check(hipMemcpy(tmp.get(), p.data(), p.coeffCount() * 8, hipMemcpyDeviceToDevice), "copy");
    return tmp.get();
}
}  // namespace detail

// ------------------------------------------------------------------- encrypt/decrypt

// EncryptorCuda (src/encryptor_cuda.cuh:50-400).  Randomness: the context's ChaCha20
// sampler stream, a fresh 64-bit nonce from std::random_device per encryption (the
// reference uses curand).
class Encryptor {
public:
    Encryptor(const SEALContext &ctx, const PublicKey &pk) : ctx_(ctx), pk_(pk) {
        ctx_.checkSet();
        has_pk_ = true;
    }
    Encryptor(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk) {
        ctx_.checkSet();
        has_sk_ = true;
    }
    Encryptor(const SEALContext &ctx, const PublicKey &pk, const SecretKey &sk) : ctx_(ctx), pk_(pk), sk_(sk) {
        ctx_.checkSet();
        has_pk_ = has_sk_ = true;
    }
    // EncryptorCuda::setPublicKey / setSecretKey (src/encryptor_cuda.cuh:87-101).
    void setPublicKey(const PublicKey &pk) {
        if (pk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("public key is not valid for encryption parameters");
        pk_ = pk;
        has_pk_ = true;
    }
    void setSecretKey(const SecretKey &sk) {
        if (sk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("secret key is not valid for encryption parameters");
        sk_ = sk;
        has_sk_ = true;
    }

    void encrypt(const Plaintext &plain, Ciphertext &ct) const {
        if (!has_pk_) throw std::logic_error("public key is not set");
        const ParmsID id = levelOf(plain);
        const uint32_t L = ctx_.levelOf(id);
        DeviceArray<uint64_t> tmp;
        const uint64_t *p = ctx_.scheme() == SchemeType::ckks ? plain.data() : detail::padded_plain(ctx_, plain, tmp);
        ct.resize(ctx_, id, 2);
        ct.isNttForm() = ctx_.scheme() == SchemeType::ckks;
        ct.scale() = ctx_.scheme() == SchemeType::ckks ? plain.scale() : 1.0;
        size_t ws = 0;
        check(shc_encrypt_workspace_size(ctx_.handle(), L, &ws));
        check(shc_encrypt(ctx_.handle(), pk_.data().data(), p, L, detail::os_random64(), ct.data(), ctx_.workspace(ws),
                          ws, nullptr));
    }
    Ciphertext encrypt(const Plaintext &plain) const {
        Ciphertext ct;
        encrypt(plain, ct);
        return ct;
    }
    // Symmetric encryption: an encryption of zero under the secret key plus the plaintext.
    void encryptSymmetric(const Plaintext &plain, Ciphertext &ct) const {
        if (!has_sk_) throw std::logic_error("secret key is not set");
        const ParmsID id = levelOf(plain);
        const uint32_t L = ctx_.levelOf(id);
        const bool ckks = ctx_.scheme() == SchemeType::ckks;
        DeviceArray<uint64_t> tmp;
        const uint64_t *p = ckks ? plain.data() : detail::padded_plain(ctx_, plain, tmp);
        ct.resize(ctx_, id, 2);
        ct.isNttForm() = ckks;
        ct.scale() = ckks ? plain.scale() : 1.0;
        check(shc_encrypt_zero_symmetric(ctx_.handle(), sk_.data().data(), L, detail::os_random64(), ct.data(),
                                         nullptr));
        if (ckks)
            check(shc_add_plain(ctx_.handle(), ct.data(), p, ct.data(), 1, 2, L, 0, nullptr));
        else
            check(shc_bfv_add_plain(ctx_.handle(), ct.data(), p, ct.data(), 1, 2, L, 0, nullptr));
    }
    Ciphertext encryptSymmetric(const Plaintext &plain) const {
        Ciphertext ct;
        encryptSymmetric(plain, ct);
        return ct;
    }
    // EncryptorCuda::encryptZero (src/encryptor_cuda.cuh:170-236, encryptZeroInternal at
    // src/encryptor_cuda.cu:92-166): an encryption of zero at the level of `id` (below the
    // first level: one level above, then divide-and-round), in the scheme's default form.
    void encryptZero(ParmsID id, Ciphertext &ct) const {
        if (!has_pk_) throw std::logic_error("public key is not set");
        const uint32_t L = ctx_.levelOrThrow(id, "parms_id");
        ct.resize(ctx_, id, 2);
        ct.isNttForm() = ctx_.scheme() == SchemeType::ckks;
        ct.scale() = 1.0;
        size_t ws = 0;
        check(shc_encrypt_workspace_size(ctx_.handle(), L, &ws));
        check(shc_encrypt(ctx_.handle(), pk_.data().data(), nullptr, L, detail::os_random64(), ct.data(),
                          ctx_.workspace(ws), ws, nullptr));
    }
    void encryptZero(Ciphertext &ct) const { encryptZero(ctx_.firstParmsID(), ct); }
    Ciphertext encryptZero(ParmsID id) const {
        Ciphertext ct;
        encryptZero(id, ct);
        return ct;
    }
    Ciphertext encryptZero() const { return encryptZero(ctx_.firstParmsID()); }
    // EncryptorCuda::encryptZeroSymmetric (src/encryptor_cuda.cuh:312-388).
    void encryptZeroSymmetric(ParmsID id, Ciphertext &ct) const {
        if (!has_sk_) throw std::logic_error("secret key is not set");
        const uint32_t L = ctx_.levelOrThrow(id, "parms_id");
        ct.resize(ctx_, id, 2);
        ct.isNttForm() = ctx_.scheme() == SchemeType::ckks;
        ct.scale() = 1.0;
        check(shc_encrypt_zero_symmetric(ctx_.handle(), sk_.data().data(), L, detail::os_random64(), ct.data(),
                                         nullptr));
    }
    void encryptZeroSymmetric(Ciphertext &ct) const { encryptZeroSymmetric(ctx_.firstParmsID(), ct); }
    Ciphertext encryptZeroSymmetric(ParmsID id) const {
        Ciphertext ct;
        encryptZeroSymmetric(id, ct);
        return ct;
    }
    Ciphertext encryptZeroSymmetric() const { return encryptZeroSymmetric(ctx_.firstParmsID()); }

private:
    ParmsID levelOf(const Plaintext &plain) const {
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!plain.isNttForm()) throw std::invalid_argument("plain must be in NTT form");
            ctx_.levelOrThrow(plain.parmsID(), "plain");
            return plain.parmsID();
        }
        if (plain.isNttForm()) throw std::invalid_argument("plain cannot be in NTT form");
        return ctx_.firstParmsID();
    }
    SEALContext ctx_;
    PublicKey pk_;
    SecretKey sk_;
    bool has_pk_ = false, has_sk_ = false;
};

// DecryptorCuda (src/decryptor_cuda.cuh:40-90).
class Decryptor {
public:
    Decryptor(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk) { ctx_.checkSet(); }
    void decrypt(const Ciphertext &ct, Plaintext &plain) {
        if (ct.size() < 2 || ct.size() > 3) throw std::invalid_argument("ciphertext size must be 2 or 3");
        const uint32_t L = ctx_.levelOrThrow(ct.parmsID(), "encrypted");
        const bool ckks = ctx_.scheme() == SchemeType::ckks;
        if (ckks != ct.isNttForm()) throw std::invalid_argument("ciphertext is in the wrong form");
        size_t ws = 0;
        check(shc_decrypt_workspace_size(ctx_.handle(), L, &ws));
        plain.parmsID() = parmsIDZero;
        plain.resize(ckks ? (size_t)L * ctx_.polyModulusDegree() : ctx_.polyModulusDegree());
        check(shc_decrypt(ctx_.handle(), ct.data(), (uint32_t)ct.size(), sk_.data().data(), L, plain.data(),
                          ctx_.workspace(ws), ws, nullptr));
        plain.parmsID() = ckks ? ct.parmsID() : parmsIDZero;
        plain.scale() = ct.scale();
    }

private:
    SEALContext ctx_;
    SecretKey sk_;
};

// ------------------------------------------------------------------------ evaluator

// EvaluatorCuda (src/evaluator_cuda.cuh:15-420).
class Evaluator {
public:
    explicit Evaluator(const SEALContext &ctx) : ctx_(ctx) { ctx_.checkSet(); }

    // ---- additive
    void negateInplace(Ciphertext &a) const {
        check(shc_negate(ctx_.handle(), a.data(), a.data(), (uint32_t)a.size(), level(a), nullptr));
    }
    void negate(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        negateInplace(out);
    }
    void addInplace(Ciphertext &a, const Ciphertext &b) const { addSub(a, b, false); }
    void add(const Ciphertext &a, const Ciphertext &b, Ciphertext &out) const {
        Ciphertext r = a;   // out may alias a or b
        addInplace(r, b);
        out = std::move(r);
    }
    void addMany(const std::vector<Ciphertext> &cts, Ciphertext &out) const {
        if (cts.empty()) throw std::invalid_argument("encrypteds cannot be empty");
        out = cts[0];
        for (size_t i = 1; i < cts.size(); i++) addInplace(out, cts[i]);
    }
    void subInplace(Ciphertext &a, const Ciphertext &b) const { addSub(a, b, true); }
    void sub(const Ciphertext &a, const Ciphertext &b, Ciphertext &out) const {
        Ciphertext r = a;
        subInplace(r, b);
        out = std::move(r);
    }

    // ---- multiplicative
    void multiplyInplace(Ciphertext &a, const Ciphertext &b) const {
        sameParms(a, b);
        if (a.size() != 2 || b.size() != 2) throw std::invalid_argument("only size-2 ciphertexts can be multiplied");
        const uint32_t L = level(a);
        Ciphertext out;
        out.resize(ctx_, a.parmsID(), 3);
        out.isNttForm() = a.isNttForm();
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!a.isNttForm() || !b.isNttForm()) throw std::invalid_argument("encrypted must be in NTT form");
            out.scale() = a.scale() * b.scale();
            checkScale(out.scale(), L);
            check(shc_ckks_multiply(ctx_.handle(), a.data(), b.data(), out.data(), 1, L, nullptr));
        } else {
            if (a.isNttForm() || b.isNttForm()) throw std::invalid_argument("encrypted cannot be in NTT form");
            size_t ws = 0;
            check(shc_multiply_workspace_size(ctx_.handle(), 1, L, &ws));
            check(shc_multiply(ctx_.handle(), a.data(), b.data(), out.data(), 1, L, ctx_.workspace(ws), ws, nullptr));
        }
        a = std::move(out);
    }
    void multiply(const Ciphertext &a, const Ciphertext &b, Ciphertext &out) const {
        Ciphertext r = a;
        multiplyInplace(r, b);
        out = std::move(r);
    }
    void squareInplace(Ciphertext &a) const {
        const Ciphertext b = a;
        multiplyInplace(a, b);
    }
    void square(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        squareInplace(out);
    }
    void relinearizeInplace(Ciphertext &a, const RelinKeys &rk) const {
        if (rk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("relin_keys is not valid for encryption parameters");
        if (a.size() == 2) return;
        if (a.size() != 3) throw std::invalid_argument("destination_size must be at least 2");
        const uint32_t L = level(a);
        size_t ws = 0;
        check(shc_keyswitch_workspace_size(ctx_.handle(), 1, L, &ws));
        check(shc_relinearize(ctx_.handle(), a.data(), keyWords(rk.key(2)), 1, L, ctx_.workspace(ws), ws, nullptr));
        a.shrink(2);
    }
    void relinearize(const Ciphertext &a, const RelinKeys &rk, Ciphertext &out) const {
        out = a;
        relinearizeInplace(out, rk);
    }

    // ---- modulus switching
    void rescaleToNext(const Ciphertext &a, Ciphertext &out) const {
        if (ctx_.scheme() != SchemeType::ckks) throw std::invalid_argument("unsupported operation for scheme type");
        const double scale = a.scale() / (double)ctx_.prime(level(a) - 1);
        divideRound(a, out);
        out.scale() = scale;
    }
    void rescaleToNextInplace(Ciphertext &a) const {
        Ciphertext out;
        rescaleToNext(a, out);
        a = std::move(out);
    }
    void rescaleToInplace(Ciphertext &a, ParmsID id) const {
        const uint32_t target = targetLevel(a, id);
        while (level(a) > target) rescaleToNextInplace(a);
    }
    void rescaleTo(const Ciphertext &a, ParmsID id, Ciphertext &out) const {
        Ciphertext r = a;
        rescaleToInplace(r, id);
        out = std::move(r);
    }
    void modSwitchToNext(const Ciphertext &a, Ciphertext &out) const {
        if (ctx_.scheme() == SchemeType::bfv) {
            const double scale = a.scale();
            divideRound(a, out);
            out.scale() = scale;
            return;
        }
        // CKKS: drop the last limb of every polynomial (modSwitchDropToNext)
        const uint32_t L = nextLevel(a);
        Ciphertext r;
        r.resize(ctx_, ctx_.idOf(L), a.size());
        r.isNttForm() = a.isNttForm();
        r.scale() = a.scale();
        const size_t n = ctx_.polyModulusDegree();
        hip_check(hipMemcpy2D(r.data(), L * n * 8, a.data(), (L + 1) * n * 8, L * n * 8, a.size(),
                              hipMemcpyDeviceToDevice),
                  "modswitch copy");
        out = std::move(r);
    }
    void modSwitchToNextInplace(Ciphertext &a) const {
        Ciphertext out;
        modSwitchToNext(a, out);
        a = std::move(out);
    }
    void modSwitchToInplace(Ciphertext &a, ParmsID id) const {
        const uint32_t target = targetLevel(a, id);
        while (level(a) > target) modSwitchToNextInplace(a);
    }
    // Plaintexts in NTT form (CKKS) drop their last limb.
    void modSwitchToNextInplace(Plaintext &p) const {
        if (!p.isNttForm()) throw std::invalid_argument("plain is not in NTT form");
        const uint32_t L0 = ctx_.levelOrThrow(p.parmsID(), "plain");
        if (L0 < 2 || L0 == ctx_.keyCount()) throw std::invalid_argument("end of modulus switching chain reached");
        const uint32_t L = L0 - 1;
        DeviceArray<uint64_t> d(L * ctx_.polyModulusDegree());
        hip_check(hipMemcpy(d.get(), p.data(), d.size() * 8, hipMemcpyDeviceToDevice), "copy");
        p.dynArray() = std::move(d);
        p.parmsID() = ctx_.idOf(L);
    }
    void modSwitchToInplace(Plaintext &p, ParmsID id) const {
        const uint32_t target = ctx_.levelOrThrow(id, "parms_id");
        if (ctx_.levelOrThrow(p.parmsID(), "plain") < target)
            throw std::invalid_argument("cannot switch to higher level modulus");
        while (ctx_.levelOf(p.parmsID()) > target) modSwitchToNextInplace(p);
    }
    void modSwitchTo(const Ciphertext &a, ParmsID id, Ciphertext &out) const {
        Ciphertext r = a;
        modSwitchToInplace(r, id);
        out = std::move(r);
    }
    void modSwitchTo(const Plaintext &p, ParmsID id, Plaintext &out) const {
        Plaintext r = p;
        modSwitchToInplace(r, id);
        out = std::move(r);
    }

    // EvaluatorCuda::multiplyMany (src/evaluator_cuda.cu:1434-1510), BFV: the same product
    // tree -- neighbours (0,1), (2,3), ... multiplied and relinearized, an odd last one
    // carried, then the front two of the list multiplied onto its back until one is left
    // -- so the result has the same words as the reference's.
    void multiplyMany(const std::vector<Ciphertext> &cts, const RelinKeys &rk, Ciphertext &out) const {
        if (cts.empty()) throw std::invalid_argument("encrypteds vector must not be empty");
        for (const Ciphertext &c : cts)
            if (&c == &out) throw std::invalid_argument("encrypteds must be different from destination");
        ctx_.levelOrThrow(cts[0].parmsID(), "encrypteds");
        if (ctx_.scheme() != SchemeType::bfv) throw std::logic_error("unsupported scheme");
        if (cts.size() == 1) {
            out = cts[0];
            return;
        }
        std::vector<Ciphertext> prod;
        prod.reserve(2 * cts.size());
        for (size_t i = 0; i + 1 < cts.size(); i += 2) {
            Ciphertext t;
            multiply(cts[i], cts[i + 1], t);
            relinearizeInplace(t, rk);
            prod.push_back(std::move(t));
        }
        if (cts.size() & 1) prod.push_back(cts.back());
        for (size_t i = 0; i + 1 < prod.size(); i += 2) {
            Ciphertext t;
            multiply(prod[i], prod[i + 1], t);
            relinearizeInplace(t, rk);
            prod.push_back(std::move(t));
        }
        out = std::move(prod.back());
    }
    // EvaluatorCuda::exponentiateInplace (src/evaluator_cuda.cu:1512-1540): the product of
    // `exponent` copies through multiplyMany.
    void exponentiateInplace(Ciphertext &a, uint64_t exponent, const RelinKeys &rk) const {
        level(a);
        if (rk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("relin_keys is not valid for encryption parameters");
        if (exponent == 0) throw std::invalid_argument("exponent cannot be 0");
        if (exponent == 1) return;
        const std::vector<Ciphertext> copies((size_t)exponent, a);
        Ciphertext r;
        multiplyMany(copies, rk, r);
        a = std::move(r);
    }
    void exponentiate(const Ciphertext &a, uint64_t exponent, const RelinKeys &rk, Ciphertext &out) const {
        Ciphertext r = a;
        exponentiateInplace(r, exponent, rk);
        out = std::move(r);
    }

    // ---- plaintext operations
    void addPlainInplace(Ciphertext &a, const Plaintext &p) const { addSubPlain(a, p, false); }
    void addPlain(const Ciphertext &a, const Plaintext &p, Ciphertext &out) const {
        out = a;
        addPlainInplace(out, p);
    }
    void subPlainInplace(Ciphertext &a, const Plaintext &p) const { addSubPlain(a, p, true); }
    void subPlain(const Ciphertext &a, const Plaintext &p, Ciphertext &out) const {
        out = a;
        subPlainInplace(out, p);
    }
    void multiplyPlainInplace(Ciphertext &a, const Plaintext &p) const {
        const uint32_t L = level(a);
        if (a.isNttForm() != p.isNttForm()) throw std::invalid_argument("NTT form mismatch");
        if (a.isNttForm()) {
            if (a.parmsID() != p.parmsID()) throw std::invalid_argument("encrypted_ntt and plain_ntt parameter mismatch");
            check(shc_multiply_plain_ntt(ctx_.handle(), a.data(), p.data(), a.data(), (uint32_t)a.size(), L, nullptr));
        } else {
            DeviceArray<uint64_t> tmp;
            const uint64_t *pp = detail::padded_plain(ctx_, p, tmp);
            size_t ws = 0;
            check(shc_bfv_multiply_plain_workspace_size(ctx_.handle(), L, &ws));
            check(shc_bfv_multiply_plain(ctx_.handle(), a.data(), pp, a.data(), (uint32_t)a.size(), L,
                                         ctx_.workspace(ws), ws, nullptr));
        }
        if (ctx_.scheme() == SchemeType::ckks) {
            a.scale() *= p.scale();
            checkScale(a.scale(), L);
        }
    }
    void multiplyPlain(const Ciphertext &a, const Plaintext &p, Ciphertext &out) const {
        out = a;
        multiplyPlainInplace(out, p);
    }

    // ---- NTT form
    void transformToNttInplace(Ciphertext &a) const {
        if (a.isNttForm()) throw std::invalid_argument("encrypted is already in NTT form");
        check(shc_ntt_forward(ctx_.handle(), a.data(), (uint32_t)a.size(), level(a), nullptr));
        a.isNttForm() = true;
    }
    void transformToNtt(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        transformToNttInplace(out);
    }
    void transformFromNttInplace(Ciphertext &a) const {
        if (!a.isNttForm()) throw std::invalid_argument("encrypted is not in NTT form");
        check(shc_ntt_inverse(ctx_.handle(), a.data(), (uint32_t)a.size(), level(a), nullptr));
        a.isNttForm() = false;
    }
    void transformFromNtt(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        transformFromNttInplace(out);
    }
    // BFV plaintext -> lifted NTT form at level `id`.
    void transformToNttInplace(Plaintext &p, ParmsID id) const {
        if (p.isNttForm()) throw std::invalid_argument("plain is already in NTT form");
        const uint32_t L = ctx_.levelOrThrow(id, "parms_id");
        DeviceArray<uint64_t> tmp;
        const uint64_t *pp = detail::padded_plain(ctx_, p, tmp);
        DeviceArray<uint64_t> d(L * ctx_.polyModulusDegree());
        check(shc_bfv_plain_to_ntt(ctx_.handle(), pp, L, d.get(), nullptr));
        p.dynArray() = std::move(d);
        p.parmsID() = id;
    }

    // ---- Galois
    void applyGaloisInplace(Ciphertext &a, uint32_t elt, const GaloisKeys &gk) const {
        if (gk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("galois_keys is not valid for encryption parameters");
        if (a.size() != 2) throw std::invalid_argument("encrypted size must be 2");
        const uint32_t L = level(a);
        size_t ws = 0;
        check(shc_galois_workspace_size(ctx_.handle(), 1, L, &ws));
        check(shc_apply_galois_keyswitch(ctx_.handle(), a.data(), keyWords(gk.key(elt)), elt, 1, L, ctx_.workspace(ws), ws,
                                         nullptr));
    }
    void applyGalois(const Ciphertext &a, uint32_t elt, const GaloisKeys &gk, Ciphertext &out) const {
        out = a;
        applyGaloisInplace(out, elt, gk);
    }
    void rotateRowsInplace(Ciphertext &a, int steps, const GaloisKeys &gk) const {
        if (ctx_.scheme() != SchemeType::bfv) throw std::logic_error("unsupported scheme");
        rotateInternal(a, steps, gk);
    }
    void rotateRows(const Ciphertext &a, int steps, const GaloisKeys &gk, Ciphertext &out) const {
        out = a;
        rotateRowsInplace(out, steps, gk);
    }
    void rotateColumnsInplace(Ciphertext &a, const GaloisKeys &gk) const {
        if (ctx_.scheme() != SchemeType::bfv) throw std::logic_error("unsupported scheme");
        applyGaloisInplace(a, (uint32_t)(2 * ctx_.polyModulusDegree() - 1), gk);
    }
    void rotateColumns(const Ciphertext &a, const GaloisKeys &gk, Ciphertext &out) const {
        out = a;
        rotateColumnsInplace(out, gk);
    }
    void rotateVectorInplace(Ciphertext &a, int steps, const GaloisKeys &gk) const {
        if (ctx_.scheme() != SchemeType::ckks) throw std::logic_error("unsupported scheme");
        rotateInternal(a, steps, gk);
    }
    void rotateVector(const Ciphertext &a, int steps, const GaloisKeys &gk, Ciphertext &out) const {
        out = a;
        rotateVectorInplace(out, steps, gk);
    }
    void complexConjugateInplace(Ciphertext &a, const GaloisKeys &gk) const {
        if (ctx_.scheme() != SchemeType::ckks) throw std::logic_error("unsupported scheme");
        applyGaloisInplace(a, (uint32_t)(2 * ctx_.polyModulusDegree() - 1), gk);
    }
    void complexConjugate(const Ciphertext &a, const GaloisKeys &gk, Ciphertext &out) const {
        out = a;
        complexConjugateInplace(out, gk);
    }

private:
    uint32_t level(const Ciphertext &a) const { return ctx_.levelOrThrow(a.parmsID(), "encrypted"); }
    // A key's words, after checking that it has this context's [K-1][2][K][N] shape (a
    // loaded key carries its own shape).
    const uint64_t *keyWords(const DeviceArray<uint64_t> &k) const {
        const size_t K = ctx_.keyCount(), n = ctx_.polyModulusDegree();
        if (k.size() != (K - 1) * 2 * K * n) throw std::invalid_argument("keyswitching key has the wrong shape");
        return k.get();
    }
    uint32_t targetLevel(const Ciphertext &a, ParmsID id) const {
        const uint32_t target = ctx_.levelOrThrow(id, "parms_id");
        if (level(a) < target) throw std::invalid_argument("cannot switch to higher level modulus");
        return target;
    }
    void sameParms(const Ciphertext &a, const Ciphertext &b) const {
        if (a.parmsID() != b.parmsID()) throw std::invalid_argument("encrypted1 and encrypted2 parameter mismatch");
        if (a.isNttForm() != b.isNttForm()) throw std::invalid_argument("NTT form mismatch");
    }
    void checkScale(double scale, uint32_t L) const {
        if (!(scale > 0) || (int)std::log2(scale) >= ctx_.levelBits(L)) throw std::invalid_argument("scale out of bounds");
    }
    static bool scalesClose(double a, double b) {
        const double d = std::fabs(a - b);
        return d <= std::max(std::fabs(a), std::fabs(b)) * 1e-9;
    }
    void addSub(Ciphertext &a, const Ciphertext &b, bool sub) const {
        sameParms(a, b);
        if (ctx_.scheme() == SchemeType::ckks && !scalesClose(a.scale(), b.scale()))
            throw std::invalid_argument("scale mismatch");
        const uint32_t L = level(a);
        const size_t common = std::min(a.size(), b.size());
        if (b.size() > a.size()) {
            Ciphertext grown;
            grown.resize(ctx_, a.parmsID(), b.size());
            grown.isNttForm() = a.isNttForm();
            grown.scale() = a.scale();
            hip_check(hipMemcpy(grown.data(), a.data(), a.size() * a.polyWords() * 8, hipMemcpyDeviceToDevice), "copy");
            if (sub)
                check(shc_negate(ctx_.handle(), b.data(common), grown.data(common), (uint32_t)(b.size() - common), L,
                                 nullptr));
            else
                hip_check(hipMemcpy(grown.data(common), b.data(common), (b.size() - common) * b.polyWords() * 8,
                                    hipMemcpyDeviceToDevice),
                          "copy");
            a = std::move(grown);
        }
        if (sub) check(shc_sub(ctx_.handle(), a.data(), b.data(), a.data(), (uint32_t)common, L, nullptr));
        else check(shc_add(ctx_.handle(), a.data(), b.data(), a.data(), (uint32_t)common, L, nullptr));
    }
    void addSubPlain(Ciphertext &a, const Plaintext &p, bool sub) const {
        const uint32_t L = level(a);
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!p.isNttForm() || !a.isNttForm() || p.parmsID() != a.parmsID())
                throw std::invalid_argument("encrypted and plain parameter mismatch");
            if (!scalesClose(a.scale(), p.scale())) throw std::invalid_argument("scale mismatch");
            check(shc_add_plain(ctx_.handle(), a.data(), p.data(), a.data(), 1, (uint32_t)a.size(), L, sub, nullptr));
        } else {
            if (p.isNttForm() || a.isNttForm()) throw std::invalid_argument("NTT form mismatch");
            DeviceArray<uint64_t> tmp;
            const uint64_t *pp = detail::padded_plain(ctx_, p, tmp);
            check(shc_bfv_add_plain(ctx_.handle(), a.data(), pp, a.data(), 1, (uint32_t)a.size(), L, sub, nullptr));
        }
    }
    uint32_t nextLevel(const Ciphertext &a) const {
        auto cd = ctx_.getContextData(a.parmsID());
        if (!cd) throw std::invalid_argument("encrypted is not valid for encryption parameters");
        if (cd->nextParmsID() == parmsIDZero) throw std::invalid_argument("end of modulus switching chain reached");
        return level(a) - 1;
    }
    void divideRound(const Ciphertext &a, Ciphertext &out) const {
        const uint32_t L = nextLevel(a) + 1;
        Ciphertext r;   // out may alias a
        r.resize(ctx_, ctx_.idOf(L - 1), a.size());
        r.isNttForm() = a.isNttForm();
        size_t ws = 0;
        check(shc_rescale_workspace_size(ctx_.handle(), 1, (uint32_t)a.size(), L, &ws));
        check(shc_rescale_to_next(ctx_.handle(), a.data(), r.data(), 1, (uint32_t)a.size(), L, ctx_.workspace(ws),
                                  ws, nullptr));
        out = std::move(r);
    }
    // EvaluatorCuda::rotateInternal (src/evaluator_cuda.cu:2106-2163): the key of the
    // step itself, else the NAF terms one by one.
    void rotateInternal(Ciphertext &a, int steps, const GaloisKeys &gk) const {
        if (steps == 0) return;
        const size_t n = ctx_.polyModulusDegree();
        const uint32_t elt = detail::elt_from_step(n, steps);
        if (gk.hasKey(elt)) {
            applyGaloisInplace(a, elt, gk);
            return;
        }
        std::vector<int> naf;
        const bool neg = steps < 0;
        int v = std::abs(steps);
        for (int i = 0; v; i++) {
            const int z = (v & 1) ? 2 - (v & 3) : 0;
            v = (v - z) >> 1;
            if (z) naf.push_back((neg ? -z : z) * (1 << i));
        }
        if (naf.size() == 1) throw std::invalid_argument("Galois key not present");
        for (int s : naf)
            if ((size_t)std::abs(s) != (n >> 1)) rotateInternal(a, s, gk);
    }

    SEALContext ctx_;
};

inline void Ciphertext::saveTerms(std::ostream &os, const Evaluator &ev, const std::vector<size_t> &term_ids) const {
    Ciphertext coeff = *this;
    if (coeff.isNttForm()) ev.transformFromNttInplace(coeff);
    const std::vector<uint64_t> h = coeff.dynArray().toHost();
    saveHeader(os, true);
    for (size_t id : term_ids) {
        if (id >= n_) throw std::invalid_argument("term index out of range");
        for (size_t j = 0; j < L_; j++) detail::put(os, h[j * n_ + id]);
    }
    const uint64_t rest = (uint64_t)((size_ - 1) * polyWords());
    detail::put(os, rest);
    os.write(reinterpret_cast<const char *>(h.data() + polyWords()), (std::streamsize)(rest * 8));
}

inline void Ciphertext::loadTerms(std::istream &is, const Evaluator &ev, const SEALContext &ctx,
                                  const std::vector<size_t> &term_ids) {
    loadHeader(is, ctx, true);
    const char *what = "stream ended before the ciphertext did";
    std::vector<uint64_t> h(size_ * polyWords(), 0);
    for (size_t id : term_ids) {
        if (id >= n_) throw std::invalid_argument("term index out of range");
        for (size_t j = 0; j < L_; j++) h[j * n_ + id] = detail::get<uint64_t>(is, what);
    }
    const uint64_t rest = detail::get<uint64_t>(is, what);
    if (rest != (size_ - 1) * polyWords()) throw std::invalid_argument("ciphertext word count does not match its shape");
    is.read(reinterpret_cast<char *>(h.data() + polyWords()), (std::streamsize)(rest * 8));
    if (!is) throw std::invalid_argument(what);
    data_.fromHost(h.data(), h.size());
    if (ntt_form_) {
        ntt_form_ = false;
        ev.transformToNttInplace(*this);
    }
}

}  // namespace troyhip
