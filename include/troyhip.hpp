// troyhip.hpp -- the troy_cuda C++ class API over the MI355X C ABI (sealhip.h).
//
// Header only.  A program written against the reference's troyn:: aliases
// (src/troy_cuda.cuh:20-42) switches by using namespace troyhip instead: the same class
// names, method names, argument meanings and exception types.  Device buffers are
// owned by RAII arrays (the reference's DeviceArray, src/utils/devicearray.cuh:106-200);
// every operation runs on the default stream, as the reference's do, and returns when
// its work is queued (host-visible results -- decode, toHost -- synchronise).
//
// Link with libsealhip.so (and the HIP runtime).
#pragma once
#include "sealhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <istream>
#include <ostream>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <random>
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

// ---------------------------------------------------------------------- device array

template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    DeviceArray(const DeviceArray &o) { *this = o; }
    DeviceArray(DeviceArray &&o) noexcept : ptr_(o.ptr_), size_(o.size_) {
        o.ptr_ = nullptr;
        o.size_ = 0;
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
            o.ptr_ = nullptr;
            o.size_ = 0;
        }
        return *this;
    }
    ~DeviceArray() { release(); }

    // Reallocates (contents are not kept) when the size changes.
    void resize(size_t n) {
        if (n == size_) return;
        release();
        if (n) hip_check(hipMalloc(reinterpret_cast<void **>(&ptr_), n * sizeof(T)), "hipMalloc");
        size_ = n;
    }
    // Grows to at least n elements, keeping nothing.
    void ensure(size_t n) {
        if (n > size_) resize(n);
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
    T *get() { return ptr_; }
    const T *get() const { return ptr_; }
    size_t size() const { return size_; }

private:
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        size_ = 0;
    }
    T *ptr_ = nullptr;
    size_t size_ = 0;
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

// A level of the modulus chain, named by its number of primes (the key level has all
// K, the first data level K-1, the last 1).
struct ParmsID {
    uint32_t level = 0;
    bool operator==(const ParmsID &o) const { return level == o.level; }
    bool operator!=(const ParmsID &o) const { return level != o.level; }
};
static const ParmsID parmsIDZero{0};

class SEALContext;

// ContextDataCuda (src/context_cuda.cuh:11-170): one level of the chain.
class ContextData {
public:
    ContextData(ParmsID id, std::vector<Modulus> q, SchemeType scheme, size_t n, Modulus t, bool has_next)
        : id_(id), q_(std::move(q)), scheme_(scheme), n_(n), t_(t), has_next_(has_next) {}
    ParmsID parmsID() const { return id_; }
    size_t chainIndex() const { return id_.level - 1; }
    const std::vector<Modulus> &coeffModulus() const { return q_; }
    SchemeType scheme() const { return scheme_; }
    size_t polyModulusDegree() const { return n_; }
    const Modulus &plainModulus() const { return t_; }
    int totalCoeffModulusBitCount() const {
        int b = 0;
        for (auto &m : q_) b += m.bitCount();
        return b;
    }
    ParmsID nextParmsID() const { return has_next_ ? ParmsID{id_.level - 1} : parmsIDZero; }

private:
    ParmsID id_;
    std::vector<Modulus> q_;
    SchemeType scheme_;
    size_t n_;
    Modulus t_;
    bool has_next_;
};

// SEALContextCuda (src/context_cuda.cuh:172-238).  Copies share the device tables.
class SEALContext {
public:
    explicit SEALContext(const EncryptionParameters &parms, bool expand_mod_chain = true,
                         SecurityLevel sec = SecurityLevel::tc128)
        : s_(std::make_shared<State>()) {
        KernelProvider::checkInitialized();
        s_->parms = parms;
        s_->expand = expand_mod_chain;
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

    ParmsID keyParmsID() const { return ParmsID{(uint32_t)keyCount()}; }
    ParmsID firstParmsID() const { return ParmsID{(uint32_t)(keyCount() > 1 ? keyCount() - 1 : 1)}; }
    ParmsID lastParmsID() const { return s_->expand ? ParmsID{1} : firstParmsID(); }
    bool usingKeyswitching() const { return keyCount() > 1; }

    std::shared_ptr<const ContextData> getContextData(ParmsID id) const {
        if (id.level == 0 || id.level > keyCount()) return nullptr;
        if (id.level != keyParmsID().level && id.level < lastParmsID().level) return nullptr;
        const auto &q = s_->parms.coeffModulus();
        const bool has_next = id.level > lastParmsID().level && id.level != keyParmsID().level;
        return std::make_shared<ContextData>(id, std::vector<Modulus>(q.begin(), q.begin() + id.level),
                                             scheme(), polyModulusDegree(), s_->parms.plainModulus(), has_next);
    }
    std::shared_ptr<const ContextData> firstContextData() const { return getContextData(firstParmsID()); }
    std::shared_ptr<const ContextData> keyContextData() const { return getContextData(keyParmsID()); }
    std::shared_ptr<const ContextData> lastContextData() const { return getContextData(lastParmsID()); }

    // ---- used by the other classes
    shc_context *handle() const {
        checkSet();
        return s_->handle;
    }
    // Scratch device memory shared by the operations of this context (one stream; the
    // reference's evaluators are not thread safe either).
    void *workspace(size_t bytes) const {
        s_->ws.ensure((bytes + 7) / 8);
        return s_->ws.get();
    }
    uint64_t prime(size_t i) const { return s_->parms.coeffModulus()[i].value(); }
    // The reference's parms_id of a level: BLAKE2b-256 of (scheme, N, the level's primes,
    // plain modulus) as little-endian words (EncryptionParameters::computeParmsID,
    // src/encryptionparams.cpp:118-156).
    std::array<uint64_t, 4> parmsHash(ParmsID id) const {
        std::vector<uint64_t> q;
        for (uint32_t i = 0; i < id.level; i++) q.push_back(prime(i));
        return detail::parms_hash((uint64_t)scheme(), polyModulusDegree(), q, s_->parms.plainModulus().value());
    }
    // The level whose parms_id is `h`, or parmsIDZero.
    ParmsID levelOfHash(const std::array<uint64_t, 4> &h) const {
        for (uint32_t l = 1; l <= keyCount(); l++)
            if (parmsHash(ParmsID{l}) == h) return ParmsID{l};
        return parmsIDZero;
    }
    int levelBits(ParmsID id) const {
        int b = 0;
        for (uint32_t i = 0; i < id.level; i++) b += s_->parms.coeffModulus()[i].bitCount();
        return b;
    }

private:
    struct State {
        EncryptionParameters parms;
        bool expand = true;
        std::string error;
        shc_context *handle = nullptr;
        DeviceArray<uint64_t> ws;
        ~State() {
            if (handle) shc_context_destroy(handle);
        }
    };
    std::shared_ptr<State> s_;
};

// ------------------------------------------------------------------ data containers

// PlaintextCuda: BFV plaintexts are N coefficients mod t (coefficient form) or, after
// transformToNtt, [L][N] in NTT form; CKKS plaintexts are [L][N] in NTT form.
class Plaintext {
public:
    Plaintext() = default;
    ParmsID &parmsID() { return parms_id_; }
    ParmsID parmsID() const { return parms_id_; }
    double &scale() { return scale_; }
    double scale() const { return scale_; }
    bool isNttForm() const { return parms_id_ != parmsIDZero; }
    size_t coeffCount() const { return data_.size(); }
    DeviceArray<uint64_t> &data() { return data_; }
    const DeviceArray<uint64_t> &data() const { return data_; }
    void resize(size_t coeff_count) { data_.resize(coeff_count); }

private:
    DeviceArray<uint64_t> data_;
    ParmsID parms_id_ = parmsIDZero;
    double scale_ = 1.0;
};

class Evaluator;

namespace detail {
template <class T>
inline void put(std::ostream &os, const T &v) {
    os.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <class T>
inline T get(std::istream &is) {
    T v{};
    is.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!is) throw std::invalid_argument("stream ended before the ciphertext did");
    return v;
}
}  // namespace detail

// CiphertextCuda (src/ciphertext_cuda.cuh:12-246): [size][L][N] words on the device.
class Ciphertext {
public:
    Ciphertext() = default;
    void resize(const SEALContext &ctx, ParmsID id, size_t size) {
        n_ = ctx.polyModulusDegree();
        if (id != parms_id_ || parms_hash_ == std::array<uint64_t, 4>{}) parms_hash_ = ctx.parmsHash(id);
        parms_id_ = id;
        size_ = size;
        data_.resize(size * id.level * n_);
    }

    // CiphertextCuda::save / load (src/ciphertext_cuda.cu:8-35, 138-183): parms_id
    // (32 B), is_ntt_form (1 B), size, N, L (8 B each), scale (double), correction
    // factor (1), seed (0), terms flag (false), word count, words -- the reference's raw
    // layout, so files move between the two.  Seeded ciphertexts (c1 expanded from a
    // curand stream) are not loadable here.
    void save(std::ostream &os) const {
        saveHeader(os, false);
        const std::vector<uint64_t> h = data_.toHost();
        const uint64_t words = (uint64_t)(size_ * polyWords());
        detail::put(os, words);
        os.write(reinterpret_cast<const char *>(h.data()), words * 8);
    }
    void load(std::istream &is, const SEALContext &ctx) {
        loadHeader(is, ctx, false);
        const uint64_t words = detail::get<uint64_t>(is);
        if (words != size_ * polyWords()) throw std::invalid_argument("ciphertext word count does not match its shape");
        std::vector<uint64_t> h(words);
        is.read(reinterpret_cast<char *>(h.data()), words * 8);
        if (!is) throw std::invalid_argument("stream ended before the ciphertext did");
        data_.fromHost(h.data(), h.size());
    }
    // saveTerms / loadTerms (src/ciphertext_cuda.cu:37-136): only the listed coefficients
    // of c0 (coefficient form), then every other polynomial in full.
    void saveTerms(std::ostream &os, const Evaluator &ev, const std::vector<size_t> &term_ids) const;
    void loadTerms(std::istream &is, const Evaluator &ev, const SEALContext &ctx, const std::vector<size_t> &term_ids);
    ParmsID parmsID() const { return parms_id_; }
    ParmsID &parmsID() { return parms_id_; }
    size_t size() const { return size_; }
    size_t polyModulusDegree() const { return n_; }
    size_t coeffModulusSize() const { return parms_id_.level; }
    bool isNttForm() const { return ntt_form_; }
    bool &isNttForm() { return ntt_form_; }
    double scale() const { return scale_; }
    double &scale() { return scale_; }
    uint64_t *data() { return data_.get(); }
    const uint64_t *data() const { return data_.get(); }
    uint64_t *data(size_t poly) { return data_.get() + poly * coeffModulusSize() * n_; }
    const uint64_t *data(size_t poly) const { return data_.get() + poly * coeffModulusSize() * n_; }
    size_t polyWords() const { return coeffModulusSize() * n_; }
    std::vector<uint64_t> toHost() const { return data_.toHost(); }
    // Keeps the first `size` polynomials (relinearization drops the third).
    void shrink(size_t size) { size_ = size; }
    DeviceArray<uint64_t> &dynArray() { return data_; }
    const std::array<uint64_t, 4> &parmsHash() const { return parms_hash_; }

private:
    void saveHeader(std::ostream &os, bool terms) const {
        detail::put(os, parms_hash_);
        detail::put(os, ntt_form_);
        detail::put(os, (uint64_t)size_);
        detail::put(os, (uint64_t)n_);
        detail::put(os, (uint64_t)coeffModulusSize());
        detail::put(os, scale_);
        detail::put(os, (uint64_t)1);  // correction factor
        detail::put(os, (uint64_t)0);  // seed
        detail::put(os, terms);
    }
    void loadHeader(std::istream &is, const SEALContext &ctx, bool terms) {
        const auto hash = detail::get<std::array<uint64_t, 4>>(is);
        const bool ntt = detail::get<bool>(is);
        const uint64_t size = detail::get<uint64_t>(is), n = detail::get<uint64_t>(is), L = detail::get<uint64_t>(is);
        const double scale = detail::get<double>(is);
        (void)detail::get<uint64_t>(is);  // correction factor
        const uint64_t seed = detail::get<uint64_t>(is);
        const bool is_terms = detail::get<bool>(is);
        if (seed != 0) throw std::invalid_argument("seeded ciphertexts are not supported");
        if (is_terms != terms)
            throw std::invalid_argument(terms ? "Trying to load a normal ciphertext, but term indices is specified"
                                              : "Trying to load a termed ciphertext, but indices is not specified");
        const ParmsID id = ctx.levelOfHash(hash);
        if (id == parmsIDZero || id.level != L || n != ctx.polyModulusDegree())
            throw std::invalid_argument("ciphertext is not valid for encryption parameters");
        resize(ctx, id, size);
        ntt_form_ = ntt;
        scale_ = scale;
    }

    DeviceArray<uint64_t> data_;
    std::array<uint64_t, 4> parms_hash_{};
    ParmsID parms_id_ = parmsIDZero;
    size_t size_ = 0, n_ = 0;
    bool ntt_form_ = false;
    double scale_ = 1.0;
};

// SecretKeyCuda: [K][N] in NTT form.
class SecretKey {
public:
    DeviceArray<uint64_t> &data() { return data_; }
    const DeviceArray<uint64_t> &data() const { return data_; }
    ParmsID parmsID() const { return parms_id_; }
    ParmsID &parmsID() { return parms_id_; }

private:
    DeviceArray<uint64_t> data_;
    ParmsID parms_id_ = parmsIDZero;
};

// PublicKeyCuda: a size-2 key-level ciphertext in NTT form.
class PublicKey {
public:
    Ciphertext &data() { return data_; }
    const Ciphertext &data() const { return data_; }
    ParmsID parmsID() const { return data_.parmsID(); }

private:
    Ciphertext data_;
};

// KSwitchKeysCuda: [K-1][2][K][N] per key.  RelinKeys hold one (for s^2); GaloisKeys
// hold one per Galois element.
class RelinKeys {
public:
    DeviceArray<uint64_t> &data() { return data_; }
    const DeviceArray<uint64_t> &data() const { return data_; }
    ParmsID parmsID() const { return parms_id_; }
    ParmsID &parmsID() { return parms_id_; }
    size_t size() const { return data_.size() ? 1 : 0; }

private:
    DeviceArray<uint64_t> data_;
    ParmsID parms_id_ = parmsIDZero;
};

class GaloisKeys {
public:
    bool hasKey(uint32_t elt) const { return keys_.count(elt) != 0; }
    const DeviceArray<uint64_t> &key(uint32_t elt) const {
        auto it = keys_.find(elt);
        if (it == keys_.end()) throw std::invalid_argument("Galois key not present");
        return it->second;
    }
    DeviceArray<uint64_t> &operator[](uint32_t elt) { return keys_[elt]; }
    size_t size() const { return keys_.size(); }
    ParmsID parmsID() const { return parms_id_; }
    ParmsID &parmsID() { return parms_id_; }

private:
    std::map<uint32_t, DeviceArray<uint64_t>> keys_;
    ParmsID parms_id_ = parmsIDZero;
};

namespace detail {
inline uint64_t fresh_seed() {
    std::random_device rd;
    return ((uint64_t)rd() << 32) ^ rd();
}
inline uint32_t elt_from_step(size_t n, int step) {
    uint32_t elt = 0;
    check(shc_galois_elt_from_step((uint32_t)n, step, &elt));
    return elt;
}
}  // namespace detail

// ---------------------------------------------------------------------- key generator

// KeyGeneratorCuda (src/keygenerator_cuda.cuh:30-300).  Keys are generated on the
// device with the library's counter-based sampler (seeded from std::random_device).
class KeyGenerator {
public:
    explicit KeyGenerator(const SEALContext &ctx) : ctx_(ctx), seed_(detail::fresh_seed()) {
        ctx_.checkSet();
        sk_.data().resize(ctx_.keyCount() * ctx_.polyModulusDegree());
        check(shc_keygen_secret(ctx_.handle(), next(), sk_.data().get(), nullptr));
        sk_.parmsID() = ctx_.keyParmsID();
    }
    KeyGenerator(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk), seed_(detail::fresh_seed()) {
        ctx_.checkSet();
    }
    const SecretKey &secretKey() const { return sk_; }

    void createPublicKey(PublicKey &pk) const {
        pk.data().resize(ctx_, ctx_.keyParmsID(), 2);
        pk.data().isNttForm() = true;
        check(shc_keygen_public(ctx_.handle(), sk_.data().get(), next(), pk.data().data(), nullptr));
    }
    PublicKey createPublicKey() const {
        PublicKey pk;
        createPublicKey(pk);
        return pk;
    }
    void createRelinKeys(RelinKeys &rk) const {
        requireKeyswitching();
        rk.data().resize(keyWords());
        size_t ws = 0;
        check(shc_keygen_workspace_size(ctx_.handle(), &ws));
        check(shc_keygen_relin(ctx_.handle(), sk_.data().get(), next(), rk.data().get(), ctx_.workspace(ws), ws,
                               nullptr));
        rk.parmsID() = ctx_.keyParmsID();
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
        for (uint32_t e : elts) {
            auto &k = gk[e];
            k.resize(keyWords());
            check(shc_keygen_galois(ctx_.handle(), sk_.data().get(), e, next(), k.get(), ctx_.workspace(ws), ws,
                                    nullptr));
        }
        gk.parmsID() = ctx_.keyParmsID();
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
    uint64_t next() const { return seed_ + (counter_++) * 0x9E3779B97F4A7C15ull; }
    size_t keyWords() const {
        return (ctx_.keyCount() - 1) * 2 * ctx_.keyCount() * ctx_.polyModulusDegree();
    }
    void requireKeyswitching() const {
        if (!ctx_.usingKeyswitching()) throw std::logic_error("keyswitching is not supported by the context");
    }
    SEALContext ctx_;
    SecretKey sk_;
    uint64_t seed_;
    mutable uint64_t counter_ = 1;
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
        plain.resize(n);
        plain.parmsID() = parmsIDZero;
        check(shc_batch_encode(ctx_.handle(), dv.get(), (uint32_t)values.size(), plain.data().get(), nullptr));
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
        DeviceArray<uint64_t> dv(n), scratch(n);
        check(shc_batch_decode(ctx_.handle(), plain.data().get(), dv.get(), scratch.get(), nullptr));
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
        if (!ctx_.getContextData(id)) throw std::invalid_argument("parms_id is not valid for encryption parameters");
        if (scale <= 0 || (int)std::log2(scale) >= ctx_.levelBits(id))
            throw std::invalid_argument("scale out of bounds");
        std::vector<double> flat(2 * values.size());
        for (size_t i = 0; i < values.size(); i++) {
            flat[2 * i] = values[i].real();
            flat[2 * i + 1] = values[i].imag();
        }
        DeviceArray<double> dv;
        dv.fromHost(flat.data(), flat.size());
        plain.resize(id.level * ctx_.polyModulusDegree());
        size_t ws = 0;
        check(shc_ckks_encode_workspace_size(ctx_.handle(), &ws));
        check(shc_ckks_encode(ctx_.handle(), dv.get(), (uint32_t)values.size(), scale, id.level,
                              plain.data().get(), ctx_.workspace(ws), ws, nullptr));
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
        const uint32_t L = plain.parmsID().level;
        const size_t n = ctx_.polyModulusDegree();
        DeviceArray<double> dv(n);
        size_t ws = 0;
        check(shc_ckks_decode_workspace_size(ctx_.handle(), L, &ws));
        check(shc_ckks_decode(ctx_.handle(), plain.data().get(), L, plain.scale(), dv.get(), ctx_.workspace(ws), ws,
                              nullptr));
        const std::vector<double> h = dv.toHost();
        out.resize(slotCount());
        for (size_t i = 0; i < out.size(); i++) out[i] = {h[2 * i], h[2 * i + 1]};
    }

private:
    SEALContext ctx_;
};

// ------------------------------------------------------------------- encrypt/decrypt

// EncryptorCuda (src/encryptor_cuda.cuh:50-400).  Randomness: the library's counter-
// based sampler, a fresh seed per call (the reference uses curand).
class Encryptor {
public:
    Encryptor(const SEALContext &ctx, const PublicKey &pk) : ctx_(ctx), pk_(pk), seed_(detail::fresh_seed()) {
        ctx_.checkSet();
        has_pk_ = true;
    }
    Encryptor(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk), seed_(detail::fresh_seed()) {
        ctx_.checkSet();
        has_sk_ = true;
    }
    Encryptor(const SEALContext &ctx, const PublicKey &pk, const SecretKey &sk)
        : ctx_(ctx), pk_(pk), sk_(sk), seed_(detail::fresh_seed()) {
        ctx_.checkSet();
        has_pk_ = has_sk_ = true;
    }

    void encrypt(const Plaintext &plain, Ciphertext &ct) const {
        if (!has_pk_) throw std::logic_error("public key is not set");
        const ParmsID id = levelOf(plain);
        ct.resize(ctx_, id, 2);
        ct.isNttForm() = ctx_.scheme() == SchemeType::ckks;
        ct.scale() = ctx_.scheme() == SchemeType::ckks ? plain.scale() : 1.0;
        size_t ws = 0;
        check(shc_encrypt_workspace_size(ctx_.handle(), id.level, &ws));
        check(shc_encrypt(ctx_.handle(), pk_.data().data(), plain.data().get(), id.level, next(), ct.data(),
                          ctx_.workspace(ws), ws, nullptr));
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
        ct.resize(ctx_, id, 2);
        const bool ckks = ctx_.scheme() == SchemeType::ckks;
        ct.isNttForm() = ckks;
        ct.scale() = ckks ? plain.scale() : 1.0;
        check(shc_encrypt_zero_symmetric(ctx_.handle(), sk_.data().get(), id.level, next(), ct.data(), nullptr));
        if (ckks)
            check(shc_add_plain(ctx_.handle(), ct.data(), plain.data().get(), ct.data(), 1, 2, id.level, 0, nullptr));
        else
            check(shc_bfv_add_plain(ctx_.handle(), ct.data(), plain.data().get(), ct.data(), 1, 2, id.level, 0,
                                    nullptr));
    }
    void encryptZero(ParmsID id, Ciphertext &ct) const {
        if (!has_pk_) throw std::logic_error("public key is not set");
        Plaintext zero;
        if (ctx_.scheme() == SchemeType::ckks) {
            zero.resize(id.level * ctx_.polyModulusDegree());
            hip_check(hipMemset(zero.data().get(), 0, zero.data().size() * 8), "memset");
            zero.parmsID() = id;
        } else {
            zero.resize(ctx_.polyModulusDegree());
            hip_check(hipMemset(zero.data().get(), 0, zero.data().size() * 8), "memset");
        }
        encrypt(zero, ct);
    }
    void encryptZero(Ciphertext &ct) const { encryptZero(ctx_.firstParmsID(), ct); }

private:
    ParmsID levelOf(const Plaintext &plain) const {
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!plain.isNttForm()) throw std::invalid_argument("plain must be in NTT form");
            return plain.parmsID();
        }
        if (plain.isNttForm()) throw std::invalid_argument("plain cannot be in NTT form");
        return ctx_.firstParmsID();
    }
    uint64_t next() const { return seed_ + (counter_++) * 0x9E3779B97F4A7C15ull; }
    SEALContext ctx_;
    PublicKey pk_;
    SecretKey sk_;
    bool has_pk_ = false, has_sk_ = false;
    uint64_t seed_;
    mutable uint64_t counter_ = 1;
};

// DecryptorCuda (src/decryptor_cuda.cuh:40-90).
class Decryptor {
public:
    Decryptor(const SEALContext &ctx, const SecretKey &sk) : ctx_(ctx), sk_(sk) { ctx_.checkSet(); }
    void decrypt(const Ciphertext &ct, Plaintext &plain) {
        if (ct.size() < 2 || ct.size() > 3) throw std::invalid_argument("ciphertext size must be 2 or 3");
        const uint32_t L = ct.parmsID().level;
        const bool ckks = ctx_.scheme() == SchemeType::ckks;
        if (ckks != ct.isNttForm()) throw std::invalid_argument("ciphertext is in the wrong form");
        size_t ws = 0;
        check(shc_decrypt_workspace_size(ctx_.handle(), L, &ws));
        plain.resize(ckks ? (size_t)L * ctx_.polyModulusDegree() : ctx_.polyModulusDegree());
        check(shc_decrypt(ctx_.handle(), ct.data(), (uint32_t)ct.size(), sk_.data().get(), L, plain.data().get(),
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
        check(shc_negate(ctx_.handle(), a.data(), a.data(), (uint32_t)a.size(), a.parmsID().level, nullptr));
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
        const uint32_t L = a.parmsID().level;
        Ciphertext out;
        out.resize(ctx_, a.parmsID(), 3);
        out.isNttForm() = a.isNttForm();
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!a.isNttForm() || !b.isNttForm()) throw std::invalid_argument("encrypted must be in NTT form");
            out.scale() = a.scale() * b.scale();
            checkScale(out.scale(), a.parmsID());
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
        const uint32_t L = a.parmsID().level;
        size_t ws = 0;
        check(shc_keyswitch_workspace_size(ctx_.handle(), 1, L, &ws));
        check(shc_relinearize(ctx_.handle(), a.data(), rk.data().get(), 1, L, ctx_.workspace(ws), ws, nullptr));
        a.shrink(2);
    }
    void relinearize(const Ciphertext &a, const RelinKeys &rk, Ciphertext &out) const {
        out = a;
        relinearizeInplace(out, rk);
    }

    // ---- modulus switching
    void rescaleToNext(const Ciphertext &a, Ciphertext &out) const {
        if (ctx_.scheme() != SchemeType::ckks) throw std::invalid_argument("unsupported operation for scheme type");
        divideRound(a, out);
        out.scale() = a.scale() / (double)ctx_.prime(a.parmsID().level - 1);
    }
    void rescaleToNextInplace(Ciphertext &a) const {
        Ciphertext out;
        rescaleToNext(a, out);
        a = std::move(out);
    }
    void rescaleToInplace(Ciphertext &a, ParmsID id) const {
        while (a.parmsID().level > id.level) rescaleToNextInplace(a);
    }
    void modSwitchToNext(const Ciphertext &a, Ciphertext &out) const {
        if (ctx_.scheme() == SchemeType::bfv) {
            divideRound(a, out);
            out.scale() = a.scale();
            return;
        }
        // CKKS: drop the last limb of every polynomial (modSwitchDropToNext)
        const uint32_t L = nextLevel(a);
        out.resize(ctx_, ParmsID{L}, a.size());
        out.isNttForm() = a.isNttForm();
        out.scale() = a.scale();
        const size_t n = ctx_.polyModulusDegree();
        hip_check(hipMemcpy2D(out.data(), L * n * 8, a.data(), (L + 1) * n * 8, L * n * 8, a.size(),
                              hipMemcpyDeviceToDevice),
                  "modswitch copy");
    }
    void modSwitchToNextInplace(Ciphertext &a) const {
        Ciphertext out;
        modSwitchToNext(a, out);
        a = std::move(out);
    }
    void modSwitchToInplace(Ciphertext &a, ParmsID id) const {
        while (a.parmsID().level > id.level) modSwitchToNextInplace(a);
    }
    // Plaintexts in NTT form (CKKS) drop their last limb.
    void modSwitchToNextInplace(Plaintext &p) const {
        if (!p.isNttForm()) throw std::invalid_argument("plain is not in NTT form");
        if (p.parmsID().level < 2) throw std::invalid_argument("end of modulus switching chain reached");
        const uint32_t L = p.parmsID().level - 1;
        DeviceArray<uint64_t> d(L * ctx_.polyModulusDegree());
        hip_check(hipMemcpy(d.get(), p.data().get(), d.size() * 8, hipMemcpyDeviceToDevice), "copy");
        p.data() = std::move(d);
        p.parmsID() = ParmsID{L};
    }
    void modSwitchToInplace(Plaintext &p, ParmsID id) const {
        while (p.parmsID().level > id.level) modSwitchToNextInplace(p);
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
        const uint32_t L = a.parmsID().level;
        if (a.isNttForm() != p.isNttForm()) throw std::invalid_argument("NTT form mismatch");
        if (a.isNttForm()) {
            if (a.parmsID() != p.parmsID()) throw std::invalid_argument("encrypted_ntt and plain_ntt parameter mismatch");
            check(shc_multiply_plain_ntt(ctx_.handle(), a.data(), p.data().get(), a.data(), (uint32_t)a.size(), L,
                                         nullptr));
        } else {
            size_t ws = 0;
            check(shc_bfv_multiply_plain_workspace_size(ctx_.handle(), L, &ws));
            check(shc_bfv_multiply_plain(ctx_.handle(), a.data(), p.data().get(), a.data(), (uint32_t)a.size(), L,
                                         ctx_.workspace(ws), ws, nullptr));
        }
        if (ctx_.scheme() == SchemeType::ckks) {
            a.scale() *= p.scale();
            checkScale(a.scale(), a.parmsID());
        }
    }
    void multiplyPlain(const Ciphertext &a, const Plaintext &p, Ciphertext &out) const {
        out = a;
        multiplyPlainInplace(out, p);
    }

    // ---- NTT form
    void transformToNttInplace(Ciphertext &a) const {
        if (a.isNttForm()) throw std::invalid_argument("encrypted is already in NTT form");
        check(shc_ntt_forward(ctx_.handle(), a.data(), (uint32_t)a.size(), a.parmsID().level, nullptr));
        a.isNttForm() = true;
    }
    void transformToNtt(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        transformToNttInplace(out);
    }
    void transformFromNttInplace(Ciphertext &a) const {
        if (!a.isNttForm()) throw std::invalid_argument("encrypted is not in NTT form");
        check(shc_ntt_inverse(ctx_.handle(), a.data(), (uint32_t)a.size(), a.parmsID().level, nullptr));
        a.isNttForm() = false;
    }
    void transformFromNtt(const Ciphertext &a, Ciphertext &out) const {
        out = a;
        transformFromNttInplace(out);
    }
    // BFV plaintext -> lifted NTT form at level `id`.
    void transformToNttInplace(Plaintext &p, ParmsID id) const {
        if (p.isNttForm()) throw std::invalid_argument("plain is already in NTT form");
        DeviceArray<uint64_t> d(id.level * ctx_.polyModulusDegree());
        check(shc_bfv_plain_to_ntt(ctx_.handle(), p.data().get(), id.level, d.get(), nullptr));
        p.data() = std::move(d);
        p.parmsID() = id;
    }

    // ---- Galois
    void applyGaloisInplace(Ciphertext &a, uint32_t elt, const GaloisKeys &gk) const {
        if (gk.parmsID() != ctx_.keyParmsID()) throw std::invalid_argument("galois_keys is not valid for encryption parameters");
        if (a.size() != 2) throw std::invalid_argument("encrypted size must be 2");
        const uint32_t L = a.parmsID().level;
        size_t ws = 0;
        check(shc_galois_workspace_size(ctx_.handle(), 1, L, &ws));
        check(shc_apply_galois_keyswitch(ctx_.handle(), a.data(), gk.key(elt).get(), elt, 1, L, ctx_.workspace(ws), ws,
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
    void sameParms(const Ciphertext &a, const Ciphertext &b) const {
        if (a.parmsID() != b.parmsID()) throw std::invalid_argument("encrypted1 and encrypted2 parameter mismatch");
        if (a.isNttForm() != b.isNttForm()) throw std::invalid_argument("NTT form mismatch");
    }
    void checkScale(double scale, ParmsID id) const {
        if (!(scale > 0) || (int)std::log2(scale) >= ctx_.levelBits(id)) throw std::invalid_argument("scale out of bounds");
    }
    static bool scalesClose(double a, double b) {
        const double d = std::fabs(a - b);
        return d <= std::max(std::fabs(a), std::fabs(b)) * 1e-9;
    }
    void addSub(Ciphertext &a, const Ciphertext &b, bool sub) const {
        sameParms(a, b);
        if (ctx_.scheme() == SchemeType::ckks && !scalesClose(a.scale(), b.scale()))
            throw std::invalid_argument("scale mismatch");
        const uint32_t L = a.parmsID().level;
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
        const uint32_t L = a.parmsID().level;
        if (ctx_.scheme() == SchemeType::ckks) {
            if (!p.isNttForm() || !a.isNttForm() || p.parmsID() != a.parmsID())
                throw std::invalid_argument("encrypted and plain parameter mismatch");
            if (!scalesClose(a.scale(), p.scale())) throw std::invalid_argument("scale mismatch");
            check(shc_add_plain(ctx_.handle(), a.data(), p.data().get(), a.data(), 1, (uint32_t)a.size(), L, sub,
                                nullptr));
        } else {
            if (p.isNttForm() || a.isNttForm()) throw std::invalid_argument("NTT form mismatch");
            check(shc_bfv_add_plain(ctx_.handle(), a.data(), p.data().get(), a.data(), 1, (uint32_t)a.size(), L, sub,
                                    nullptr));
        }
    }
    uint32_t nextLevel(const Ciphertext &a) const {
        auto cd = ctx_.getContextData(a.parmsID());
        if (!cd) throw std::invalid_argument("encrypted is not valid for encryption parameters");
        if (cd->nextParmsID() == parmsIDZero) throw std::invalid_argument("end of modulus switching chain reached");
        return a.parmsID().level - 1;
    }
    void divideRound(const Ciphertext &a, Ciphertext &out) const {
        const uint32_t L = nextLevel(a) + 1;
        out.resize(ctx_, ParmsID{L - 1}, a.size());
        out.isNttForm() = a.isNttForm();
        size_t ws = 0;
        check(shc_rescale_workspace_size(ctx_.handle(), 1, (uint32_t)a.size(), L, &ws));
        check(shc_rescale_to_next(ctx_.handle(), a.data(), out.data(), 1, (uint32_t)a.size(), L, ctx_.workspace(ws),
                                  ws, nullptr));
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
    const std::vector<uint64_t> h = coeff.toHost();
    saveHeader(os, true);
    const size_t L = coeffModulusSize();
    for (size_t id : term_ids) {
        if (id >= n_) throw std::invalid_argument("term index out of range");
        for (size_t j = 0; j < L; j++) detail::put(os, h[j * n_ + id]);
    }
    const uint64_t rest = (uint64_t)((size_ - 1) * polyWords());
    detail::put(os, rest);
    os.write(reinterpret_cast<const char *>(h.data() + polyWords()), rest * 8);
}

inline void Ciphertext::loadTerms(std::istream &is, const Evaluator &ev, const SEALContext &ctx,
                                  const std::vector<size_t> &term_ids) {
    loadHeader(is, ctx, true);
    const size_t L = coeffModulusSize();
    std::vector<uint64_t> h(size_ * polyWords(), 0);
    for (size_t id : term_ids) {
        if (id >= n_) throw std::invalid_argument("term index out of range");
        for (size_t j = 0; j < L; j++) h[j * n_ + id] = detail::get<uint64_t>(is);
    }
    const uint64_t rest = detail::get<uint64_t>(is);
    if (rest != (size_ - 1) * polyWords()) throw std::invalid_argument("ciphertext word count does not match its shape");
    is.read(reinterpret_cast<char *>(h.data() + polyWords()), rest * 8);
    if (!is) throw std::invalid_argument("stream ended before the ciphertext did");
    data_.fromHost(h.data(), h.size());
    if (ntt_form_) {
        ntt_form_ = false;
        ev.transformToNttInplace(*this);
    }
}

}  // namespace troyhip
