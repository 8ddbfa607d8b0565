// hostmath.h -- host-side number theory that builds the device tables.
//
// The reference computes every table on the CPU at context creation
// (SEALContext::validate, src/context.cpp:138-424; NTTTables, src/utils/ntt.cpp:17-66;
// RNSTool, src/utils/rns.cpp:581-803) and uploads them; this library does the same
// in its own host code.
#pragma once
#include <stdint.h>
#include <vector>

namespace shc {
namespace hm {

using u128 = unsigned __int128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t q);
uint64_t invmod(uint64_t a, uint64_t q);              // 0 if not invertible
inline uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }
void barrett_ratio(uint64_t q, uint64_t &lo, uint64_t &hi);  // floor(2^128 / q)
uint32_t reverse_bits(uint32_t x, int bits);
int bit_count(uint64_t v);
bool is_prime(uint64_t v);
// getPrimes (src/utils/numth.cpp:261-283): `count` primes = 1 mod factor, descending from 2^bits
std::vector<uint64_t> get_primes(uint64_t factor, int bit_size, size_t count);
// CoeffModulus::Create (src/modulus.cpp:80-165); plain_modulus 0 for the variant without t
std::vector<uint64_t> coeff_modulus_create(size_t n, const std::vector<int> &bit_sizes, uint64_t plain_modulus = 0);
// tryMinimalPrimitiveRoot (src/utils/numth.cpp:335-362)
bool minimal_primitive_root(uint64_t degree, uint64_t q, uint64_t &root);

// GaloisTool::getEltFromStep (src/utils/galois.cpp:44-86); throws on |step| >= N/2.
uint32_t galois_elt_from_step(uint64_t n, int step);

struct NttTables {
    uint64_t q = 0, root = 0;
    std::vector<uint64_t> fwd;  // interleaved (w, shoup(w)) pairs, [2N]
    std::vector<uint64_t> inv;  // [2N]
    uint64_t inv_n = 0, inv_n_shoup = 0;
};
// Throws shc::Error(invalid_argument) when q has no primitive 2N-th root.
NttTables make_ntt_tables(int logn, uint64_t q);

// ----- multi-precision helpers (little-endian 64-bit words)
using Big = std::vector<uint64_t>;
Big big_mul_word(const Big &a, uint64_t w);
uint64_t big_mod_word(const Big &a, uint64_t q);
Big big_product(const std::vector<uint64_t> &ws);        // product of words
Big big_product_except(const std::vector<uint64_t> &ws, size_t skip);
// floor(a / d) and a mod d for a word divisor
Big big_div_word(const Big &a, uint64_t d, uint64_t &rem);
int big_bit_count(const Big &a);

}  // namespace hm
}  // namespace shc
