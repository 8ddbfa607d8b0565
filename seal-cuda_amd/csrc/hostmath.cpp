// hostmath.cpp -- see hostmath.h.
#include "hostmath.h"
#include "common.h"
#include <algorithm>
#include <map>

namespace shc {
namespace hm {

uint64_t powmod(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    a %= q;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod(r, a, q);
        a = mulmod(a, a, q);
    }
    return r;
}

uint64_t invmod(uint64_t a, uint64_t q) {
    if (q < 2) return 0;
    a %= q;
    if (!a) return 0;
    __int128 r0 = q, r1 = a, s0 = 0, s1 = 1;
    while (r1) {
        __int128 k = r0 / r1, t;
        t = r0 - k * r1; r0 = r1; r1 = t;
        t = s0 - k * s1; s0 = s1; s1 = t;
    }
    if (r0 != 1) return 0;
    if (s0 < 0) s0 += q;
    return (uint64_t)s0;
}

void barrett_ratio(uint64_t q, uint64_t &lo, uint64_t &hi) {
    u128 top = ((u128)1 << 64) / q;
    u128 rem = ((u128)1 << 64) % q;
    hi = (uint64_t)top;
    lo = (uint64_t)((rem << 64) / q);
}

uint32_t reverse_bits(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; i++) r = (r << 1) | ((x >> i) & 1u);
    return r;
}

int bit_count(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

bool is_prime(uint64_t v) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (v < 2) return false;
    for (uint64_t b : bases) {
        if (v == b) return true;
        if (v % b == 0) return false;
    }
    uint64_t d = v - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (uint64_t b : bases) {
        uint64_t x = powmod(b, d, v);
        if (x == 1 || x == v - 1) continue;
        bool composite = true;
        for (int i = 1; i < s && composite; i++) {
            x = mulmod(x, x, v);
            if (x == v - 1) composite = false;
        }
        if (composite) return false;
    }
    return true;
}

std::vector<uint64_t> get_primes(uint64_t factor, int bit_size, size_t count) {
    std::vector<uint64_t> out;
    uint64_t value = ((uint64_t(1) << bit_size) - 1) / factor * factor + 1;
    const uint64_t lower = uint64_t(1) << (bit_size - 1);
    while (out.size() < count && value > lower) {
        if (is_prime(value)) out.push_back(value);
        value -= factor;
    }
    if (out.size() < count) fail(ERR_LOGIC, "failed to find enough qualifying primes");
    return out;
}

static uint64_t gcd(uint64_t a, uint64_t b) {
    while (b) { uint64_t t = a % b; a = b; b = t; }
    return a;
}

std::vector<uint64_t> coeff_modulus_create(size_t n, const std::vector<int> &bit_sizes, uint64_t plain_modulus) {
    if (n < 2 || n > 131072 || (n & (n - 1))) fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree is invalid");
    if (bit_sizes.size() > 64) fail(ERR_INVALID_ARGUMENT, "bit_sizes is invalid");
    for (int b : bit_sizes)
        if (b < 2 || b > 60) fail(ERR_INVALID_ARGUMENT, "bit_sizes is invalid");
    uint64_t factor = 2 * (uint64_t)n;
    if (plain_modulus) factor *= plain_modulus / gcd(plain_modulus, factor);
    std::map<int, size_t> count;
    for (int b : bit_sizes) count[b]++;
    // primes of each size, descending; handed out from the back (ascending per size)
    std::map<int, std::vector<uint64_t>> table;
    for (auto &kv : count) table[kv.first] = get_primes(factor, kv.first, kv.second);
    std::vector<uint64_t> out;
    for (int b : bit_sizes) {
        out.push_back(table[b].back());
        table[b].pop_back();
    }
    return out;
}

bool minimal_primitive_root(uint64_t degree, uint64_t q, uint64_t &root) {
    if (q < 3 || (q - 1) % degree) return false;
    const uint64_t cofactor = (q - 1) / degree;
    uint64_t g = 0;
    for (uint64_t c = 2; c < q && c < (1u << 22); c++) {
        uint64_t r = powmod(c, cofactor, q);
        if (powmod(r, degree >> 1, q) == q - 1) { g = r; break; }
    }
    if (!g) return false;
    // the primitive degree-th roots are the odd powers of g; keep the smallest
    const uint64_t g2 = mulmod(g, g, q);
    uint64_t best = g, cur = g;
    for (uint64_t i = 0; i < degree; i += 2) {
        best = std::min(best, cur);
        cur = mulmod(cur, g2, q);
    }
    root = best;
    return true;
}

uint32_t galois_elt_from_step(uint64_t n, int step) {
    const uint64_t m = 2 * n;
    if (step == 0) return (uint32_t)(m - 1);
    const uint64_t pos = (uint64_t)(step < 0 ? -(int64_t)step : step);
    if (pos >= (n >> 1)) fail(ERR_INVALID_ARGUMENT, "step count too large");
    uint64_t e = step < 0 ? (n >> 1) - pos : pos;
    uint64_t g = 1;
    while (e--) g = (g * 3) & (m - 1);
    return (uint32_t)g;
}

NttTables make_ntt_tables(int logn, uint64_t q) {
    NttTables t;
    const uint64_t n = uint64_t(1) << logn;
    t.q = q;
    if (!minimal_primitive_root(2 * n, q, t.root)) fail(ERR_INVALID_ARGUMENT, "invalid modulus");
    const uint64_t inv_root = invmod(t.root, q);
    std::vector<uint64_t> rp(n), irp(n);
    uint64_t p = t.root;
    for (uint64_t i = 1; i < n; i++) {
        rp[reverse_bits((uint32_t)i, logn)] = p;
        p = mulmod(p, t.root, q);
    }
    rp[0] = 1;
    p = inv_root;
    for (uint64_t i = 1; i < n; i++) {
        irp[reverse_bits((uint32_t)(i - 1), logn) + 1] = p;
        p = mulmod(p, inv_root, q);
    }
    irp[0] = 1;
    t.fwd.resize(2 * n);
    t.inv.resize(2 * n);
    for (uint64_t i = 0; i < n; i++) {
        t.fwd[2 * i] = rp[i];
        t.fwd[2 * i + 1] = shoup(rp[i], q);
        t.inv[2 * i] = irp[i];
        t.inv[2 * i + 1] = shoup(irp[i], q);
    }
    t.inv_n = invmod(n % q, q);
    t.inv_n_shoup = shoup(t.inv_n, q);
    return t;
}

// ---------------------------------------------------------------- multi-precision

Big big_mul_word(const Big &a, uint64_t w) {
    Big r(a.size() + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < a.size(); i++) {
        u128 p = (u128)a[i] * w + carry;
        r[i] = (uint64_t)p;
        carry = (uint64_t)(p >> 64);
    }
    r[a.size()] = carry;
    while (r.size() > 1 && r.back() == 0) r.pop_back();
    return r;
}

uint64_t big_mod_word(const Big &a, uint64_t q) {
    u128 r = 0;
    for (size_t i = a.size(); i-- > 0;) r = ((r << 64) | a[i]) % q;
    return (uint64_t)r;
}

Big big_product(const std::vector<uint64_t> &ws) {
    Big r{1};
    for (uint64_t w : ws) r = big_mul_word(r, w);
    return r;
}

Big big_product_except(const std::vector<uint64_t> &ws, size_t skip) {
    Big r{1};
    for (size_t i = 0; i < ws.size(); i++)
        if (i != skip) r = big_mul_word(r, ws[i]);
    return r;
}

Big big_div_word(const Big &a, uint64_t d, uint64_t &rem) {
    Big q(a.size(), 0);
    u128 r = 0;
    for (size_t i = a.size(); i-- > 0;) {
        r = (r << 64) | a[i];
        q[i] = (uint64_t)(r / d);
        r %= d;
    }
    rem = (uint64_t)r;
    while (q.size() > 1 && q.back() == 0) q.pop_back();
    return q;
}

int big_bit_count(const Big &a) {
    for (size_t i = a.size(); i-- > 0;)
        if (a[i]) return (int)(64 * i) + bit_count(a[i]);
    return 0;
}

}  // namespace hm
}  // namespace shc
