// Host check of the NTT's lazy arithmetic (seal-cuda_amd/csrc/butterfly.h):
//   * mulhi64_approx(x, y) is floor(x y / 2^64) minus 0, 1 or 2;
//   * shoup_lazier(x, w, wq, 2^64 - q) is congruent to x w mod q and below 4q;
//   * the forward butterfly maps [0, 8q)^2 into [0, 8q)^2 with x + w y, x - w y;
//   * the inverse butterfly maps [0, 4q)^2 into [0, 4q)^2 with x + y, (x - y) w;
//   * for q < 2^60, the wide forward schedule: reducing by 8q maps x in [0, 16q) to
//     outputs in [0, 12q), no reduction maps x in [0, 12q) to outputs in [0, 16q), for
//     any 64-bit y; sub_if_ge(v, 8q) and sub_if_ge_neg(v, 2^64 - 8q) take [0, 16q) to
//     [0, 8q);
//   * the key-switch product sums (accum.h): 31 products t k with t < 8q or 63 with
//     t < 4q (k < q), plus one correction word, equal the exact 128-bit sum;
// for primes from 17 to 61 bits, random and extreme operands (0, range ends,
// multiples of q +- 1, all-ones words).  Exact products come from unsigned __int128.
// Prints "ok <cases>" and exits 0, or prints the first failing case and exits 1.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "accum.h"
#include "butterfly.h"

using u64 = uint64_t;
using wide = unsigned __int128;
using shc::butterfly;
using shc::mulhi64_approx;
using shc::shoup_lazier;
using shc::sub_if_ge;
using shc::sub_if_ge_neg;

static u64 state = 0x9e3779b97f4a7c15ull;
static u64 next() {  // splitmix64
    u64 z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static u64 mulmod(u64 a, u64 b, u64 q) { return (u64)((wide)a * b % q); }
static bool is_prime(u64 n) {
    if (n < 2) return false;
    for (u64 p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
        if (n % p == 0) return n == p;
    }
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) d >>= 1, s++;
    for (u64 a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
        u64 x = 1, b = a % n, e = d;
        while (e) {
            if (e & 1) x = mulmod(x, b, n);
            b = mulmod(b, b, n);
            e >>= 1;
        }
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s && comp; r++) {
            x = mulmod(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}
// operand generator over [0, bound): extremes first, then random
static u64 operand(u64 bound, u64 q, int k) {
    switch (k) {
        case 0: return 0;
        case 1: return bound - 1;
        case 2: return q - 1;
        case 3: return q;
        case 4: return bound >= 2 * q ? 2 * q - 1 : bound - 1;
        case 5: return bound - q;
        default: return next() % bound;
    }
}

int main() {
    std::vector<u64> primes;
    // the largest prime below 2^b (b = 17..61), an NTT prime 1 mod 2^17 near 2^b, 2^61 - 1
    for (int b = 17; b <= 61; b++) {
        u64 p = (u64(1) << b) - 1;
        while (!is_prime(p)) p -= 2;
        primes.push_back(p);
        u64 m = u64(1) << 17;
        u64 c = (((u64(1) << b) - 1) / m) * m + 1;
        while (c > m && !is_prime(c)) c -= m;
        if (c > m) primes.push_back(c);
    }
    long cases = 0;
    for (u64 q : primes) {
        const u64 neg_q = 0 - q, four_q = q << 2;
        for (int wi = 0; wi < 24; wi++) {
            const u64 w = wi == 0 ? 0 : wi == 1 ? 1 : wi == 2 ? q - 1 : next() % q;
            const u64 wq = (u64)(((wide)w << 64) / q);
            for (int k = 0; k < 400; k++) {
                // mulhi64_approx over full 64-bit words
                const u64 a = k == 0 ? ~0ull : next(), c = k == 1 ? ~0ull : next();
                const u64 hi = (u64)(((wide)a * c) >> 64), ap = mulhi64_approx(a, c);
                if (ap > hi || hi - ap > 2) {
                    printf("mulhi64_approx(%llu, %llu) = %llu, exact %llu\n", (unsigned long long)a,
                           (unsigned long long)c, (unsigned long long)ap, (unsigned long long)hi);
                    return 1;
                }
                // shoup_lazier over any 64-bit x
                const u64 x = k == 0 ? ~0ull : k == 1 ? 0 : next();
                const u64 r = shoup_lazier(x, w, wq, neg_q);
                if (r >= four_q || r % q != mulmod(x % q, w, q)) {
                    printf("shoup_lazier q=%llu w=%llu x=%llu -> %llu\n", (unsigned long long)q,
                           (unsigned long long)w, (unsigned long long)x, (unsigned long long)r);
                    return 1;
                }
                // forward butterfly on [0, 8q)
                u64 fx = operand(8 * q, q, k % 8), fy = operand(8 * q, q, (k / 8) % 8);
                const u64 sx = fx % q, wy = mulmod(fy % q, w, q);
                butterfly<false>(fx, fy, ulonglong2{w, wq}, four_q, neg_q);
                if (fx >= 8 * q || fy >= 8 * q || fx % q != (sx + wy) % q || fy % q != (sx + q - wy) % q) {
                    printf("forward butterfly q=%llu w=%llu case %d\n", (unsigned long long)q, (unsigned long long)w, k);
                    return 1;
                }
                // inverse butterfly on [0, 4q)
                u64 ix = operand(4 * q, q, k % 8), iy = operand(4 * q, q, (k / 8) % 8);
                const u64 ax = ix % q, ay = iy % q;
                butterfly<true>(ix, iy, ulonglong2{w, wq}, four_q, neg_q);
                if (ix >= 4 * q || iy >= 4 * q || ix % q != (ax + ay) % q || iy % q != mulmod((ax + q - ay) % q, w, q)) {
                    printf("inverse butterfly q=%llu w=%llu case %d\n", (unsigned long long)q, (unsigned long long)w, k);
                    return 1;
                }
                if (q < (u64(1) << 60)) {
                    // wide schedule, reducing layer: x < 16q, y any word
                    u64 rx = operand(16 * q, q, k % 8), ry = k % 3 == 0 ? ~0ull - k : next();
                    const u64 rsx = rx % q, rwy = mulmod(ry % q, w, q);
                    butterfly<false>(rx, ry, ulonglong2{w, wq}, four_q, neg_q, true, 0 - 8 * q);
                    if (rx >= 12 * q || ry >= 12 * q || rx % q != (rsx + rwy) % q || ry % q != (rsx + q - rwy) % q) {
                        printf("wide reducing butterfly q=%llu w=%llu case %d\n", (unsigned long long)q,
                               (unsigned long long)w, k);
                        return 1;
                    }
                    // wide schedule, layer without reduction: x < 12q
                    u64 nx = operand(12 * q, q, k % 8), ny = operand(16 * q, q, (k / 8) % 8);
                    const u64 nsx = nx % q, nwy = mulmod(ny % q, w, q);
                    butterfly<false>(nx, ny, ulonglong2{w, wq}, four_q, neg_q, false, 0 - 8 * q);
                    if (nx >= 16 * q || ny >= 16 * q || nx % q != (nsx + nwy) % q || ny % q != (nsx + q - nwy) % q) {
                        printf("wide butterfly q=%llu w=%llu case %d\n", (unsigned long long)q, (unsigned long long)w, k);
                        return 1;
                    }
                    // end of a wide pass
                    const u64 v = operand(16 * q, q, k % 8), e = sub_if_ge(v, 8 * q);
                    const u64 en = sub_if_ge_neg(v, 0 - 8 * q);
                    if (e >= 8 * q || e % q != v % q || en != e) {
                        printf("sub_if_ge q=%llu v=%llu\n", (unsigned long long)q, (unsigned long long)v);
                        return 1;
                    }
                    // product sums: 31 terms below 8q or 63 below 4q, then a correction
                    for (int bound = 8; bound >= 4; bound -= 4) {
                        shc::Acc3 acc{0, 0, 0};
                        wide exact = 0;
                        const int terms = bound == 8 ? 31 : 63;
                        for (int m = 0; m < terms; m++) {
                            // case 0: every term at its maximum
                            const int kc = (k * 3 + m) % 8;
                            const u64 t = k == 0 ? bound * q - 1 : operand(bound * q, q, (k + m) % 8);
                            const u64 kw = k == 0 || kc == 3 ? q - 1 : operand(q, q, kc);
                            shc::mac(acc, t, kw);
                            exact += (wide)t * kw;
                        }
                        const u64 corr = k % 2 ? ~0ull : next();
                        shc::acc_add(acc, corr);
                        exact += corr;
                        const shc::u128 got = shc::acc_value(acc);
                        if (got.lo != (u64)exact || got.hi != (u64)(exact >> 64)) {
                            printf("product sum q=%llu bound %dq case %d\n", (unsigned long long)q, bound, k);
                            return 1;
                        }
                    }
                    cases += 5;
                }
                cases += 4;
            }
        }
    }
    printf("ok %ld\n", cases);
    return 0;
}
