// ntt.hip -- batched negacyclic NTT / inverse NTT for gfx950.
//
// Same transform as the reference's nttNegacyclicHarvey / inverseNttNegacyclicHarvey
// (src/utils/ntt.cpp:158-213, src/utils/dwthandler.h:88-372; GPU launchers
// src/kernelutils.cu:330-476), scheduled for CDNA4 instead of one launch per layer:
//
//   * A length-2^logn limb is viewed as a 2^K1 x 2^K2 matrix.  The first K1 layers
//     act along columns, the last K2 layers inside contiguous rows, so a forward
//     transform is two kernels (columns pass, rows pass) and each reads and writes
//     the limb once.  Up to 2^12 points the whole limb is one tile (one kernel).
//   * A tile lives in LDS; each thread pulls 2^E values into VGPRs and runs E
//     butterfly layers there before writing them back, so one LDS round trip
//     covers E layers.
//   * Harvey-style lazy butterflies with Shoup twiddles keep values in [0,8q)
//     (forward) or [0,4q) (inverse); the twiddle product uses a quotient estimate
//     that may be low by up to 2 (9 32-bit multiplies instead of 10), which those
//     wider ranges absorb.  The full reduction and the N^-1 scaling are folded into
//     the store of the last pass.
//   * The key switch's ModUp transforms take the wide schedule, which needs q < 2^60
//     (contexts hold only such primes): values may grow to 16q inside a pass (local
//     layers 0, 1 and odd layers skip the conditional subtraction, even layers >= 2
//     subtract 8q) and one subtraction of 8q at the end of the pass brings them back
//     below 8q, so passes still hand each other [0, 8q).
//
// Twiddle indexing: forward layer l (m = 2^l groups) uses root_powers[m + group];
// inverse layer l uses inv_root_powers[N - 2m + 1 + group], group = x >> (logn - l).
#include "ntt.h"
#include "accum.h"
#include "butterfly.h"
#include <vector>

namespace shc {
namespace {

// Twiddle tables are global memory; saying so lets the compiler use global loads
// (flat loads would also hold up the LDS waits).
__device__ __forceinline__ ulonglong2 load_tw(const ulonglong2 *tw, int i) {
    typedef uint64_t g2 __attribute__((ext_vector_type(2)));
    const g2 v = ((const __attribute__((address_space(1))) g2 *)tw)[i];
    return ulonglong2{v.x, v.y};
}

// Position inside a 2^K-point sub-transform of register j of thread t, for the phase
// that runs local layers [S0, S0+EP): the EP butterfly bits sit at bit positions
// [K-S0-EP, K-S0) of the index and come from the low EP bits of j.
template <int K, int E, int S0, int EP>
__device__ __forceinline__ int phase_index(int t, int j) {
    constexpr int LOWB = K - S0 - EP;
    const int jl = j & ((1 << EP) - 1);
    const int jh = j >> EP;
    const int other = (t << (E - EP)) | jh;
    return ((other >> LOWB) << (K - S0)) | (jl << LOWB) | (other & ((1 << LOWB) - 1));
}

// A prime's butterfly constants: 4q, 2^64 - q, 2^64 - 4q, 2^64 - 8q.
struct BfConst {
    uint64_t four_q, neg_q, neg_four_q, neg_eight_q;
};
__device__ __forceinline__ BfConst bf_const(const DevPrime &P) {
    return BfConst{P.q << 2, 0 - P.q, P.neg_four_q, P.neg_eight_q};
}

// One phase on the thread's registers.  g = index of this sub-transform among the
// 2^L0 sub-transforms the full transform splits into at global layer L0.  W selects
// the wide forward schedule (q < 2^60).
template <int LOGN, int K, int E, int S0, int EP, bool INV, bool W>
__device__ __forceinline__ void phase_compute(uint64_t (&r)[1 << E], int t, int g, int L0,
                                              const ulonglong2 *__restrict__ tw, const BfConst &bc) {
#pragma unroll
    for (int ss = 0; ss < EP; ss++) {
        const int s = INV ? (S0 + EP - 1 - ss) : (S0 + ss);
        const int b = S0 + EP - 1 - s;
#pragma unroll
        for (int j = 0; j < (1 << E); j++) {
            if (j & (1 << b)) continue;
            // in the first phase a register's group does not depend on the thread (its
            // index bits above K - s all come from j): with t = 0 it is a constant, so
            // the twiddle address is uniform and the load scalar
            const int idx = phase_index<K, E, S0, EP>(S0 == 0 ? 0 : t, j);
            const int grp = (g << s) + (idx >> (K - s));
            const int l = L0 + s;
            const ulonglong2 w = load_tw(tw, INV ? (1 << LOGN) - (2 << l) + 1 + grp : (1 << l) + grp);
            const bool red = !W || (s >= 2 && !(s & 1));
            butterfly<INV>(r[j], r[j | (1 << b)], w, bc.four_q, bc.neg_q, red, W ? bc.neg_eight_q : bc.neg_four_q);
        }
    }
}

// Runs all phases of a sub-transform.  The first phase reads its registers straight
// from global memory through `load(idx, j)` (j: the register) and the last writes them
// back through
// `store(idx, value)`; in between, values cross threads through LDS at `addr(idx)`,
// so a pass makes one LDS round trip per phase boundary and one barrier each.
template <int LOGN, int K, int E, bool INV, bool W, int P = 0, class Addr, class Load, class Store>
__device__ __forceinline__ void run_phases(uint64_t *tile, const Addr &addr, const Load &load, const Store &store,
                                           int t, int g, int L0, const ulonglong2 *__restrict__ tw,
                                           const BfConst &bc, bool active) {
    constexpr int NP = (K + E - 1) / E;
    if constexpr (P < NP) {
        // forward runs phases from layer 0 upwards, inverse from the top layer down
        constexpr int PI = INV ? (NP - 1 - P) : P;
        constexpr int S0 = PI * E;
        constexpr int EP = (S0 + E <= K) ? E : (K - S0);
        uint64_t r[1 << E];
        if (active) {
#pragma unroll
            for (int j = 0; j < (1 << E); j++) {
                const int idx = phase_index<K, E, S0, EP>(t, j);
                r[j] = P == 0 ? load(idx, j) : tile[addr(idx)];
            }
            phase_compute<LOGN, K, E, S0, EP, INV, W>(r, t, g, L0, tw, bc);
#pragma unroll
            for (int j = 0; j < (1 << E); j++) {
                const int idx = phase_index<K, E, S0, EP>(t, j);
                if (P == NP - 1) store(idx, W ? sub_if_ge_neg(r[j], bc.neg_eight_q) : r[j]);  // wide: [0,16q) -> [0,8q)
                else tile[addr(idx)] = r[j];
            }
        }
        if constexpr (P + 1 < NP) {
            __syncthreads();
            run_phases<LOGN, K, E, INV, W, P + 1>(tile, addr, load, store, t, g, L0, tw, bc, active);
        }
    }
}

// LDS slot of element idx of a row: XOR-swizzled by idx's high nibble so that both
// the "stride 16" and the "16 consecutive" phase patterns hit distinct banks.
__host__ __device__ constexpr int swz(int idx) { return idx ^ ((idx >> 4) & 15); }

// Final touch of a value leaving the last pass: forward [0, 8q) -> [0, 4q) (lazy) or
// [0, q); inverse [0, 4q) times N^-1 (exact Shoup) -> [0, 2q) (lazy) or [0, q).
template <bool INV, bool FULL>
__device__ __forceinline__ uint64_t finish(uint64_t x, const DevPrime &P) {
    if (!INV) {
        const uint64_t four_q = P.q << 2;
        x = x >= four_q ? x - four_q : x;
        if (FULL) {
            x = x >= P.two_q ? x - P.two_q : x;
            x = x >= P.q ? x - P.q : x;
        }
    } else {
        x = shoup_lazy(x, P.inv_n, P.inv_n_shoup, P.q);
        if (FULL) x = x >= P.q ? x - P.q : x;
    }
    return x;
}

// Rows pass: one block takes row g (the last K layers; the whole limb when K == LOGN)
// of SUBS limbs that share prime i -- limbs (i, j) for SUBS consecutive j -- so every
// row in the block uses the same twiddles (broadcast loads, L1/L2 hits).  Blocks run
// with the limb group fastest, so consecutive blocks also share them.  jn_short != 0:
// primes i < I - 1 have only jn_short limbs (the grid is sized for jn).
// MX (fewer than SUBS limbs per prime, e.g. one ciphertext): a block takes SUBS
// (limb, row) pairs of prime i instead, limb fastest, so that its threads are not idle;
// the rows of a block then differ and so do their twiddles.
template <int LOGN, int K, int E, int SUBS>
__host__ __device__ constexpr uint32_t rows_blocks_per_prime(uint32_t jn, bool mx) {
    return mx ? ((jn << (LOGN - K)) + SUBS - 1) / SUBS : (1u << (LOGN - K)) * ((jn + SUBS - 1) / SUBS);
}
template <int LOGN, int K, int SUBS>
__host__ __device__ constexpr bool rows_mixed(uint32_t jn) {
    return LOGN > K && jn < (uint32_t)SUBS;
}

// (tile: SUBS << K words of LDS, the caller's)
template <int LOGN, int K, int E, int SUBS, bool INV, bool W = false, bool MX = false, class Load, class Store>
__device__ __forceinline__ void rows_body_t(uint64_t *tile, uint32_t bid, const DevPrime *__restrict__ primes,
                                            uint32_t I, uint32_t j0, uint32_t jn, const Load &load_elem,
                                            const Store &store_elem, uint32_t jn_short = 0) {
    constexpr int T = 1 << (K - E);
    constexpr int SZ = 1 << K;
    constexpr int RPL = LOGN - K;  // log2(rows per limb)
    constexpr int NP = (K + E - 1) / E;
    static_assert(!MX || (NP > 1 && RPL > 0), "pairs of (limb, row) need rows and a staged store");
    const int tid = threadIdx.x;
    const int sub = tid / T, t = tid % T;
    uint32_t i, jfirst = 0, q0 = 0;
    int g, nsub;
    bool active;
    uint32_t j;
    if constexpr (MX) {
        const uint32_t per = rows_blocks_per_prime<LOGN, K, E, SUBS>(jn, true);
        i = bid / per;
        q0 = (bid % per) * SUBS;
        nsub = (int)min<uint32_t>(SUBS, (jn << RPL) - q0);
        active = sub < nsub;
        const uint32_t q = q0 + (uint32_t)(active ? sub : 0);
        j = j0 + q % jn;
        g = (int)(q / jn);
    } else {
        const uint32_t jc_count = (jn + SUBS - 1) / SUBS;
        const uint32_t jc = bid % jc_count;
        const uint32_t r = bid / jc_count;
        g = (int)(r & ((1u << RPL) - 1));
        i = r >> RPL;
        jfirst = j0 + jc * SUBS;
        const uint32_t jend = j0 + (jn_short && i + 1 < I ? jn_short : jn);
        if (jfirst >= jend) return;  // uniform: a whole block past a short row
        nsub = (int)min<uint32_t>(SUBS, jend - jfirst);
        active = sub < nsub;
        j = jfirst + (uint32_t)(active ? sub : 0);
    }
    const size_t row_off = (size_t)g << K;
    const DevPrime &P = primes[i];
    const ulonglong2 *tw = INV ? P.inv : P.fwd;
    auto addr = [sub](int idx) { return sub * SZ + swz(idx); };
    auto load = [&](int idx, int) { return load_elem(i, j, row_off + idx); };
    if constexpr (NP == 1) {
        auto store = [&](int idx, uint64_t x) { store_elem(i, j, row_off + idx, x); };
        run_phases<LOGN, K, E, INV, W>(tile, addr, load, store, t, g, RPL, tw, bf_const(P), active);
    } else {
        // the last phase leaves 16 consecutive values per thread; stage them in LDS so
        // that the block writes whole rows with consecutive lanes on consecutive words
        auto stage = [&](int idx, uint64_t x) { tile[addr(idx)] = x; };
        run_phases<LOGN, K, E, INV, W>(tile, addr, load, stage, t, g, RPL, tw, bf_const(P), active);
        __syncthreads();
        if constexpr (MX) {
            for (int e = tid; e < nsub * SZ; e += SUBS * T) {
                const int sb = e / SZ, off = e % SZ;
                const uint32_t q = q0 + (uint32_t)sb;
                store_elem(i, j0 + q % jn, ((size_t)(q / jn) << K) + off, tile[sb * SZ + swz(off)]);
            }
        } else if (nsub == SUBS) {
            // a full block: a fixed number of elements per thread, unrolled so that every
            // row's base address is uniform (scalar) and only the offset varies by lane
            constexpr int PER = SUBS * SZ / (SUBS * T);
#pragma unroll
            for (int k = 0; k < PER; k++) {
                const int e = tid + k * SUBS * T;
                const int sb = e / SZ, off = e % SZ;
                store_elem(i, jfirst + (uint32_t)sb, row_off + off, tile[sb * SZ + swz(off)]);
            }
        } else {
            for (int e = tid; e < nsub * SZ; e += SUBS * T) {
                const int sb = e / SZ, off = e % SZ;
                store_elem(i, jfirst + (uint32_t)sb, row_off + off, tile[sb * SZ + swz(off)]);
            }
        }
    }
}

template <int LOGN, int K, int E, int SUBS, bool INV, bool W = false, bool MX = false, class Load, class Store>
__device__ __forceinline__ void rows_body(uint32_t bid, const DevPrime *__restrict__ primes, uint32_t I, uint32_t j0,
                                          uint32_t jn, const Load &load_elem, const Store &store_elem,
                                          uint32_t jn_short = 0) {
    constexpr int NP = (K + E - 1) / E;
    __shared__ uint64_t tile[NP > 1 ? SUBS << K : 1];
    rows_body_t<LOGN, K, E, SUBS, INV, W, MX>(tile, bid, primes, I, j0, jn, load_elem, store_elem, jn_short);
}

// Columns pass: SUBS adjacent columns of one limb per block; a column is the first
// K layers of the limb (stride 2^(LOGN-K) between its values).
// Blocks run limb (i, j) by limb with j fastest, or with i fastest when I_FAST (then I
// is the prime count) -- the order for loads shared by all primes of one j.
template <int LOGN, int K, int E, int SUBS, bool INV, bool I_FAST = false, class Load, class Store>
__device__ __forceinline__ void cols_body_t(uint64_t *tile, uint32_t bid, const DevPrime *__restrict__ primes,
                                            uint32_t j0, uint32_t jn, const Load &load_elem, const Store &store_elem,
                                            uint32_t I = 1) {
    constexpr int C = 1 << (LOGN - K);
    constexpr int TILES = C / SUBS;
    const int tid = threadIdx.x;
    const int col = tid % SUBS, t = tid / SUBS;
    const uint32_t r = bid / TILES;
    const size_t c = (size_t)(bid % TILES) * SUBS + col;
    const uint32_t i = I_FAST ? r % I : r / jn, j = j0 + (I_FAST ? r / I : r % jn);
    const DevPrime &P = primes[i];
    const ulonglong2 *tw = INV ? P.inv : P.fwd;
    auto addr = [col](int idx) { return idx * SUBS + col; };
    auto load = [&](int idx, int) { return load_elem(i, j, (size_t)idx * C + c); };
    auto store = [&](int idx, uint64_t x) { store_elem(i, j, (size_t)idx * C + c, x); };
    run_phases<LOGN, K, E, INV, false>(tile, addr, load, store, t, 0, 0, tw, bf_const(P), true);
}

template <int LOGN, int K, int E, int SUBS, bool INV, bool I_FAST = false, class Load, class Store>
__device__ __forceinline__ void cols_body(uint32_t bid, const DevPrime *__restrict__ primes, uint32_t j0, uint32_t jn,
                                          const Load &load_elem, const Store &store_elem, uint32_t I = 1) {
    __shared__ uint64_t tile[SUBS << K];
    cols_body_t<LOGN, K, E, SUBS, INV, I_FAST>(tile, bid, primes, j0, jn, load_elem, store_elem, I);
}

// The key switch's ModUp rows pass fused with the products by the key (N = 2^15, 2^16):
// a block owns RW adjacent rows of target prime i for CG ciphertexts (NS = CG x RW
// slots).  Each step transforms SUBS rows of T -- NS slots x SUBS/NS digits, the last K
// layers with the wide schedule -- stages them in LDS, and then every thread multiplies
// its coefficient of those rows by the key words of their digits and adds into
// three-word sums (one per slot and key component), so the transformed limbs never
// return to HBM:
//   P[ct][k][i][c] = sum_j NTT_i(T[i][ct][j])[c] * key[j][k][ki][c]   (S for i == L).
// With diag (CKKS, T without its own-prime limbs), the term j == i is the NTT-form input
// limb diag[ct * diag_stride + i * N].  The groups start at ciphertext ct_base.  A step
// past the last digit, or a ciphertext past the batch, transforms a repeated row and
// does not add or store it.  Groups of 4 and 2 ciphertexts take one row (RW = 1); a
// single ciphertext takes two rows (RW = 2), so that a step holds 8 of its digits
// instead of 16 (24 digits in 3 steps, not 2 with a third of the rows wasted).  With
// RW = 2 the slots run slot-major (each wave's rows share one row index, so its twiddle
// loads stay scalar).  Blocks of one tile run one after another on one XCD, so its key
// slice comes from that XCD's L2.
// Sum bound: up to 31 digits the transformed words (< 8q) enter the sums directly
// (31 * 8q^2 + q^2 < 2^128 for q < 2^60); beyond, R4 takes them below 4q first.
template <int LOGN, int K, int E, int SUBS, int CG, int RW, bool R4>
__device__ __forceinline__ void rows_keyprod_body(
    uint32_t bid, const uint64_t *__restrict__ T, const uint64_t *__restrict__ key, uint64_t *__restrict__ Pout,
    uint64_t *__restrict__ Sout, const DevPrime *__restrict__ primes, const uint64_t *__restrict__ diag,
    size_t diag_stride, uint32_t count, uint32_t L, uint32_t K_key, uint32_t ct_base, uint32_t groups,
    uint32_t tiles, const KsFold &fold) {
    constexpr int TT = 1 << (K - E);  // threads per row
    constexpr int SZ = 1 << K;        // row length
    static_assert(SZ == SUBS * TT, "one coefficient per thread in the product phase");
    constexpr int NS = CG * RW;
    static_assert(SUBS % NS == 0, "whole digits per step");
    constexpr int RPL = LOGN - K;
    constexpr int GPL = (1 << RPL) / RW;  // row groups per limb
    static_assert(GPL * RW == (1 << RPL), "whole row groups");
    constexpr int DJ = SUBS / NS;  // digits per step
    static_assert(RW == 1 || DJ % (64 / TT) == 0, "slot-major: one slot per wave");
    __shared__ uint64_t tile[SUBS * SZ];
    // bid & 7 is the XCD (launches, and groups of a grouped launch, start at multiples of 8)
    const uint32_t xcd = bid & 7, kk = bid >> 3;
    const uint32_t tl = (kk / groups) * 8 + xcd;
    if (tl >= tiles) return;  // uniform: the whole block
    const uint32_t grp = kk % groups;
    const uint32_t i = tl / GPL;
    const int g0 = (int)(tl % GPL) * RW;  // first row of the group
    const uint32_t ct0 = ct_base + grp * CG;
    const uint32_t gv = min<uint32_t>(CG, count - ct0);
    const bool short_row = diag && i < L;
    const uint32_t Lt = short_row ? L - 1 : L;
    const int tid = threadIdx.x;
    const int sub = tid / TT, t = tid % TT;
    // this thread's transform row: slot gg (ciphertext gg % CG, row gg / CG), digit dj;
    // tile row of (slot c, digit d): slot(c, d)
    const int gg = RW > 1 ? sub / DJ : sub % NS, dj = RW > 1 ? sub % DJ : sub / NS;
    auto slot = [](int c, int d) { return RW > 1 ? c * DJ + d : d * NS + c; };
    const size_t n = (size_t)1 << LOGN;
    auto row_off = [g0](int r) { return (size_t)(g0 + r) << K; };
    // the row index is uniform across a wave (one slot per wave when RW > 1)
    const int g_t = RW > 1 ? __builtin_amdgcn_readfirstlane(g0 + gg / CG) : g0;
    const DevPrime &P = primes[i];
    const BfConst bc = bf_const(P);
    const uint32_t ki = i == L ? K_key - 1 : i;
    const size_t kstride = (size_t)K_key << LOGN;  // key[j][k] -> key[j][k+1]
    const uint64_t *kp = key + ((size_t)ki << LOGN) + row_off(0) + tid;  // row r: + r * SZ
    Acc3 a0[NS], a1[NS];
#pragma unroll
    for (int c = 0; c < NS; c++) a0[c] = a1[c] = Acc3{0, 0, 0};
    const uint64_t *trow = T + (((size_t)i * count * L + (size_t)(ct0 + min<uint32_t>(gg % CG, gv - 1)) * Lt) << LOGN) +
                           row_off(gg / CG);
    const uint32_t steps = (Lt + DJ - 1) / DJ;
    for (uint32_t s = 0; s < steps; s++) {
        const uint64_t *row = trow + ((size_t)min<uint32_t>(s * DJ + dj, Lt - 1) << LOGN);
        auto load = [row](int idx, int) { return row[idx]; };
        if (s) __syncthreads();  // the previous step's products have read the tile
        // the twiddles are the same in every step; reloading them (L1/L2 hits) keeps them
        // from being held in registers across steps
        const ulonglong2 *tw = P.fwd;
        asm volatile("" : "+s"(tw));
        // small groups (short, latency-bound launches) also recompute their LDS addresses
        // each step, for fewer registers and more waves; a group of 4 keeps them (fewer
        // instructions; measured faster)
        int sub_s = sub, t_s = t;
        if constexpr (CG < 4) asm volatile("" : "+v"(sub_s), "+v"(t_s));
        auto addr_s = [sub_s](int idx) { return sub_s * SZ + swz(idx); };
        auto stage_s = [&](int idx, uint64_t x) { tile[addr_s(idx)] = x; };
        // rows past the last digit (the last step of a short target) are not transformed:
        // with a group of 4 that is a whole wave, which leaves the SIMD to other waves
        run_phases<LOGN, K, E, false, true>(tile, addr_s, load, stage_s, t_s, g_t, RPL, tw, bc, s * DJ + dj < Lt);
        // the key words of a digit are requested one digit ahead (the first before the
        // barrier), at clamped indices rather than behind a branch
        auto key_j = [&](uint32_t jt) {
            jt = min<uint32_t>(jt, Lt - 1);
            return short_row && jt >= i ? jt + 1 : jt;
        };
        uint64_t kn0[RW], kn1[RW];
        auto fetch_keys = [&](uint32_t j) {
#pragma unroll
            for (int r = 0; r < RW; r++) {
                kn0[r] = kp[(2 * j) * kstride + r * SZ];
                kn1[r] = kp[(2 * j + 1) * kstride + r * SZ];
            }
        };
        fetch_keys(key_j(s * DJ));
        __syncthreads();
#pragma unroll 1
        for (int d = 0; d < DJ; d++) {
            const uint32_t jt = s * DJ + d;
            if (jt >= Lt) break;  // uniform
            uint64_t k0[RW], k1[RW];
#pragma unroll
            for (int r = 0; r < RW; r++) {
                k0[r] = kn0[r];
                k1[r] = kn1[r];
            }
            fetch_keys(key_j(jt + 1));
#pragma unroll
            for (int c = 0; c < NS; c++) {
                uint64_t x = tile[slot(c, d) * SZ + swz(tid)];
                if (R4) x = sub_if_ge_neg(x, bc.neg_four_q);
                mac(a0[c], x, k0[c / CG]);
                mac(a1[c], x, k1[c / CG]);
            }
        }
    }
    if (short_row) {
#pragma unroll
        for (int c = 0; c < NS; c++) {
            const uint64_t k0 = kp[(2 * i) * kstride + (c / CG) * SZ], k1 = kp[(2 * i + 1) * kstride + (c / CG) * SZ];
            const uint64_t x = diag[(ct0 + min<uint32_t>(c % CG, gv - 1)) * diag_stride + ((size_t)i << LOGN) +
                                    row_off(c / CG) + tid];
            mac(a0[c], x, k0);
            mac(a1[c], x, k1);
        }
    }
    // the ciphertext the switch adds into, times p (CKKS, i < L): one more term per sum,
    // and the ModDown reads no ciphertext
    if (fold.ct && i < L) {
        const uint64_t pm = fold.dc[i].p_mod;
#pragma unroll
        for (int c = 0; c < NS; c++) {
            const uint64_t *cw = fold.ct + (size_t)(ct0 + min<uint32_t>(c % CG, gv - 1)) * fold.ct_stride +
                                 ((size_t)i << LOGN) + row_off(c / CG) + tid;
            mac<true>(a0[c], cw[0], pm);
            if (!fold.c1_zero) mac<true>(a1[c], cw[(size_t)L << LOGN], pm);
        }
    }
#pragma unroll
    for (int c = 0; c < NS; c++) {
        if ((uint32_t)(c % CG) >= gv) continue;
        const size_t ct = ct0 + c % CG;
        const size_t off = row_off(c / CG) + tid;
        const uint64_t r0 = acc_reduce(a0[c], P), r1 = acc_reduce(a1[c], P);
        if (i == L) {
            Sout[(ct * 2 + 0) * n + off] = r0;
            Sout[(ct * 2 + 1) * n + off] = r1;
        } else {
            Pout[((ct * 2 + 0) * L + i) * n + off] = r0;
            Pout[((ct * 2 + 1) * L + i) * n + off] = r1;
        }
    }
}

template <int LOGN, int K, int E, int SUBS, int CG, int RW, bool R4>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_keyprod_kernel(
    const uint64_t *__restrict__ T, const uint64_t *__restrict__ key, uint64_t *__restrict__ Pout,
    uint64_t *__restrict__ Sout, const DevPrime *__restrict__ primes, const uint64_t *__restrict__ diag,
    size_t diag_stride, uint32_t count, uint32_t L, uint32_t K_key, uint32_t ct_base, uint32_t groups,
    uint32_t tiles, KsFold fold) {
    rows_keyprod_body<LOGN, K, E, SUBS, CG, RW, R4>(blockIdx.x, T, key, Pout, Sout, primes, diag, diag_stride, count,
                                                    L, K_key, ct_base, groups, tiles, fold);
}

// ---- grouped launches: ciphertexts at several levels in one launch per stage.  Group
// g owns blocks [start[g], start[g+1]) and runs the one-level body with its own
// arguments on block (blockIdx.x - start[g]); the argument tables travel by value (in
// the kernel-argument segment; the group index is uniform, so they are scalar loads).
struct KeyprodGroup {
    const uint64_t *T;
    uint64_t *P, *S;
    const DevPrime *primes;
    const uint64_t *diag;
    size_t diag_stride;
    uint32_t count, L, ct_base, groups, tiles;
    KsFold fold;
};
struct KeyprodGroups {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    const uint64_t *key;
    uint32_t K_key;
    KeyprodGroup g[MAX_GROUPS];
};
__device__ __forceinline__ uint32_t group_of(const uint32_t *start, uint32_t n, uint32_t b) {
    uint32_t g = 0;
#pragma unroll 1
    while (g + 1 < n && b >= start[g + 1]) g++;
    return g;
}

template <int LOGN, int K, int E, int SUBS, int CG, int RW, bool R4>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_keyprod_grouped_kernel(KeyprodGroups kg) {
    const uint32_t gi = group_of(kg.start, kg.n, blockIdx.x);
    const KeyprodGroup &a = kg.g[gi];
    rows_keyprod_body<LOGN, K, E, SUBS, CG, RW, R4>(blockIdx.x - kg.start[gi], a.T, kg.key, a.P, a.S, a.primes,
                                                    a.diag, a.diag_stride, a.count, a.L, kg.K_key, a.ct_base,
                                                    a.groups, a.tiles, a.fold);
}

// Plain in-place passes over a LimbGrid buffer.
struct GridAccess {
    uint64_t *data;
    LimbGrid lg;
    int logn;
    __device__ uint64_t *at(uint32_t i, uint32_t j, size_t off) const {
        return data + ((size_t)(i * lg.A + j * lg.B) << logn) + off;
    }
};

template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL, bool MX = false>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_kernel(uint64_t *__restrict__ data,
                                                                  const DevPrime *__restrict__ primes, LimbGrid lg,
                                                                  uint32_t j0, uint32_t jn) {
    const GridAccess g{data, lg, LOGN};
    rows_body<LOGN, K, E, SUBS, INV, false, MX>(
        blockIdx.x, primes, lg.I, j0, jn, [&](uint32_t i, uint32_t j, size_t off) { return *g.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

// Rows pass of a transform whose input lies elsewhere: reads limb (i, j) from `src`
// (grid slg) and writes it to `data` (grid lg), same shape.
template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL, bool MX = false>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_from_kernel(const uint64_t *__restrict__ src,
                                                                       LimbGrid slg, uint64_t *__restrict__ data,
                                                                       const DevPrime *__restrict__ primes,
                                                                       LimbGrid lg, uint32_t j0, uint32_t jn) {
    const GridAccess g{data, lg, LOGN};
    const GridAccess s{const_cast<uint64_t *>(src), slg, LOGN};
    rows_body<LOGN, K, E, SUBS, INV, false, MX>(
        blockIdx.x, primes, lg.I, j0, jn, [&](uint32_t i, uint32_t j, size_t off) { return *s.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

// The key switch's ModUp rows pass: the same code as a kernel of its own, so that
// profiles tell it apart from the other forward transforms.  Wide schedule; outputs
// left in [0, 4q) (FIN) or [0, 8q), ranges the key products' 128-bit sums absorb.
// jn_short: see rows_body (targets without their own digit's limb).
template <int LOGN, int K, int E, int SUBS, bool FIN>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_modup_kernel(uint64_t *__restrict__ data,
                                                                        const DevPrime *__restrict__ primes,
                                                                        LimbGrid lg, uint32_t jn_short) {
    const GridAccess g{data, lg, LOGN};
    rows_body<LOGN, K, E, SUBS, false, true>(
        blockIdx.x, primes, lg.I, 0, lg.J, [&](uint32_t i, uint32_t j, size_t off) { return *g.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = FIN ? finish<false, false>(x, primes[i]) : x;
        },
        jn_short);
}

template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_kernel(uint64_t *__restrict__ data,
                                                                  const DevPrime *__restrict__ primes, LimbGrid lg,
                                                                  uint32_t j0, uint32_t jn) {
    const GridAccess g{data, lg, LOGN};
    cols_body<LOGN, K, E, SUBS, INV>(
        blockIdx.x, primes, j0, jn, [&](uint32_t i, uint32_t j, size_t off) { return *g.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

// ModUp columns pass into T[i][ct][j] (target-major, so that the rows pass walks one
// prime's limbs at a time): a block takes one tile of 16 columns of digit j of
// ciphertext ct (read from d[ct][j]), keeps it in registers, and for each of tc target
// primes i in turn reduces it modulo q_i when the digit's prime exceeds 8 q_i (the
// forward butterflies take inputs < 8q) and runs the first K layers -- one LDS round
// trip and two barriers per target -- so the loads' latency is paid once per tile and
// chunk of targets, not once per target.  Large batches take all L + 1 targets per block
// (measured 0.6 ms faster per 128 ciphertexts than one block per target); small ones
// split the targets into chunks so that the grid still fills the GPU, down to one target
// per block (ONE: nothing held in registers, 60 VGPRs instead of 132).
// skip_diag: limb (i, i) is not produced (its NTT is the NTT-form input limb) and the
// targets i < L keep L - 1 limbs per ciphertext, T[i][ct][j - (j > i)].
template <int LOGN, int K, int E, int SUBS, bool ONE>
__device__ __forceinline__ void cols_modup_body(uint32_t bid, const uint64_t *__restrict__ d, size_t d_ct_stride,
                                                uint64_t *__restrict__ T, const DevPrime *__restrict__ targets,
                                                const DevPrime *__restrict__ digits, uint32_t count, uint32_t L,
                                                bool skip_diag, uint32_t tc) {
    static_assert(K >= E, "the first phase fills all 2^E registers");
    constexpr int C = 1 << (LOGN - K);
    constexpr int TILES = C / SUBS;
    __shared__ uint64_t tile[SUBS * (1 << K)];
    const int tid = threadIdx.x;
    const int col = tid % SUBS, t = tid / SUBS;
    const uint32_t chunks = (L + tc) / tc;  // ceil((L + 1) / tc)
    const uint32_t rc = bid / TILES;  // (ct * L + j) * chunks + chunk
    const uint32_t r = rc / chunks;          // ct * L + j
    const uint32_t i0 = (rc % chunks) * tc, i1 = min<uint32_t>(i0 + tc, L + 1);
    const int c0 = (int)(bid % TILES) * SUBS;
    const uint32_t j = r % L;
    const size_t ct = r / L;
    const uint64_t *in = d + ct * d_ct_stride + ((size_t)j << LOGN) + c0 + col;
    const uint64_t qj = digits[j].q;
    auto addr = [col](int idx) { return idx * SUBS + col; };
    if constexpr (ONE) {
        // one target per block (small batches): the first phase loads straight from the
        // digit, nothing held across targets (fewer registers, more waves)
        const uint32_t i = i0;
        if (skip_diag && i == j) return;  // uniform: the whole block
        const bool short_row = skip_diag && i < L;
        const size_t limb = (size_t)i * count * L + ct * (short_row ? L - 1 : L) + (short_row && j > i ? j - 1 : j);
        uint64_t *out = T + (limb << LOGN) + c0 + col;
        const DevPrime &P = targets[i];
        const bool reduce = qj > 8 * P.q;
        auto load = [&](int idx, int) {
            const uint64_t x = in[(size_t)idx * C];
            return reduce ? barrett64(x, P.q, P.ratio_hi) : x;
        };
        auto store = [out](int idx, uint64_t x) { out[(size_t)idx * C] = x; };
        run_phases<LOGN, K, E, false, true>(tile, addr, load, store, t, 0, 0, P.fwd, bf_const(P), true);
        return;
    }
    uint64_t dv[1 << E];
#pragma unroll
    for (int k = 0; k < (1 << E); k++) dv[k] = in[(size_t)phase_index<K, E, 0, E>(t, k) * C];
    for (uint32_t i = i0; i < i1; i++) {
        if (skip_diag && i == j) continue;  // uniform
        const bool short_row = skip_diag && i < L;
        const size_t limb = (size_t)i * count * L + ct * (short_row ? L - 1 : L) + (short_row && j > i ? j - 1 : j);
        uint64_t *out = T + (limb << LOGN) + c0 + col;
        const DevPrime &P = targets[i];
        const bool reduce = qj > 8 * P.q;
        auto load = [&](int, int k) { return reduce ? barrett64(dv[k], P.q, P.ratio_hi) : dv[k]; };
        auto store = [out](int idx, uint64_t x) { out[(size_t)idx * C] = x; };
        __syncthreads();  // the previous target's second phase has read the tile
        run_phases<LOGN, K, E, false, true>(tile, addr, load, store, t, 0, 0, targets[i].fwd, bf_const(P), true);
    }
}

template <int LOGN, int K, int E, int SUBS, bool ONE>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_modup_kernel(const uint64_t *__restrict__ d,
                                                                        size_t d_ct_stride,
                                                                        uint64_t *__restrict__ T,
                                                                        const DevPrime *__restrict__ targets,
                                                                        const DevPrime *__restrict__ digits,
                                                                        uint32_t count, uint32_t L, bool skip_diag,
                                                                        uint32_t tc) {
    cols_modup_body<LOGN, K, E, SUBS, ONE>(blockIdx.x, d, d_ct_stride, T, targets, digits, count, L, skip_diag, tc);
}

struct ModupGroup {
    const uint64_t *d;
    size_t d_ct_stride;
    uint64_t *T;
    const DevPrime *targets;
    uint32_t count, L, tc;
    bool skip_diag;
};
struct ModupGroups {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    const DevPrime *digits;
    ModupGroup g[MAX_GROUPS];
};

template <int LOGN, int K, int E, int SUBS, bool ONE>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_modup_grouped_kernel(ModupGroups mg) {
    const uint32_t gi = group_of(mg.start, mg.n, blockIdx.x);
    const ModupGroup &a = mg.g[gi];
    cols_modup_body<LOGN, K, E, SUBS, ONE>(blockIdx.x - mg.start[gi], a.d, a.d_ct_stride, a.T, a.targets, mg.digits,
                                           a.count, a.L, a.skip_diag, a.tc);
}

// Grouped transform passes: group g's limbs (i, j) of its grid, j < lg.J, all primes of
// its grid; the first pass of an inverse reads from src when the group has one.
struct NttGroupDev {
    uint64_t *data;
    const uint64_t *src;
    LimbGrid lg, slg;
    const DevPrime *primes;
};
struct NttGroups {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    NttGroupDev g[MAX_GROUPS];
};

template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL, bool MX>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_grouped_kernel(NttGroups ng) {
    const uint32_t gi = group_of(ng.start, ng.n, blockIdx.x);
    const NttGroupDev &a = ng.g[gi];
    const GridAccess g{a.data, a.lg, LOGN};
    const GridAccess s{const_cast<uint64_t *>(a.src ? a.src : a.data), a.src ? a.slg : a.lg, LOGN};
    const DevPrime *primes = a.primes;
    rows_body<LOGN, K, E, SUBS, INV, false, MX>(
        blockIdx.x - ng.start[gi], primes, a.lg.I, 0, a.lg.J,
        [&](uint32_t i, uint32_t j, size_t off) { return *s.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

template <int LOGN, int K, int E, int SUBS, bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_grouped_kernel(NttGroups ng) {
    const uint32_t gi = group_of(ng.start, ng.n, blockIdx.x);
    const NttGroupDev &a = ng.g[gi];
    const GridAccess g{a.data, a.lg, LOGN};
    const GridAccess s{const_cast<uint64_t *>(a.src ? a.src : a.data), a.src ? a.slg : a.lg, LOGN};
    const DevPrime *primes = a.primes;
    cols_body<LOGN, K, E, SUBS, INV>(
        blockIdx.x - ng.start[gi], primes, 0, a.lg.J, [&](uint32_t i, uint32_t j, size_t off) { return *s.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

// ---- Transforms over several ring degrees (N = 2^13 .. 2^17) in one launch per pass.
// Every degree runs 256-thread blocks over a 32 KB LDS tile -- the rows pass 4096 >> K2
// rows per block, the columns pass 4096 >> K1 columns -- so blocks of every degree share
// one kernel; a group's degree is uniform over its blocks.
struct NttMultiDev {
    NttGroupDev d;
    int32_t logn;
};
struct NttMulti {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    NttMultiDev g[MAX_GROUPS];
};

template <int LOGN, bool INV, bool LAST, bool FULL>
__device__ __forceinline__ void rows_multi(uint64_t *tile, uint32_t bid, const NttGroupDev &a) {
    constexpr int K2 = LOGN - LOGN / 2;
    constexpr int RS = 4096 >> K2;
    const GridAccess g{a.data, a.lg, LOGN};
    const GridAccess s{const_cast<uint64_t *>(a.src ? a.src : a.data), a.src ? a.slg : a.lg, LOGN};
    const DevPrime *primes = a.primes;
    rows_body_t<LOGN, K2, 4, RS, INV>(
        tile, bid, primes, a.lg.I, 0, a.lg.J, [&](uint32_t i, uint32_t j, size_t off) { return *s.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

template <int LOGN, bool INV, bool LAST, bool FULL>
__device__ __forceinline__ void cols_multi(uint64_t *tile, uint32_t bid, const NttGroupDev &a) {
    constexpr int K1 = LOGN / 2;
    constexpr int CS = 4096 >> K1;
    const GridAccess g{a.data, a.lg, LOGN};
    const GridAccess s{const_cast<uint64_t *>(a.src ? a.src : a.data), a.src ? a.slg : a.lg, LOGN};
    const DevPrime *primes = a.primes;
    cols_body_t<LOGN, K1, 4, CS, INV>(
        tile, bid, primes, 0, a.lg.J, [&](uint32_t i, uint32_t j, size_t off) { return *s.at(i, j, off); },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            *g.at(i, j, off) = LAST ? finish<INV, FULL>(x, primes[i]) : x;
        });
}

template <bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(256) void ntt_rows_multi_kernel(NttMulti ng) {
    __shared__ uint64_t tile[4096];
    const uint32_t gi = group_of(ng.start, ng.n, blockIdx.x);
    const NttMultiDev &m = ng.g[gi];
    const uint32_t bid = blockIdx.x - ng.start[gi];
    switch (m.logn) {  // uniform: the whole block
    case 13: rows_multi<13, INV, LAST, FULL>(tile, bid, m.d); break;
    case 14: rows_multi<14, INV, LAST, FULL>(tile, bid, m.d); break;
    case 15: rows_multi<15, INV, LAST, FULL>(tile, bid, m.d); break;
    case 16: rows_multi<16, INV, LAST, FULL>(tile, bid, m.d); break;
    default: rows_multi<17, INV, LAST, FULL>(tile, bid, m.d); break;
    }
}

template <bool INV, bool LAST, bool FULL>
__global__ __launch_bounds__(256) void ntt_cols_multi_kernel(NttMulti ng) {
    __shared__ uint64_t tile[4096];
    const uint32_t gi = group_of(ng.start, ng.n, blockIdx.x);
    const NttMultiDev &m = ng.g[gi];
    const uint32_t bid = blockIdx.x - ng.start[gi];
    switch (m.logn) {
    case 13: cols_multi<13, INV, LAST, FULL>(tile, bid, m.d); break;
    case 14: cols_multi<14, INV, LAST, FULL>(tile, bid, m.d); break;
    case 15: cols_multi<15, INV, LAST, FULL>(tile, bid, m.d); break;
    case 16: cols_multi<16, INV, LAST, FULL>(tile, bid, m.d); break;
    default: cols_multi<17, INV, LAST, FULL>(tile, bid, m.d); break;
    }
}

// ---- ModDown by the special prime p, fused into the transforms (switch_key step 5):
//   u_i = ([s + floor(p/2)]_p mod q_i) - (floor(p/2) mod q_i)     (s: special limb, coeff)
//   ct_k[i] += (P_k[i] - u_i) * p^-1 mod q_i
// CKKS runs it as the forward NTT of u (first pass computes u from s on load, last pass
// adds into ct on store); BFV as the inverse NTT of P (last pass subtracts u and adds).
// Limb (i, j): prime i < L, j = 2 ct + k.
__device__ __forceinline__ uint64_t moddown_prep(uint64_t s, uint64_t p, const DevPrime &Q, const DivRoundConst &d) {
    uint64_t v = s + (p >> 1);
    v = v >= p ? v - p : v;
    return submod(barrett64(v, Q.q, Q.ratio_hi), d.half_mod, Q.q);
}

// Word (limb i of polynomial j, coefficient off) of a ModDown's P.
template <int LOGN>
__device__ __forceinline__ uint64_t moddown_p(const ModDownArgs &a, uint32_t i, uint32_t j, size_t off) {
    if (a.p_lpp)  // uniform
        return a.P[(size_t)(j >> 1) * a.p_ct_stride + ((size_t)((j & 1) * a.p_lpp + i) << LOGN) + off];
    return a.P[((size_t)(j * a.L + i) << LOGN) + off];
}

template <int LOGN>
__device__ __forceinline__ void moddown_add(const ModDownArgs &a, uint32_t i, uint32_t j, size_t off, uint64_t pv,
                                            uint64_t u, const DevPrime &Q) {
    uint64_t *dst = a.ct + (size_t)(j >> 1) * a.ct_stride + ((size_t)((j & 1) * a.L + i) << LOGN) + off;
    const uint64_t v = shoup(submod(pv, u, Q.q), a.dc[i].inv, a.dc[i].inv_shoup, Q.q);
    if (a.folded) *dst = v;  // uniform: the ciphertext's words are already in P
    else *dst = addmod(*dst, v, Q.q);
}

template <int LOGN, int K, int E, int SUBS>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_moddown_prep_kernel(uint64_t *__restrict__ U,
                                                                               const DevPrime *__restrict__ primes,
                                                                               ModDownArgs a, uint32_t jn) {
    // every prime reads the same special limb: i fastest keeps it in L2
    cols_body<LOGN, K, E, SUBS, false, true>(
        blockIdx.x, primes, 0, jn,
        [&](uint32_t i, uint32_t j, size_t off) {
            return moddown_prep(a.S[((size_t)j << LOGN) + off], a.p, primes[i], a.dc[i]);
        },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) { U[((size_t)(j * a.L + i) << LOGN) + off] = x; }, a.L);
}

template <int LOGN, int K, int E, int SUBS, bool MX>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_moddown_kernel(const uint64_t *__restrict__ U,
                                                                          const DevPrime *__restrict__ primes,
                                                                          ModDownArgs a, uint32_t jn) {
    rows_body<LOGN, K, E, SUBS, false, false, MX>(
        blockIdx.x, primes, a.L, 0, jn, [&](uint32_t i, uint32_t j, size_t off) { return U[((size_t)(j * a.L + i) << LOGN) + off]; },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            const DevPrime &Q = primes[i];
            moddown_add<LOGN>(a, i, j, off, moddown_p<LOGN>(a, i, j, off), finish<false, true>(x, Q), Q);
        });
}

// ---- CKKS public-key encryption fused into the error term's forward transform ----
__device__ __forceinline__ uint64_t enc_lift(int64_t v, uint64_t q) { return v < 0 ? q - (uint64_t)(-v) : (uint64_t)v; }

template <int LOGN, int K, int E, int SUBS>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_encrypt_kernel(const DevPrime *__restrict__ primes,
                                                                          EncArgs a, uint32_t jn) {
    // every prime of polynomial j reads the same error and last-limb words: i fastest
    cols_body<LOGN, K, E, SUBS, false, true>(
        blockIdx.x, primes, 0, jn,
        [&](uint32_t i, uint32_t j, size_t off) {
            const DevPrime &Q = primes[i];
            const size_t w = ((size_t)j << LOGN) + off;
            const int64_t e = a.Es[w];
            uint64_t x = enc_lift(e, Q.q);
            if (a.drop) x = submod(x, moddown_prep(addmod(a.Pl[w], enc_lift(e, a.p), a.p), a.p, Q, a.dc[i]), Q.q);
            return x;
        },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) { a.out[((size_t)(j * a.L + i) << LOGN) + off] = x; },
        a.L);
}

template <int LOGN, int K, int E, int SUBS, bool MX>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_encrypt_kernel(const DevPrime *__restrict__ primes,
                                                                          EncArgs a, uint32_t jn) {
    rows_body<LOGN, K, E, SUBS, false, false, MX>(
        blockIdx.x, primes, a.L, 0, jn,
        [&](uint32_t i, uint32_t j, size_t off) { return a.out[((size_t)(j * a.L + i) << LOGN) + off]; },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            const DevPrime &Q = primes[i];
            const uint32_t g = j >> 1, k = j & 1;
            const uint64_t pu = mulmod(a.pk[((size_t)(k * a.K + i) << LOGN) + off],
                                       a.U[((size_t)(g * a.L1 + i) << LOGN) + off], Q.q, Q.ratio_lo, Q.ratio_hi);
            uint64_t v = addmod(pu, finish<false, true>(x, Q), Q.q);
            if (a.drop) v = shoup(v, a.dc[i].inv, a.dc[i].inv_shoup, Q.q);
            if (a.plain && k == 0) v = addmod(v, a.plain[((size_t)(g * a.L + i) << LOGN) + off], Q.q);
            a.out[((size_t)(j * a.L + i) << LOGN) + off] = v;
        });
}

template <int LOGN>
void launch_encrypt(const EncArgs &a, uint32_t jn, const DevPrime *primes, hipStream_t st) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    const uint32_t col_blocks = a.L * jn * ((1u << K2) / CS);
    const bool mx = rows_mixed<LOGN, K2, RS>(jn);
    const uint32_t row_blocks = a.L * rows_blocks_per_prime<LOGN, K2, E, RS>(jn, mx);
    ntt_cols_encrypt_kernel<LOGN, K1, E, CS><<<col_blocks, CS << (K1 - E), 0, st>>>(primes, a, jn);
    if (mx)
        ntt_rows_encrypt_kernel<LOGN, K2, E, RS, true><<<row_blocks, RS << (K2 - E), 0, st>>>(primes, a, jn);
    else
        ntt_rows_encrypt_kernel<LOGN, K2, E, RS, false><<<row_blocks, RS << (K2 - E), 0, st>>>(primes, a, jn);
}

struct ModDownGroupDev {
    uint64_t *U;
    ModDownArgs a;
    uint32_t jn;
};
struct ModDownGroups {
    uint32_t n;
    uint32_t start[MAX_GROUPS + 1];
    ModDownGroupDev g[MAX_GROUPS];
};

template <int LOGN, int K, int E, int SUBS>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_moddown_prep_grouped_kernel(
    const DevPrime *__restrict__ primes, ModDownGroups mg) {
    const uint32_t gi = group_of(mg.start, mg.n, blockIdx.x);
    const ModDownGroupDev &gr = mg.g[gi];
    const ModDownArgs &a = gr.a;
    uint64_t *U = gr.U;
    cols_body<LOGN, K, E, SUBS, false, true>(
        blockIdx.x - mg.start[gi], primes, 0, gr.jn,
        [&](uint32_t i, uint32_t j, size_t off) {
            return moddown_prep(a.S[((size_t)j << LOGN) + off], a.p, primes[i], a.dc[i]);
        },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) { U[((size_t)(j * a.L + i) << LOGN) + off] = x; }, a.L);
}

template <int LOGN, int K, int E, int SUBS, bool MX>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_rows_moddown_grouped_kernel(const DevPrime *__restrict__ primes,
                                                                                  ModDownGroups mg) {
    const uint32_t gi = group_of(mg.start, mg.n, blockIdx.x);
    const ModDownGroupDev &gr = mg.g[gi];
    const ModDownArgs &a = gr.a;
    const uint64_t *U = gr.U;
    rows_body<LOGN, K, E, SUBS, false, false, MX>(
        blockIdx.x - mg.start[gi], primes, a.L, 0, gr.jn,
        [&](uint32_t i, uint32_t j, size_t off) { return U[((size_t)(j * a.L + i) << LOGN) + off]; },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            const DevPrime &Q = primes[i];
            moddown_add<LOGN>(a, i, j, off, moddown_p<LOGN>(a, i, j, off), finish<false, true>(x, Q), Q);
        });
}

template <int LOGN, int K, int E, int SUBS>
__global__ __launch_bounds__(SUBS << (K - E)) void ntt_cols_moddown_kernel(const uint64_t *__restrict__ Pt,
                                                                          const DevPrime *__restrict__ primes,
                                                                          ModDownArgs a, uint32_t jn) {
    cols_body<LOGN, K, E, SUBS, true>(
        blockIdx.x, primes, 0, jn, [&](uint32_t i, uint32_t j, size_t off) { return Pt[((size_t)(j * a.L + i) << LOGN) + off]; },
        [&](uint32_t i, uint32_t j, size_t off, uint64_t x) {
            const DevPrime &Q = primes[i];
            const uint64_t u = moddown_prep(a.S[((size_t)j << LOGN) + off], a.p, Q, a.dc[i]);
            moddown_add<LOGN>(a, i, j, off, finish<true, true>(x, Q), u, Q);
        });
}

template <int LOGN>
void launch_moddown(bool ckks, uint64_t *buf, const ModDownArgs &a, uint32_t jn, const DevPrime *primes,
                    hipStream_t st) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    const uint32_t col_blocks = a.L * jn * ((1u << K2) / CS);
    const bool mx = rows_mixed<LOGN, K2, RS>(jn);
    const uint32_t row_blocks = a.L * rows_blocks_per_prime<LOGN, K2, E, RS>(jn, mx);
    if (ckks) {
        ntt_cols_moddown_prep_kernel<LOGN, K1, E, CS><<<col_blocks, CS << (K1 - E), 0, st>>>(buf, primes, a, jn);
        if (mx)
            ntt_rows_moddown_kernel<LOGN, K2, E, RS, true><<<row_blocks, RS << (K2 - E), 0, st>>>(buf, primes, a, jn);
        else
            ntt_rows_moddown_kernel<LOGN, K2, E, RS, false><<<row_blocks, RS << (K2 - E), 0, st>>>(buf, primes, a, jn);
    } else {
        const LimbGrid lg{1, a.L, a.L, jn};
        if (mx)
            ntt_rows_kernel<LOGN, K2, E, RS, true, false, true, true>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(buf, primes, lg, 0, jn);
        else
            ntt_rows_kernel<LOGN, K2, E, RS, true, false, true, false>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(buf, primes, lg, 0, jn);
        ntt_cols_moddown_kernel<LOGN, K1, E, CS><<<col_blocks, CS << (K1 - E), 0, st>>>(buf, primes, a, jn);
    }
}

template <int LOGN, bool INV, bool FULL>
void launch_ntt(uint64_t *data, LimbGrid lg, uint32_t j0, uint32_t jn, const DevPrime *primes, hipStream_t st) {
    if constexpr (LOGN <= 12) {
        constexpr int K = LOGN;
        constexpr int E = K < 4 ? K : 4;
        constexpr int T = 1 << (K - E);
        constexpr int SUBS = (T >= 256) ? 1 : (256 / T);
        const uint32_t blocks = lg.I * ((jn + SUBS - 1) / SUBS);
        ntt_rows_kernel<LOGN, K, E, SUBS, INV, true, FULL><<<blocks, SUBS * T, 0, st>>>(data, primes, lg, j0, jn);
    } else {
        constexpr int K1 = LOGN / 2;
        constexpr int K2 = LOGN - K1;
        constexpr int E = 4;
        constexpr int CS = 16;                                    // columns per block
        constexpr int RS = (4096 >> K2) < 1 ? 1 : (4096 >> K2);  // rows per block
        const uint32_t col_blocks = lg.I * jn * ((1u << K2) / CS);
        const bool mx = rows_mixed<LOGN, K2, RS>(jn);
        const uint32_t row_blocks = lg.I * rows_blocks_per_prime<LOGN, K2, E, RS>(jn, mx);
        if (!INV) {
            ntt_cols_kernel<LOGN, K1, E, CS, false, false, FULL>
                <<<col_blocks, CS << (K1 - E), 0, st>>>(data, primes, lg, j0, jn);
            if (mx)
                ntt_rows_kernel<LOGN, K2, E, RS, false, true, FULL, true>
                    <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lg, j0, jn);
            else
                ntt_rows_kernel<LOGN, K2, E, RS, false, true, FULL, false>
                    <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lg, j0, jn);
        } else {
            if (mx)
                ntt_rows_kernel<LOGN, K2, E, RS, true, false, FULL, true>
                    <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lg, j0, jn);
            else
                ntt_rows_kernel<LOGN, K2, E, RS, true, false, FULL, false>
                    <<<row_blocks, RS << (K2 - E), 0, st>>>(data, primes, lg, j0, jn);
            ntt_cols_kernel<LOGN, K1, E, CS, true, true, FULL>
                <<<col_blocks, CS << (K1 - E), 0, st>>>(data, primes, lg, j0, jn);
        }
    }
}

// Inverse transform of src's limbs into data (rows pass from src, columns pass in place).
template <int LOGN, bool FULL>
void launch_inverse_from(const uint64_t *src, LimbGrid slg, uint64_t *data, LimbGrid lg, uint32_t j0, uint32_t jn,
                         const DevPrime *primes, hipStream_t st) {
    if constexpr (LOGN <= 12) {
        constexpr int K = LOGN;
        constexpr int E = K < 4 ? K : 4;
        constexpr int T = 1 << (K - E);
        constexpr int SUBS = (T >= 256) ? 1 : (256 / T);
        const uint32_t blocks = lg.I * ((jn + SUBS - 1) / SUBS);
        ntt_rows_from_kernel<LOGN, K, E, SUBS, true, true, FULL>
            <<<blocks, SUBS * T, 0, st>>>(src, slg, data, primes, lg, j0, jn);
    } else {
        constexpr int K1 = LOGN / 2;
        constexpr int K2 = LOGN - K1;
        constexpr int E = 4;
        constexpr int CS = 16;
        constexpr int RS = (4096 >> K2) < 1 ? 1 : (4096 >> K2);
        const uint32_t col_blocks = lg.I * jn * ((1u << K2) / CS);
        const bool mx = rows_mixed<LOGN, K2, RS>(jn);
        const uint32_t row_blocks = lg.I * rows_blocks_per_prime<LOGN, K2, E, RS>(jn, mx);
        if (mx)
            ntt_rows_from_kernel<LOGN, K2, E, RS, true, false, FULL, true>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(src, slg, data, primes, lg, j0, jn);
        else
            ntt_rows_from_kernel<LOGN, K2, E, RS, true, false, FULL, false>
                <<<row_blocks, RS << (K2 - E), 0, st>>>(src, slg, data, primes, lg, j0, jn);
        ntt_cols_kernel<LOGN, K1, E, CS, true, true, FULL>
            <<<col_blocks, CS << (K1 - E), 0, st>>>(data, primes, lg, j0, jn);
    }
}

template <bool FULL>
void dispatch_inverse_from(const uint64_t *src, LimbGrid slg, uint64_t *data, LimbGrid lg, const DevPrime *primes,
                           int logn, hipStream_t st) {
    if (logn < 1 || logn > 17) fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    if (slg.I != lg.I || slg.J != lg.J) fail(ERR_LOGIC, "transform source and destination shapes differ");
    if ((uint64_t)lg.I * lg.J == 0) return;
    for (const LimbGrid &g : {slg, lg})
        if ((uint64_t)(g.I - 1) * g.A + (uint64_t)(g.J - 1) * g.B > UINT32_MAX)
            fail(ERR_INVALID_ARGUMENT, "too many limbs for one transform call");
    const uint64_t max_limbs = (uint64_t(1) << 34) >> logn;
    if (lg.I > max_limbs) fail(ERR_INVALID_ARGUMENT, "too many primes for one transform call");
    const uint32_t step = (uint32_t)(max_limbs / lg.I < lg.J ? max_limbs / lg.I : lg.J);
    for (uint32_t j0 = 0; j0 < lg.J; j0 += step) {
        const uint32_t jn = lg.J - j0 < step ? lg.J - j0 : step;
        switch (logn) {
#define SHC_NTT_CASE(L) \
    case L: launch_inverse_from<L, FULL>(src, slg, data, lg, j0, jn, primes, st); break;
            SHC_NTT_CASE(1) SHC_NTT_CASE(2) SHC_NTT_CASE(3) SHC_NTT_CASE(4) SHC_NTT_CASE(5) SHC_NTT_CASE(6)
            SHC_NTT_CASE(7) SHC_NTT_CASE(8) SHC_NTT_CASE(9) SHC_NTT_CASE(10) SHC_NTT_CASE(11) SHC_NTT_CASE(12)
            SHC_NTT_CASE(13) SHC_NTT_CASE(14) SHC_NTT_CASE(15) SHC_NTT_CASE(16) SHC_NTT_CASE(17)
#undef SHC_NTT_CASE
        }
        launch_check("ntt launch");
    }
}

template <bool INV, bool FULL>
void dispatch_chunk(uint64_t *data, LimbGrid lg, uint32_t j0, uint32_t jn, const DevPrime *primes, int logn,
                    hipStream_t st) {
    switch (logn) {
#define SHC_NTT_CASE(L) \
    case L: launch_ntt<L, INV, FULL>(data, lg, j0, jn, primes, st); break;
        SHC_NTT_CASE(1) SHC_NTT_CASE(2) SHC_NTT_CASE(3) SHC_NTT_CASE(4) SHC_NTT_CASE(5) SHC_NTT_CASE(6)
        SHC_NTT_CASE(7) SHC_NTT_CASE(8) SHC_NTT_CASE(9) SHC_NTT_CASE(10) SHC_NTT_CASE(11) SHC_NTT_CASE(12)
        SHC_NTT_CASE(13) SHC_NTT_CASE(14) SHC_NTT_CASE(15) SHC_NTT_CASE(16) SHC_NTT_CASE(17)
#undef SHC_NTT_CASE
        default: fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    }
    launch_check("ntt launch");
}

// Splits very large batches over j so that one launch covers at most 2^34 / N limbs
// (every pass runs N/16 work-items per limb, so a launch stays below 2^30).
template <bool INV, bool FULL>
void dispatch(uint64_t *data, LimbGrid lg, const DevPrime *primes, int logn, hipStream_t st) {
    if (logn < 1 || logn > 17) fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    if ((uint64_t)lg.I * lg.J == 0) return;
    if ((uint64_t)(lg.I - 1) * lg.A + (uint64_t)(lg.J - 1) * lg.B > UINT32_MAX)
        fail(ERR_INVALID_ARGUMENT, "too many limbs for one transform call");
    const uint64_t max_limbs = (uint64_t(1) << 34) >> logn;
    if (lg.I > max_limbs) fail(ERR_INVALID_ARGUMENT, "too many primes for one transform call");
    const uint32_t step = (uint32_t)(max_limbs / lg.I < lg.J ? max_limbs / lg.I : lg.J);
    for (uint32_t j0 = 0; j0 < lg.J; j0 += step) {
        const uint32_t jn = lg.J - j0 < step ? lg.J - j0 : step;
        dispatch_chunk<INV, FULL>(data, lg, j0, jn, primes, logn, st);
    }
}

// Targets per block of the ModUp columns pass: all L + 1 when the (digit, tile) blocks
// alone fill the GPU, fewer (chunks of them) so that a small batch still has about
// 16384 blocks.
inline uint32_t modup_targets_per_block(uint32_t tile_blocks, uint32_t L, uint32_t forced = 0) {
    if (forced) return forced < L + 1 ? forced : L + 1;
    const uint32_t want = (16384 + tile_blocks - 1) / tile_blocks;
    const uint32_t chunks = want < 1 ? 1 : (want > L + 1 ? L + 1 : want);
    return (L + chunks) / chunks;  // ceil((L + 1) / chunks)
}

template <int LOGN>
void launch_modup(const uint64_t *d, size_t d_stride, uint64_t *T, uint32_t cts, uint32_t L, const DevPrime *targets,
                  const DevPrime *digits, bool skip_diag, bool reduce_4q, const KernelProbe *probe, hipStream_t st,
                  KsShape shape) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    const uint32_t tc = modup_targets_per_block(cts * L * ((1u << K2) / CS), L, shape.cols_targets);
    const uint32_t col_grid = cts * L * ((L + tc) / tc) * ((1u << K2) / CS);
    if (tc == 1)
        ntt_cols_modup_kernel<LOGN, K1, E, CS, true><<<col_grid, CS << (K1 - E), 0, st>>>(d, d_stride, T, targets, digits,
                                                                                       cts, L, skip_diag, tc);
    else
        ntt_cols_modup_kernel<LOGN, K1, E, CS, false><<<col_grid, CS << (K1 - E), 0, st>>>(d, d_stride, T, targets, digits,
                                                                                        cts, L, skip_diag, tc);
    const LimbGrid lg{cts * L, 1, L + 1, cts * L};
    const uint32_t row_blocks = ((L + 1) << K1) * ((lg.J + RS - 1) / RS);
    if (probe) SHC_HIP(hipEventRecord(probe->begin, st));
    const uint32_t jn_short = skip_diag ? cts * (L - 1) : 0;
    if (reduce_4q)
        ntt_rows_modup_kernel<LOGN, K2, E, RS, true><<<row_blocks, RS << (K2 - E), 0, st>>>(T, targets, lg, jn_short);
    else
        ntt_rows_modup_kernel<LOGN, K2, E, RS, false><<<row_blocks, RS << (K2 - E), 0, st>>>(T, targets, lg, jn_short);
    if (probe) SHC_HIP(hipEventRecord(probe->end, st));
}

}  // namespace

bool ntt_forward_modup(const uint64_t *d, size_t d_ct_stride, uint64_t *T, uint32_t count, uint32_t L,
                       const DevPrime *targets, const DevPrime *digits, int logn, bool skip_diag, bool reduce_4q,
                       hipStream_t st, const KernelProbe *probe, KsShape shape) {
    if (logn < 13 || logn > 17 || count == 0) return false;
    if ((uint64_t)count * L * (L + 1) > ((uint64_t(1) << 34) >> logn))
        fail(ERR_INVALID_ARGUMENT, "key-switch batch too large for one launch");
    switch (logn) {
        case 13: launch_modup<13>(d, d_ct_stride, T, count, L, targets, digits, skip_diag, reduce_4q, probe, st, shape); break;
        case 14: launch_modup<14>(d, d_ct_stride, T, count, L, targets, digits, skip_diag, reduce_4q, probe, st, shape); break;
        case 15: launch_modup<15>(d, d_ct_stride, T, count, L, targets, digits, skip_diag, reduce_4q, probe, st, shape); break;
        case 16: launch_modup<16>(d, d_ct_stride, T, count, L, targets, digits, skip_diag, reduce_4q, probe, st, shape); break;
        case 17: launch_modup<17>(d, d_ct_stride, T, count, L, targets, digits, skip_diag, reduce_4q, probe, st, shape); break;
    }
    launch_check("ntt modup launch");
    return true;
}

namespace {
// The fused rows pass + key products over a batch: groups of 4 ciphertexts, then the
// rest as a group of 2 and/or 1, so that no row is transformed for a ciphertext that is
// not there.
template <int LOGN, bool R4>
void launch_keyprod_groups(const uint64_t *T, const uint64_t *key, uint64_t *P, uint64_t *S, const uint64_t *diag,
                           size_t diag_stride, uint32_t cts, uint32_t L, uint32_t K_key, const DevPrime *targets,
                           uint32_t tiles, uint64_t tile_slots, uint32_t group, const KsFold &fold, hipStream_t st) {
    constexpr int K2 = LOGN - LOGN / 2;
    constexpr int E = 4;
    constexpr int RS = 4096 >> K2;
    // groups of `group` ciphertexts (4 by default), then a group of 2, then a single one
    const uint32_t g = group == 1 || group == 2 ? group : 4;
    const uint32_t g4 = g == 4 ? cts / 4 : 0;
    const uint32_t g2 = g == 1 ? 0 : (cts - 4 * g4) / 2;
    const uint32_t singles = cts - 4 * g4 - 2 * g2;
    if (g4)
        ntt_rows_keyprod_kernel<LOGN, K2, E, RS, 4, 1, R4><<<(uint32_t)(tile_slots * g4), RS << (K2 - E), 0, st>>>(
            T, key, P, S, targets, diag, diag_stride, cts, L, K_key, 0, g4, tiles, fold);
    if (g2)
        ntt_rows_keyprod_kernel<LOGN, K2, E, RS, 2, 1, R4><<<(uint32_t)(tile_slots * g2), RS << (K2 - E), 0, st>>>(
            T, key, P, S, targets, diag, diag_stride, cts, L, K_key, g4 * 4, g2, tiles, fold);
    if (singles) {
        // single ciphertexts: two rows per block
        const uint32_t tiles2 = tiles / 2;
        const uint32_t slots2 = (tiles2 + 7) / 8 * 8;
        ntt_rows_keyprod_kernel<LOGN, K2, E, RS, 1, 2, R4><<<slots2 * singles, RS << (K2 - E), 0, st>>>(
            T, key, P, S, targets, diag, diag_stride, cts, L, K_key, cts - singles, singles, tiles2, fold);
    }
}

template <int LOGN>
void launch_modup_keyprod(const uint64_t *d, size_t d_stride, uint64_t *T, const uint64_t *key, uint64_t *P,
                          uint64_t *S, const uint64_t *diag, size_t diag_stride, uint32_t cts, uint32_t L,
                          uint32_t K_key, const DevPrime *targets, const DevPrime *digits, const KernelProbe *probe,
                          hipStream_t st, KsShape shape, const KsFold &fold) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    const uint32_t tc = modup_targets_per_block(cts * L * ((1u << K2) / CS), L, shape.cols_targets);
    const uint32_t col_grid = cts * L * ((L + tc) / tc) * ((1u << K2) / CS);
    if (tc == 1)
        ntt_cols_modup_kernel<LOGN, K1, E, CS, true><<<col_grid, CS << (K1 - E), 0, st>>>(d, d_stride, T, targets, digits,
                                                                                       cts, L, diag != nullptr, tc);
    else
        ntt_cols_modup_kernel<LOGN, K1, E, CS, false><<<col_grid, CS << (K1 - E), 0, st>>>(d, d_stride, T, targets, digits,
                                                                                        cts, L, diag != nullptr, tc);
    const uint32_t tiles = (L + 1) << K1;
    const uint64_t tile_slots = (uint64_t)((tiles + 7) / 8) * 8;
    if (tile_slots * cts >= (uint64_t(1) << 32))
        fail(ERR_INVALID_ARGUMENT, "key-switch batch too large for one launch");
    if (probe) SHC_HIP(hipEventRecord(probe->begin, st));
    // up to 31 digits the transformed words (< 8q) enter the sums unreduced
    if (L > 31)
        launch_keyprod_groups<LOGN, true>(T, key, P, S, diag, diag_stride, cts, L, K_key, targets, tiles, tile_slots,
                                          shape.rows_group, fold, st);
    else
        launch_keyprod_groups<LOGN, false>(T, key, P, S, diag, diag_stride, cts, L, K_key, targets, tiles,
                                           tile_slots, shape.rows_group, fold, st);
    if (probe) SHC_HIP(hipEventRecord(probe->end, st));
}

// Block prefix table of a grouped launch; sizes[g] blocks for group g.
template <class Table>
uint32_t fill_starts(Table &t, const std::vector<uint64_t> &sizes) {
    uint64_t acc = 0;
    t.n = (uint32_t)sizes.size();
    for (size_t g = 0; g < sizes.size(); g++) {
        t.start[g] = (uint32_t)acc;
        acc += sizes[g];
    }
    t.start[sizes.size()] = (uint32_t)acc;
    if (acc >= (uint64_t(1) << 32) / 1024) fail(ERR_INVALID_ARGUMENT, "grouped launch too large");
    return (uint32_t)acc;
}

template <int LOGN, bool INV, bool FULL>
void launch_ntt_groups(const NttGroup *gs, uint32_t n, hipStream_t st) {
    static_assert(LOGN > 12, "grouped transforms run the two-pass schedule");
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    // the first pass reads src where a group has one; the second runs in place
    NttGroups cols{}, rows_mx{}, rows_full{};
    std::vector<uint64_t> csz, msz, fsz;
    for (uint32_t g = 0; g < n; g++) {
        const NttGroup &a = gs[g];
        const bool first_is_cols = !INV;
        NttGroupDev d{a.data, nullptr, a.lg, a.lg, a.primes};
        NttGroupDev d_first{a.data, a.src, a.lg, a.src ? a.slg : a.lg, a.primes};
        cols.g[csz.size()] = first_is_cols ? d_first : d;
        csz.push_back((uint64_t)a.lg.I * a.lg.J * ((1u << K2) / CS));
        const bool mx = rows_mixed<LOGN, K2, RS>(a.lg.J);
        const uint64_t rb = (uint64_t)a.lg.I * rows_blocks_per_prime<LOGN, K2, E, RS>(a.lg.J, mx);
        if (mx) {
            rows_mx.g[msz.size()] = first_is_cols ? d : d_first;
            msz.push_back(rb);
        } else {
            rows_full.g[fsz.size()] = first_is_cols ? d : d_first;
            fsz.push_back(rb);
        }
    }
    const uint32_t cb = fill_starts(cols, csz);
    auto rows = [&](bool last) {
        if (!msz.empty()) {
            const uint32_t b = fill_starts(rows_mx, msz);
            if (last) ntt_rows_grouped_kernel<LOGN, K2, E, RS, INV, true, FULL, true><<<b, RS << (K2 - E), 0, st>>>(rows_mx);
            else ntt_rows_grouped_kernel<LOGN, K2, E, RS, INV, false, FULL, true><<<b, RS << (K2 - E), 0, st>>>(rows_mx);
        }
        if (!fsz.empty()) {
            const uint32_t b = fill_starts(rows_full, fsz);
            if (last) ntt_rows_grouped_kernel<LOGN, K2, E, RS, INV, true, FULL, false><<<b, RS << (K2 - E), 0, st>>>(rows_full);
            else ntt_rows_grouped_kernel<LOGN, K2, E, RS, INV, false, FULL, false><<<b, RS << (K2 - E), 0, st>>>(rows_full);
        }
    };
    if (!INV) {
        ntt_cols_grouped_kernel<LOGN, K1, E, CS, false, false, FULL><<<cb, CS << (K1 - E), 0, st>>>(cols);
        rows(true);
    } else {
        rows(false);
        ntt_cols_grouped_kernel<LOGN, K1, E, CS, true, true, FULL><<<cb, CS << (K1 - E), 0, st>>>(cols);
    }
    launch_check("grouped ntt launch");
}

template <bool INV, bool FULL>
void launch_ntt_multi(const NttGroup *gs, const int *logns, uint32_t n, hipStream_t st) {
    NttMulti cols{}, rows{};
    std::vector<uint64_t> csz, rsz;
    for (uint32_t g = 0; g < n; g++) {
        const NttGroup &a = gs[g];
        const int L = logns[g];
        const int K1 = L / 2, K2 = L - K1;
        const uint32_t CS = 4096u >> K1, RS = 4096u >> K2;
        const NttGroupDev d{a.data, nullptr, a.lg, a.lg, a.primes};
        const NttGroupDev d_first{a.data, a.src, a.lg, a.src ? a.slg : a.lg, a.primes};
        // the first pass reads src where a group has one; the second runs in place
        cols.g[g] = NttMultiDev{INV ? d : d_first, L};
        rows.g[g] = NttMultiDev{INV ? d_first : d, L};
        csz.push_back((uint64_t)a.lg.I * a.lg.J * ((1u << (L - K1)) / CS));
        rsz.push_back((uint64_t)a.lg.I * (1u << (L - K2)) * ((a.lg.J + RS - 1) / RS));
    }
    const uint32_t cb = fill_starts(cols, csz), rb = fill_starts(rows, rsz);
    if (!INV) {
        ntt_cols_multi_kernel<false, false, FULL><<<cb, 256, 0, st>>>(cols);
        ntt_rows_multi_kernel<false, true, FULL><<<rb, 256, 0, st>>>(rows);
    } else {
        ntt_rows_multi_kernel<true, false, FULL><<<rb, 256, 0, st>>>(rows);
        ntt_cols_multi_kernel<true, true, FULL><<<cb, 256, 0, st>>>(cols);
    }
    launch_check("multi-degree ntt launch");
}

template <int LOGN>
void launch_modup_keyprod_groups(const KsGroup *gs, uint32_t n, const uint64_t *key, uint32_t K_key,
                                 const DevPrime *digits, hipStream_t st) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    // columns pass: targets per block chosen for the whole launch's tile blocks
    uint64_t tile_blocks = 0;
    uint32_t maxL = 0;
    for (uint32_t g = 0; g < n; g++) {
        tile_blocks += (uint64_t)gs[g].count * gs[g].L * ((1u << K2) / CS);
        maxL = gs[g].L > maxL ? gs[g].L : maxL;
    }
    const uint64_t want = (16384 + tile_blocks - 1) / tile_blocks;
    ModupGroups mg{};
    mg.digits = digits;
    std::vector<uint64_t> msz;
    bool all_one = true;
    for (uint32_t g = 0; g < n; g++) {
        const KsGroup &a = gs[g];
        const uint32_t chunks = (uint32_t)(want < 1 ? 1 : (want > a.L + 1 ? a.L + 1 : want));
        const uint32_t tc = (a.L + chunks) / chunks;
        all_one = all_one && tc == 1;
        mg.g[g] = ModupGroup{a.d, a.d_ct_stride, a.T, a.targets, a.count, a.L, tc, a.diag != nullptr};
        msz.push_back((uint64_t)a.count * a.L * ((a.L + tc) / tc) * ((1u << K2) / CS));
    }
    const uint32_t cb = fill_starts(mg, msz);
    if (all_one) ntt_cols_modup_grouped_kernel<LOGN, K1, E, CS, true><<<cb, CS << (K1 - E), 0, st>>>(mg);
    else ntt_cols_modup_grouped_kernel<LOGN, K1, E, CS, false><<<cb, CS << (K1 - E), 0, st>>>(mg);
    // fused rows pass + key products: groups of 4, then of 2, then single ciphertexts
    KeyprodGroups t4{}, t2{}, t1{};
    t4.key = t2.key = t1.key = key;
    t4.K_key = t2.K_key = t1.K_key = K_key;
    std::vector<uint64_t> s4, s2, s1;
    for (uint32_t g = 0; g < n; g++) {
        const KsGroup &a = gs[g];
        const uint32_t tiles = (a.L + 1) << K1;
        const uint64_t tile_slots = (uint64_t)((tiles + 7) / 8) * 8;
        const uint32_t g4 = a.count / 4, g2 = (a.count - 4 * g4) / 2, singles = a.count - 4 * g4 - 2 * g2;
        if (g4) {
            t4.g[s4.size()] =
                KeyprodGroup{a.T, a.P, a.S, a.targets, a.diag, a.diag_stride, a.count, a.L, 0, g4, tiles, a.fold};
            s4.push_back(tile_slots * g4);
        }
        if (g2) {
            t2.g[s2.size()] =
                KeyprodGroup{a.T, a.P, a.S, a.targets, a.diag, a.diag_stride, a.count, a.L, 4 * g4, g2, tiles, a.fold};
            s2.push_back(tile_slots * g2);
        }
        if (singles) {
            const uint32_t tiles2 = tiles / 2;
            const uint64_t slots2 = (tiles2 + 7) / 8 * 8;
            t1.g[s1.size()] = KeyprodGroup{a.T, a.P, a.S, a.targets, a.diag, a.diag_stride, a.count, a.L,
                                           a.count - singles, singles, tiles2, a.fold};
            s1.push_back(slots2 * singles);
        }
    }
    const bool r4 = maxL > 31;
    if (!s4.empty()) {
        const uint32_t b = fill_starts(t4, s4);
        if (r4) ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 4, 1, true><<<b, RS << (K2 - E), 0, st>>>(t4);
        else ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 4, 1, false><<<b, RS << (K2 - E), 0, st>>>(t4);
    }
    if (!s2.empty()) {
        const uint32_t b = fill_starts(t2, s2);
        if (r4) ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 2, 1, true><<<b, RS << (K2 - E), 0, st>>>(t2);
        else ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 2, 1, false><<<b, RS << (K2 - E), 0, st>>>(t2);
    }
    if (!s1.empty()) {
        const uint32_t b = fill_starts(t1, s1);
        if (r4) ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 1, 2, true><<<b, RS << (K2 - E), 0, st>>>(t1);
        else ntt_rows_keyprod_grouped_kernel<LOGN, K2, E, RS, 1, 2, false><<<b, RS << (K2 - E), 0, st>>>(t1);
    }
}

template <int LOGN>
void launch_moddown_groups(const ModDownGroup *gs, uint32_t n, const DevPrime *primes, hipStream_t st) {
    constexpr int K1 = LOGN / 2;
    constexpr int K2 = LOGN - K1;
    constexpr int E = 4;
    constexpr int CS = 16;
    constexpr int RS = 4096 >> K2;
    ModDownGroups cols{}, rmx{}, rfull{};
    std::vector<uint64_t> csz, msz, fsz;
    for (uint32_t g = 0; g < n; g++) {
        const uint32_t jn = 2 * gs[g].count;
        const ModDownGroupDev d{gs[g].U, gs[g].a, jn};
        cols.g[g] = d;
        csz.push_back((uint64_t)gs[g].a.L * jn * ((1u << K2) / CS));
        const bool mx = rows_mixed<LOGN, K2, RS>(jn);
        const uint64_t rb = (uint64_t)gs[g].a.L * rows_blocks_per_prime<LOGN, K2, E, RS>(jn, mx);
        if (mx) {
            rmx.g[msz.size()] = d;
            msz.push_back(rb);
        } else {
            rfull.g[fsz.size()] = d;
            fsz.push_back(rb);
        }
    }
    const uint32_t cb = fill_starts(cols, csz);
    ntt_cols_moddown_prep_grouped_kernel<LOGN, K1, E, CS><<<cb, CS << (K1 - E), 0, st>>>(primes, cols);
    if (!msz.empty()) {
        const uint32_t b = fill_starts(rmx, msz);
        ntt_rows_moddown_grouped_kernel<LOGN, K2, E, RS, true><<<b, RS << (K2 - E), 0, st>>>(primes, rmx);
    }
    if (!fsz.empty()) {
        const uint32_t b = fill_starts(rfull, fsz);
        ntt_rows_moddown_grouped_kernel<LOGN, K2, E, RS, false><<<b, RS << (K2 - E), 0, st>>>(primes, rfull);
    }
}

LimbGrid poly_grid(uint32_t count, uint32_t lpp) {
    if (lpp == 0 || count % lpp) fail(ERR_INVALID_ARGUMENT, "limb count must be a multiple of limbs per polynomial");
    return LimbGrid{1, lpp, lpp, count / lpp};
}
}  // namespace

bool ntt_modup_keyprod(const uint64_t *d, size_t d_ct_stride, uint64_t *T, const uint64_t *key, uint64_t *P,
                       uint64_t *S, const uint64_t *diag, size_t diag_stride, uint32_t count, uint32_t L,
                       uint32_t key_count, const DevPrime *targets, const DevPrime *digits, int logn,
                       hipStream_t st, const KernelProbe *probe, KsShape shape, KsFold fold) {
    if ((logn != 15 && logn != 16) || count == 0) return false;
    if ((uint64_t)count * L * (L + 1) > ((uint64_t(1) << 34) >> logn))
        fail(ERR_INVALID_ARGUMENT, "key-switch batch too large for one launch");
    if (logn == 15)
        launch_modup_keyprod<15>(d, d_ct_stride, T, key, P, S, diag, diag_stride, count, L, key_count, targets,
                                 digits, probe, st, shape, fold);
    else
        launch_modup_keyprod<16>(d, d_ct_stride, T, key, P, S, diag, diag_stride, count, L, key_count, targets,
                                 digits, probe, st, shape, fold);
    launch_check("ntt modup keyprod launch");
    return true;
}

namespace {
bool moddown_jn(bool ckks, uint64_t *buf, const ModDownArgs &a, uint32_t jn, const DevPrime *primes, int logn,
                hipStream_t st) {
    if (logn < 13 || logn > 17 || jn == 0) return false;
    if ((uint64_t)jn * a.L > ((uint64_t(1) << 34) >> logn)) fail(ERR_INVALID_ARGUMENT, "batch too large for one launch");
    if (a.p_lpp && !ckks) fail(ERR_LOGIC, "a strided ModDown input is for the NTT form");
    switch (logn) {
        case 13: launch_moddown<13>(ckks, buf, a, jn, primes, st); break;
        case 14: launch_moddown<14>(ckks, buf, a, jn, primes, st); break;
        case 15: launch_moddown<15>(ckks, buf, a, jn, primes, st); break;
        case 16: launch_moddown<16>(ckks, buf, a, jn, primes, st); break;
        case 17: launch_moddown<17>(ckks, buf, a, jn, primes, st); break;
    }
    launch_check("ntt moddown launch");
    return true;
}
}  // namespace

bool ntt_moddown(bool ckks, uint64_t *buf, const ModDownArgs &a, uint32_t count, const DevPrime *primes, int logn,
                 hipStream_t st) {
    return moddown_jn(ckks, buf, a, 2 * count, primes, logn, st);
}

bool ntt_encrypt_ckks(const EncArgs &a, uint32_t m, const DevPrime *primes, int logn, hipStream_t st) {
    if (logn < 13 || logn > 17 || m == 0) return false;
    const uint32_t jn = 2 * m;
    if ((uint64_t)jn * a.L > ((uint64_t(1) << 34) >> logn)) return false;
    switch (logn) {
        case 13: launch_encrypt<13>(a, jn, primes, st); break;
        case 14: launch_encrypt<14>(a, jn, primes, st); break;
        case 15: launch_encrypt<15>(a, jn, primes, st); break;
        case 16: launch_encrypt<16>(a, jn, primes, st); break;
        case 17: launch_encrypt<17>(a, jn, primes, st); break;
    }
    launch_check("ntt encrypt launch");
    return true;
}

bool ntt_moddown_polys(uint64_t *U, const ModDownArgs &a, uint32_t polys, const DevPrime *primes, int logn,
                       hipStream_t st) {
    return moddown_jn(true, U, a, polys, primes, logn, st);
}

void ntt_forward(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn, NttOut out,
                 hipStream_t st) {
    if (count == 0) return;
    ntt_forward_grid(data, poly_grid(count, lpp), primes, logn, out, st);
}

void ntt_forward_grid(uint64_t *data, LimbGrid lg, const DevPrime *primes, int logn, NttOut out, hipStream_t st) {
    if (out == NTT_FULL) dispatch<false, true>(data, lg, primes, logn, st);
    else dispatch<false, false>(data, lg, primes, logn, st);
}

void ntt_inverse(uint64_t *data, uint32_t count, uint32_t lpp, const DevPrime *primes, int logn, NttOut out,
                 hipStream_t st) {
    if (count == 0) return;
    ntt_inverse_grid(data, poly_grid(count, lpp), primes, logn, out, st);
}

void ntt_inverse_grid(uint64_t *data, LimbGrid lg, const DevPrime *primes, int logn, NttOut out, hipStream_t st) {
    if (out == NTT_FULL) dispatch<true, true>(data, lg, primes, logn, st);
    else dispatch<true, false>(data, lg, primes, logn, st);
}

void ntt_inverse_from(const uint64_t *src, LimbGrid src_grid, uint64_t *data, LimbGrid grid, const DevPrime *primes,
                      int logn, NttOut out, hipStream_t st) {
    if (out == NTT_FULL) dispatch_inverse_from<true>(src, src_grid, data, grid, primes, logn, st);
    else dispatch_inverse_from<false>(src, src_grid, data, grid, primes, logn, st);
}

void ntt_forward_groups(const NttGroup *gs, uint32_t n, int logn, NttOut out, hipStream_t st) {
    if (n > MAX_GROUPS) fail(ERR_INVALID_ARGUMENT, "too many groups for one launch");
    if (n == 0) return;
    if (logn <= 12) {
        // one-tile transforms: a group with a source is copied first, then transformed
        for (uint32_t g = 0; g < n; g++) {
            if (gs[g].src) {
                const LimbGrid &d = gs[g].lg, &s = gs[g].slg;
                if (d.A == 1 && s.A == 1 && d.B == d.I && s.B == s.I) {  // both dense: one copy
                    SHC_HIP(hipMemcpyAsync(gs[g].data, gs[g].src, ((size_t)d.I * d.J * 8) << logn,
                                           hipMemcpyDeviceToDevice, st));
                    ntt_forward_grid(gs[g].data, gs[g].lg, gs[g].primes, logn, out, st);
                    continue;
                }
                for (uint32_t j = 0; j < d.J; j++)
                    for (uint32_t i = 0; i < d.I; i++)
                        SHC_HIP(hipMemcpyAsync(gs[g].data + ((size_t)(i * d.A + j * d.B) << logn),
                                               gs[g].src + ((size_t)(i * s.A + j * s.B) << logn), (size_t)8 << logn,
                                               hipMemcpyDeviceToDevice, st));
            }
            ntt_forward_grid(gs[g].data, gs[g].lg, gs[g].primes, logn, out, st);
        }
        return;
    }
    switch (logn) {
#define SHC_NTT_CASE(L)                                                          \
    case L:                                                                      \
        if (out == NTT_FULL) launch_ntt_groups<L, false, true>(gs, n, st);       \
        else launch_ntt_groups<L, false, false>(gs, n, st);                      \
        break;
        SHC_NTT_CASE(13) SHC_NTT_CASE(14) SHC_NTT_CASE(15) SHC_NTT_CASE(16) SHC_NTT_CASE(17)
#undef SHC_NTT_CASE
        default: fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    }
}

void ntt_inverse_groups(const NttGroup *gs, uint32_t n, int logn, NttOut out, hipStream_t st) {
    if (n > MAX_GROUPS) fail(ERR_INVALID_ARGUMENT, "too many groups for one launch");
    if (n == 0) return;
    if (logn <= 12) {
        for (uint32_t g = 0; g < n; g++) {
            if (gs[g].src) ntt_inverse_from(gs[g].src, gs[g].slg, gs[g].data, gs[g].lg, gs[g].primes, logn, out, st);
            else ntt_inverse_grid(gs[g].data, gs[g].lg, gs[g].primes, logn, out, st);
        }
        return;
    }
    switch (logn) {
#define SHC_NTT_CASE(L)                                                          \
    case L:                                                                      \
        if (out == NTT_FULL) launch_ntt_groups<L, true, true>(gs, n, st);        \
        else launch_ntt_groups<L, true, false>(gs, n, st);                       \
        break;
        SHC_NTT_CASE(13) SHC_NTT_CASE(14) SHC_NTT_CASE(15) SHC_NTT_CASE(16) SHC_NTT_CASE(17)
#undef SHC_NTT_CASE
        default: fail(ERR_INVALID_ARGUMENT, "poly_modulus_degree must be 2^1 .. 2^17");
    }
}

namespace {
void check_multi(const NttGroup *gs, const int *logns, uint32_t n) {
    if (n > MAX_GROUPS) fail(ERR_INVALID_ARGUMENT, "too many groups for one launch");
    uint64_t cols = 0, rows = 0;
    for (uint32_t g = 0; g < n; g++) {
        const int L = logns[g];
        if (L < 13 || L > 17) fail(ERR_INVALID_ARGUMENT, "multi-degree transforms take N = 2^13 .. 2^17");
        const LimbGrid &lg = gs[g].lg;
        if (!lg.I || !lg.J) fail(ERR_INVALID_ARGUMENT, "empty transform group");
        const int K1 = L / 2, K2 = L - K1;
        cols += (uint64_t)lg.I * lg.J * ((1u << (L - K1)) / (4096u >> K1));
        rows += (uint64_t)lg.I * (1u << (L - K2)) * ((lg.J + (4096u >> K2) - 1) / (4096u >> K2));
    }
    if (cols >= (uint64_t(1) << 32) || rows >= (uint64_t(1) << 32))
        fail(ERR_INVALID_ARGUMENT, "transform batch too large for one launch");
}
}  // namespace

void ntt_forward_multi(const NttGroup *gs, const int *logns, uint32_t n, NttOut out, hipStream_t st) {
    check_multi(gs, logns, n);
    if (!n) return;
    if (out == NTT_FULL) launch_ntt_multi<false, true>(gs, logns, n, st);
    else launch_ntt_multi<false, false>(gs, logns, n, st);
}

void ntt_inverse_multi(const NttGroup *gs, const int *logns, uint32_t n, NttOut out, hipStream_t st) {
    check_multi(gs, logns, n);
    if (!n) return;
    if (out == NTT_FULL) launch_ntt_multi<true, true>(gs, logns, n, st);
    else launch_ntt_multi<true, false>(gs, logns, n, st);
}

bool ntt_modup_keyprod_groups(const KsGroup *gs, uint32_t n, const uint64_t *key, uint32_t key_count,
                              const DevPrime *digits, int logn, hipStream_t st) {
    if (logn != 15 && logn != 16) return false;
    if (n > MAX_GROUPS) fail(ERR_INVALID_ARGUMENT, "too many groups for one launch");
    if (n == 0) return true;
    if (logn == 15) launch_modup_keyprod_groups<15>(gs, n, key, key_count, digits, st);
    else launch_modup_keyprod_groups<16>(gs, n, key, key_count, digits, st);
    launch_check("grouped modup keyprod launch");
    return true;
}

bool ntt_moddown_groups(const ModDownGroup *gs, uint32_t n, const DevPrime *primes, int logn, hipStream_t st) {
    if (logn < 13 || logn > 17) return false;
    if (n > MAX_GROUPS) fail(ERR_INVALID_ARGUMENT, "too many groups for one launch");
    if (n == 0) return true;
    switch (logn) {
        case 13: launch_moddown_groups<13>(gs, n, primes, st); break;
        case 14: launch_moddown_groups<14>(gs, n, primes, st); break;
        case 15: launch_moddown_groups<15>(gs, n, primes, st); break;
        case 16: launch_moddown_groups<16>(gs, n, primes, st); break;
        case 17: launch_moddown_groups<17>(gs, n, primes, st); break;
    }
    launch_check("grouped moddown launch");
    return true;
}

}  // namespace shc
