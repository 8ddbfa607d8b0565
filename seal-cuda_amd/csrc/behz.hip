// behz.hip -- BFV multiplication by the BEHZ method on gfx950.
//
// Same arithmetic as Evaluator::bfvMultiply (src/evaluator.cpp:385-624) with the
// RNSTool steps (src/utils/rns.cpp:849-960); GPU twins src/evaluator_cuda.cu:283-382,
// src/utils/rns_cuda.cu:300-443.  Per coefficient, each base conversion is a short
// dot product over the input base.  A block takes 64 coefficients, one per lane, and its
// two waves split the work: each stages half of the scaled inputs in LDS, and then
// computes its own half of the output primes, OB at a time (each staged word read once
// per group), with three-word product sums (accum.h; inputs < 2^61, at most 63 terms:
// 62 inputs and the correction product that smMrq / Shenoy-Kumaresan add,
// so every sum stays below 2^128), so intermediate residues never leave the kernel.
// Sharing the staged words between two waves halves the LDS per wave, which is what
// bounds the waves a CU holds.  Scalar multiplies that precede a conversion in the
// reference (x m_tilde, x t) are folded into the conversion's first constant.
#include "behz.h"
#include "accum.h"
#include "context.h"
#include "modarith.h"
#include "ntt.h"
#include "rns_ops.h"

#include <type_traits>

namespace shc {
namespace {

constexpr int CW = 64;       // coefficients per block of the conversion kernels, one per lane
constexpr int WV = 2;        // waves per block: each takes a slice of the output primes
constexpr int BT = CW * WV;  // threads per block
constexpr int MAXB = 64;     // LDS rows per block (>= |q| and >= |B|)

inline uint32_t grid_for(size_t coefficients) {
    const size_t b = (coefficients + CW - 1) / CW;
    return (uint32_t)(b < (size_t(1) << 20) ? b : (size_t(1) << 20));
}

// Output primes per pass for a wave's slice of `outputs`: the fewest sums computed, then
// the fewest passes over the staged inputs.
inline int out_group(uint32_t outputs) {
    const uint32_t S = (outputs + WV - 1) / WV;
    int best = 4;
    uint32_t best_slots = ~0u, best_passes = ~0u;
    for (int ob : {7, 6, 5, 4}) {
        const uint32_t passes = (S + ob - 1) / ob, slots = passes * ob;
        if (slots < best_slots || (slots == best_slots && passes < best_passes)) {
            best = ob;
            best_slots = slots;
            best_passes = passes;
        }
    }
    return best;
}

// Wave w's outputs [ob, oe) out of `outputs` (contiguous slices of ceil(outputs / WV)).
__device__ __forceinline__ void slice_of(uint32_t outputs, int w, uint32_t &ob, uint32_t &oe) {
    const uint32_t S = (outputs + WV - 1) / WV;
    ob = min<uint32_t>(outputs, (uint32_t)w * S);
    oe = min<uint32_t>(outputs, ob + S);
}

__device__ __forceinline__ uint64_t shoup_pair(uint64_t x, const uint64_t *pair, uint64_t q) {
    return shoup(x, pair[0], pair[1], q);
}

// acc[u] = sum_j v[j * CW + lane] * rows[(o0 + u) * terms + j] for u < OB (rows past
// `outputs` repeat the last one; the caller does not use those sums).  The rows are
// read-only tables with wave-uniform addresses: read through the constant address
// space, so that they become scalar loads.
typedef const __attribute__((address_space(4))) uint64_t const_u64;
template <int OB>
__device__ __forceinline__ void dot_rows(Acc3 (&acc)[OB], const uint64_t *v, const uint64_t *__restrict__ rows,
                                         uint32_t terms, uint32_t o0, uint32_t outputs, int lane) {
    const_u64 *r[OB];
#pragma unroll
    for (int u = 0; u < OB; u++) {
        r[u] = (const_u64 *)(rows + (size_t)min<uint32_t>(o0 + u, outputs - 1) * terms);
        acc[u] = Acc3{0, 0, 0};
    }
    // four terms at a time: their LDS words and table entries are requested together, so
    // one wait covers four terms
    uint32_t j = 0;
    for (; j + 4 <= terms; j += 4) {
        uint64_t y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) y[k] = v[(j + k) * CW + lane];
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int u = 0; u < OB; u++) mac<true>(acc[u], y[k], r[u][j + k]);
    }
    for (; j < terms; j++) {
        const uint64_t y = v[j * CW + lane];
#pragma unroll
        for (int u = 0; u < OB; u++) mac<true>(acc[u], y, r[u][j]);
    }
}

// (1)-(2): fastbconvmTilde + smMrq.  x: [P][L][N] base q -> out: [P][bsk][N] base Bsk.
template <int OB>
__global__ __launch_bounds__(BT) void behz_extend_kernel(const uint64_t *__restrict__ x, uint64_t *__restrict__ out,
                                                         BehzDev B, const DevPrime *__restrict__ qp, int logn,
                                                         size_t total) {
    extern __shared__ uint64_t ys[];  // [L][CW]
    const size_t n = (size_t)1 << logn;
    // the wave index, uniform across the wave (so that the slice and its table rows are scalar)
    const int lane = threadIdx.x % CW, w = __builtin_amdgcn_readfirstlane(threadIdx.x / CW);
    uint32_t ob, oe;
    slice_of(B.bsk, w, ob, oe);
    const uint64_t *row_mt = B.qhat_ext + (size_t)B.bsk * B.L;
    for (size_t base = (size_t)blockIdx.x * CW; base < total; base += (size_t)gridDim.x * CW) {
        const size_t idx = base + lane;
        const bool live = idx < total;
        const size_t p = idx >> logn, c = idx & (n - 1);
        for (uint32_t j = w; j < B.L; j += WV)
            ys[j * CW + lane] = live ? shoup_pair(x[(p * B.L + j) * n + c], B.mt_qhat_inv + 2 * j, qp[j].q) : 0;
        __syncthreads();
        // conversion into m_tilde = 2^32 (each wave): sums wrap modulo 2^64, a multiple of 2^32
        uint64_t s_mt = 0;
        for (uint32_t j = 0; j < B.L; j++) s_mt += ys[j * CW + lane] * row_mt[j];
        const uint64_t r = ((s_mt & 0xffffffffull) * B.m_tilde_neg_inv_q) & 0xffffffffull;
        for (uint32_t i0 = ob; i0 < oe; i0 += OB) {
            // the rows carry m_tilde^-1: the sums are (s + r q) m_tilde^-1 mod p directly
            Acc3 acc[OB];
            dot_rows(acc, ys, B.qhat_ext, B.L, i0, B.bsk, lane);
#pragma unroll
            for (int u = 0; u < OB; u++) {
                const uint32_t i = i0 + u;
                if (i >= oe) break;
                const DevPrime &P = B.bsk_primes[i];
                // the centred r (m_tilde is a power of two, so r >= m_tilde/2 is negative)
                // times q m_tilde^-1 joins the sum (one more term below 2^122), so one
                // reduction gives the output residue
                const uint64_t rc = r >= 0x80000000ull ? r + P.q - 0x100000000ull : r;
                mac(acc[u], rc, B.q_ext[i]);
                if (live) out[(p * B.bsk + i) * n + c] = acc_reduce(acc[u], P);
            }
        }
        __syncthreads();  // the staged words are read before the next coefficients' overwrite them
    }
}

// (6)-(7): multiply by t and fastFloor.  Tq: [P][L][N], Tb: [P][bsk][N] -> Z: [P][bsk][N].
template <int OB>
__global__ __launch_bounds__(BT) void behz_floor_kernel(const uint64_t *__restrict__ Tq, const uint64_t *__restrict__ Tb,
                                                        uint64_t *__restrict__ Z, BehzDev B,
                                                        const DevPrime *__restrict__ qp, int logn, size_t total) {
    extern __shared__ uint64_t ws[];  // [L][CW]
    const size_t n = (size_t)1 << logn;
    // the wave index, uniform across the wave (so that the slice and its table rows are scalar)
    const int lane = threadIdx.x % CW, w = __builtin_amdgcn_readfirstlane(threadIdx.x / CW);
    uint32_t ob, oe;
    slice_of(B.bsk, w, ob, oe);
    for (size_t base = (size_t)blockIdx.x * CW; base < total; base += (size_t)gridDim.x * CW) {
        const size_t idx = base + lane;
        const bool live = idx < total;
        const size_t p = idx >> logn, c = idx & (n - 1);
        for (uint32_t j = w; j < B.L; j += WV)
            ws[j * CW + lane] = live ? shoup_pair(Tq[(p * B.L + j) * n + c], B.t_qhat_inv + 2 * j, qp[j].q) : 0;
        __syncthreads();
        for (uint32_t i0 = ob; i0 < oe; i0 += OB) {
            // the rows carry prod(q)^-1: Z = t Tb q^-1 - conv q^-1 mod p
            Acc3 acc[OB];
            dot_rows(acc, ws, B.qhat_floor, B.L, i0, B.bsk, lane);
#pragma unroll
            for (int u = 0; u < OB; u++) {
                const uint32_t i = i0 + u;
                if (i >= oe) break;
                if (!live) continue;
                const DevPrime &P = B.bsk_primes[i];
                const uint64_t conv = acc_reduce(acc[u], P);
                const uint64_t tb = shoup_pair(Tb[(p * B.bsk + i) * n + c], B.t_inv_q_bsk + 2 * i, P.q);
                Z[(p * B.bsk + i) * n + c] = submod(tb, conv, P.q);
            }
        }
        __syncthreads();
    }
}

// (8): fastbconvSk.  Z: [P][bsk][N] -> out: [P][L][N] base q.
template <int OB>
__global__ __launch_bounds__(BT) void behz_sk_kernel(const uint64_t *__restrict__ Z, uint64_t *__restrict__ out,
                                                     BehzDev B, const DevPrime *__restrict__ qp, int logn,
                                                     size_t total) {
    extern __shared__ uint64_t zs[];  // [bsz][CW]
    const size_t n = (size_t)1 << logn;
    // the wave index, uniform across the wave (so that the slice and its table rows are scalar)
    const int lane = threadIdx.x % CW, w = __builtin_amdgcn_readfirstlane(threadIdx.x / CW);
    uint32_t ob, oe;
    slice_of(B.L, w, ob, oe);
    const DevPrime &Msk = B.bsk_primes[B.bsz];
    const_u64 *bm = (const_u64 *)B.bhat_mod_msk;
    for (size_t base = (size_t)blockIdx.x * CW; base < total; base += (size_t)gridDim.x * CW) {
        const size_t idx = base + lane;
        const bool live = idx < total;
        const size_t p = idx >> logn, c = idx & (n - 1);
        for (uint32_t j = w; j < B.bsz; j += WV)
            zs[j * CW + lane] =
                live ? shoup_pair(Z[(p * B.bsk + j) * n + c], B.bhat_inv + 2 * j, B.bsk_primes[j].q) : 0;
        __syncthreads();
        // alpha_sk (each wave)
        Acc3 acc{0, 0, 0};
        for (uint32_t j = 0; j < B.bsz; j++) mac<true>(acc, zs[j * CW + lane], bm[j]);
        const uint64_t conv = acc_reduce(acc, Msk);
        const uint64_t zmsk = live ? Z[(p * B.bsk + B.bsz) * n + c] : 0;
        const uint64_t alpha = shoup(conv + B.m_sk - zmsk, B.inv_prod_B_msk, B.inv_prod_B_msk_shoup, B.m_sk);
        const bool neg = alpha > B.m_sk_half;
        for (uint32_t i0 = ob; i0 < oe; i0 += OB) {
            Acc3 s[OB];
            dot_rows(s, zs, B.bhat_mod_q, B.bsz, i0, B.L, lane);
#pragma unroll
            for (int u = 0; u < OB; u++) {
                const uint32_t i = i0 + u;
                if (i >= oe) break;
                if (!live) continue;
                const DevPrime &Q = qp[i];
                // Shenoy-Kumaresan correction with the centred alpha_sk: one more term of
                // the sum (below 2^122), so one reduction gives the output residue
                const uint64_t pb = B.prod_B_mod_q[i];
                mac(s[u], neg ? B.m_sk - alpha : alpha, neg ? pb : Q.q - pb);
                out[(p * B.L + i) * n + c] = acc_reduce(s[u], Q);
            }
        }
        __syncthreads();
    }
}

// Runs f(std::integral_constant<int, OB>) for the OB out_group picks.
template <class F>
void with_group(int ob, F &&f) {
    switch (ob) {
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    default: f(std::integral_constant<int, 4>{}); break;
    }
}

struct BfvLayout {
    size_t aq, bq, ab, bb, tq, tb, z, total;
};

BfvLayout bfv_layout(const Context &c, uint32_t count, uint32_t L, uint32_t bsk) {
    const size_t limb = c.n * 8;
    BfvLayout l;
    size_t off = 0;
    auto take = [&](size_t limbs) { size_t o = off; off += (limbs * limb + 255) & ~size_t(255); return o; };
    l.aq = take((size_t)count * 2 * L);
    l.bq = take((size_t)count * 2 * L);
    l.ab = take((size_t)count * 2 * bsk);
    l.bb = take((size_t)count * 2 * bsk);
    l.tq = take((size_t)count * 3 * L);
    l.tb = take((size_t)count * 3 * bsk);
    l.z = take((size_t)count * 3 * bsk);
    l.total = off;
    return l;
}

}  // namespace

void build_behz_level(Context &c, uint32_t L, BehzLevel &out) {
    using namespace hm;
    const std::vector<uint64_t> q(c.moduli.begin(), c.moduli.begin() + L);
    const uint64_t t = c.plain;
    if (L > MAXB - 2) fail(ERR_INVALID_ARGUMENT, "too many primes for BEHZ tables");
    // RNSTool::initialize (src/utils/rns.cpp:587-646)
    size_t bsz = L;
    if (32 + bit_count(t) + big_bit_count(big_product(q)) >= 61 * (int)L + 61) bsz++;
    const std::vector<uint64_t> pr = get_primes(2 * c.n, 61, bsz + 2);
    const uint64_t msk = pr[0];
    out.gamma = pr[1];
    std::vector<uint64_t> Bv(pr.begin() + 2, pr.end());
    std::vector<uint64_t> Bsk = Bv;
    Bsk.push_back(msk);
    out.bsk_moduli = Bsk;
    const uint64_t mt = uint64_t(1) << 32;
    const uint32_t bsk = (uint32_t)Bsk.size();

    auto prod_mod = [](const std::vector<uint64_t> &base, size_t skip, uint64_t p) {
        uint64_t r = 1 % p;
        for (size_t k = 0; k < base.size(); k++)
            if (k != skip) r = mulmod(r, base[k] % p, p);
        return r;
    };
    const size_t NONE = (size_t)-1;
    std::vector<uint64_t> words;
    auto put = [&](std::vector<uint64_t> v) { size_t o = words.size(); words.insert(words.end(), v.begin(), v.end()); return o; };
    auto shoup_vec = [](const std::vector<uint64_t> &w, const std::vector<uint64_t> &mods) {
        std::vector<uint64_t> r;
        for (size_t i = 0; i < w.size(); i++) { r.push_back(w[i]); r.push_back(shoup(w[i], mods[i])); }
        return r;
    };
    std::vector<uint64_t> qhat_inv(L), v;
    for (uint32_t j = 0; j < L; j++) qhat_inv[j] = invmod(prod_mod(q, j, q[j]), q[j]);
    v.clear();
    for (uint32_t j = 0; j < L; j++) v.push_back(mulmod(mt % q[j], qhat_inv[j], q[j]));
    const size_t o_mt_qhat_inv = put(shoup_vec(v, q));
    // the q -> Bsk conversion rows (qhat_j mod b) with the step that follows each
    // conversion folded in: smMrq's m_tilde^-1 (extend, whose last row is qhat_j mod
    // m_tilde) and fastFloor's prod(q)^-1 (floor)
    std::vector<uint64_t> inv_mt(bsk), inv_q(bsk);
    for (uint32_t i = 0; i < bsk; i++) {
        inv_mt[i] = invmod(mt % Bsk[i], Bsk[i]);
        inv_q[i] = invmod(prod_mod(q, NONE, Bsk[i]), Bsk[i]);
    }
    v.clear();
    for (uint32_t i = 0; i < bsk; i++)
        for (uint32_t j = 0; j < L; j++) v.push_back(mulmod(prod_mod(q, j, Bsk[i]), inv_mt[i], Bsk[i]));
    for (uint32_t j = 0; j < L; j++) v.push_back(prod_mod(q, j, mt));
    const size_t o_qhat_ext = put(v);
    v.clear();
    for (uint32_t i = 0; i < bsk; i++) v.push_back(mulmod(prod_mod(q, NONE, Bsk[i]), inv_mt[i], Bsk[i]));
    const size_t o_q_ext = put(v);
    v.clear();
    for (uint32_t i = 0; i < bsk; i++)
        for (uint32_t j = 0; j < L; j++) v.push_back(mulmod(prod_mod(q, j, Bsk[i]), inv_q[i], Bsk[i]));
    const size_t o_qhat_floor = put(v);
    v.clear();
    for (uint32_t i = 0; i < bsk; i++) v.push_back(mulmod(t % Bsk[i], inv_q[i], Bsk[i]));
    const size_t o_t_inv_q = put(shoup_vec(v, Bsk));
    v.clear();
    for (uint32_t j = 0; j < L; j++) v.push_back(mulmod(t % q[j], qhat_inv[j], q[j]));
    const size_t o_t_qhat_inv = put(shoup_vec(v, q));
    v.clear();
    for (size_t j = 0; j < bsz; j++) v.push_back(invmod(prod_mod(Bv, j, Bv[j]), Bv[j]));
    const size_t o_bhat_inv = put(shoup_vec(v, Bv));
    v.clear();
    for (uint32_t i = 0; i < L; i++)
        for (size_t j = 0; j < bsz; j++) v.push_back(prod_mod(Bv, j, q[i]));
    const size_t o_bhat_q = put(v);
    v.clear();
    for (size_t j = 0; j < bsz; j++) v.push_back(prod_mod(Bv, j, msk));
    const size_t o_bhat_msk = put(v);
    v.clear();
    for (uint32_t i = 0; i < L; i++) v.push_back(prod_mod(Bv, NONE, q[i]));
    const size_t o_prodB_q = put(v);
    // decrypt: base q -> {t, gamma} (RNSTool::initialize, src/utils/rns.cpp:721-760)
    const uint64_t g = out.gamma;
    v.clear();
    for (uint32_t j = 0; j < L; j++) v.push_back(mulmod(mulmod(t % q[j], g % q[j], q[j]), qhat_inv[j], q[j]));
    const size_t o_tg_qhat_inv = put(shoup_vec(v, q));
    v.clear();
    for (uint32_t j = 0; j < L; j++) v.push_back(prod_mod(q, j, t));
    const size_t o_qhat_t = put(v);
    v.clear();
    for (uint32_t j = 0; j < L; j++) v.push_back(prod_mod(q, j, g));
    const size_t o_qhat_g = put(v);
    // Delta = floor(prod(q) / t) mod q_j and prod(q) mod t (SEALContext::validate,
    // src/context.cpp:264-284)
    uint64_t q_mod_t = 0;
    const Big qd = big_div_word(big_product(q), t, q_mod_t);
    v.clear();
    std::vector<uint64_t> dl;
    for (uint32_t j = 0; j < L; j++) dl.push_back(big_mod_word(qd, q[j]));
    const size_t o_delta = put(shoup_vec(dl, q));

    out.tables = DevBuf(words.size() * 8);
    SHC_HIP(hipMemcpy(out.tables.ptr, words.data(), words.size() * 8, hipMemcpyHostToDevice));
    std::vector<DevPrime> bp;
    for (uint64_t b : Bsk) bp.push_back(c.aux_prime(b));
    out.bsk_primes = DevBuf(bp.size() * sizeof(DevPrime));
    SHC_HIP(hipMemcpy(out.bsk_primes.ptr, bp.data(), out.bsk_primes.bytes, hipMemcpyHostToDevice));

    const uint64_t *base = out.tables.as<const uint64_t>();
    BehzDev &d = out.dev;
    d.L = L;
    d.bsz = (uint32_t)bsz;
    d.bsk = bsk;
    d.m_sk = msk;
    d.m_sk_half = msk >> 1;
    const uint64_t inv_q_mt = invmod(prod_mod(q, NONE, mt), mt);
    d.m_tilde_neg_inv_q = (mt - inv_q_mt) % mt;
    d.inv_prod_B_msk = invmod(prod_mod(Bv, NONE, msk), msk);
    d.inv_prod_B_msk_shoup = shoup(d.inv_prod_B_msk, msk);
    d.t = t;
    d.bsk_primes = out.bsk_primes.as<const DevPrime>();
    d.mt_qhat_inv = base + o_mt_qhat_inv;
    d.qhat_ext = base + o_qhat_ext;
    d.q_ext = base + o_q_ext;
    d.qhat_floor = base + o_qhat_floor;
    d.t_inv_q_bsk = base + o_t_inv_q;
    d.t_qhat_inv = base + o_t_qhat_inv;
    d.bhat_inv = base + o_bhat_inv;
    d.bhat_mod_q = base + o_bhat_q;
    d.bhat_mod_msk = base + o_bhat_msk;
    d.prod_B_mod_q = base + o_prodB_q;
    d.gamma = g;
    uint64_t inv = invmod(prod_mod(q, NONE, t), t);
    d.neg_inv_q_t = inv ? t - inv : 0;
    d.neg_inv_q_t_shoup = shoup(d.neg_inv_q_t, t);
    inv = invmod(prod_mod(q, NONE, g), g);
    d.neg_inv_q_g = inv ? g - inv : 0;
    d.neg_inv_q_g_shoup = shoup(d.neg_inv_q_g, g);
    d.inv_gamma_t = invmod(g % t, t);
    d.inv_gamma_t_shoup = shoup(d.inv_gamma_t, t);
    barrett_ratio(t, d.t_ratio_lo, d.t_ratio_hi);
    barrett_ratio(g, d.g_ratio_lo, d.g_ratio_hi);
    d.tg_qhat_inv = base + o_tg_qhat_inv;
    d.qhat_mod_t = base + o_qhat_t;
    d.qhat_mod_g = base + o_qhat_g;
    d.q_mod_t = q_mod_t;
    d.delta = base + o_delta;
}

size_t bfv_multiply_workspace(const Context &c, uint32_t count, uint32_t L) {
    return bfv_layout(c, count, L, c.behz(L).dev.bsk).total;
}

void bfv_multiply(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count, uint32_t L,
                  void *ws, size_t ws_bytes, hipStream_t st) {
    if (c.scheme != SCHEME_BFV) fail(ERR_LOGIC, "BFV multiplication needs a BFV context");
    if (!count) return;
    const BehzLevel &lv = c.behz(L);
    const BehzDev &B = lv.dev;
    const uint32_t bsk = B.bsk;
    const BfvLayout lay = bfv_layout(c, count, L, bsk);
    if (!ws || ws_bytes < lay.total) fail(ERR_INVALID_ARGUMENT, "workspace too small for BFV multiplication");
    char *w = static_cast<char *>(ws);
    auto at = [&](size_t off) { return reinterpret_cast<uint64_t *>(w + off); };
    uint64_t *Aq = at(lay.aq), *Bq = at(lay.bq), *Ab = at(lay.ab), *Bb = at(lay.bb);
    uint64_t *Tq = at(lay.tq), *Tb = at(lay.tb), *Z = at(lay.z);
    const int logn = c.logn;
    const size_t n = c.n;
    const DevPrime *qp = c.primes();
    const DevPrime *bp = B.bsk_primes;
    const size_t ext = (size_t)count * 2 * n;
    const size_t lds = (size_t)(L > B.bsz ? L : B.bsz) * CW * 8;
    with_group(out_group(B.bsk), [&](auto ob) {
        behz_extend_kernel<decltype(ob)::value><<<grid_for(ext), BT, lds, st>>>(a, Ab, B, qp, logn, ext);
        behz_extend_kernel<decltype(ob)::value><<<grid_for(ext), BT, lds, st>>>(b, Bb, B, qp, logn, ext);
    });
    launch_check("behz_extend");
    // the four forward transforms in one launch per pass, a and b read in place (their
    // transforms land in Aq, Bq); then the two tensor products and the two inverse
    // transforms
    const LimbGrid gq{1, L, L, count * 2}, gb{1, bsk, bsk, count * 2};
    const NttGroup fwd[4] = {{Aq, a, gq, gq, qp}, {Bq, b, gq, gq, qp}, {Ab, nullptr, gb, gb, bp},
                             {Bb, nullptr, gb, gb, bp}};
    if ((uint64_t)count * 2 * (L > bsk ? L : bsk) <= ((uint64_t(1) << 30) >> logn)) {
        ntt_forward_groups(fwd, 4, logn, NTT_FULL, st);
    } else {
        const size_t qwords = (size_t)count * 2 * L * n;
        SHC_HIP(hipMemcpyAsync(Aq, a, qwords * 8, hipMemcpyDeviceToDevice, st));
        SHC_HIP(hipMemcpyAsync(Bq, b, qwords * 8, hipMemcpyDeviceToDevice, st));
        ntt_forward(Aq, count * 2 * L, L, qp, logn, NTT_FULL, st);
        ntt_forward(Bq, count * 2 * L, L, qp, logn, NTT_FULL, st);
        ntt_forward(Ab, count * 2 * bsk, bsk, bp, logn, NTT_FULL, st);
        ntt_forward(Bb, count * 2 * bsk, bsk, bp, logn, NTT_FULL, st);
    }
    tensor_strided(qp, Aq, 2 * L * n, Bq, 2 * L * n, Tq, 3 * L * n, count, L, logn, st);
    tensor_strided(bp, Ab, 2 * bsk * n, Bb, 2 * bsk * n, Tb, 3 * bsk * n, count, bsk, logn, st);
    const LimbGrid iq{1, L, L, count * 3}, ib{1, bsk, bsk, count * 3};
    const NttGroup inv[2] = {{Tq, nullptr, iq, iq, qp}, {Tb, nullptr, ib, ib, bp}};
    if ((uint64_t)count * 3 * (L > bsk ? L : bsk) <= ((uint64_t(1) << 30) >> logn)) {
        ntt_inverse_groups(inv, 2, logn, NTT_FULL, st);
    } else {
        ntt_inverse(Tq, count * 3 * L, L, qp, logn, NTT_FULL, st);
        ntt_inverse(Tb, count * 3 * bsk, bsk, bp, logn, NTT_FULL, st);
    }
    const size_t fl = (size_t)count * 3 * n;
    with_group(out_group(B.bsk), [&](auto ob) {
        behz_floor_kernel<decltype(ob)::value><<<grid_for(fl), BT, lds, st>>>(Tq, Tb, Z, B, qp, logn, fl);
    });
    launch_check("behz_floor");
    with_group(out_group(B.L), [&](auto ob) {
        behz_sk_kernel<decltype(ob)::value><<<grid_for(fl), BT, lds, st>>>(Z, out, B, qp, logn, fl);
    });
    launch_check("behz_sk");
}

}  // namespace shc
