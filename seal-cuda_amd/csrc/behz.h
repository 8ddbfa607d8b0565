// behz.h -- device tables and launchers for BFV multiplication (BEHZ).
//
// Tables mirror RNSTool (src/utils/rns.cpp:581-803) for one level (base q of size L):
// the auxiliary base Bsk = B u {m_sk}, m_tilde = 2^32, and the constants of every
// fast base conversion, folded where a scalar multiply precedes a conversion.
#pragma once
#include "common.h"
#include <vector>

namespace shc {

struct Context;

// Offsets (in 64-bit words) into one device buffer; all Shoup operands are stored as
// (w, floor(w 2^64 / p)) pairs.
struct BehzDev {
    uint32_t L, bsz, bsk;            // |q|, |B|, |Bsk| = |B| + 1
    uint64_t m_sk, m_tilde_neg_inv_q; // m_sk; -prod(q)^-1 mod 2^32
    uint64_t m_sk_half;
    uint64_t inv_prod_B_msk, inv_prod_B_msk_shoup;
    uint64_t t;
    const DevPrime *bsk_primes;      // [bsk]: B..., m_sk (with NTT tables)
    const uint64_t *mt_qhat_inv;     // [L][2]   (m_tilde * qhat_j^-1) mod q_j
    const uint64_t *qhat_ext;        // [bsk+1][L] qhat_j m_tilde^-1 mod b (b in Bsk), then qhat_j mod m_tilde
    const uint64_t *q_ext;           // [bsk]    prod(q) m_tilde^-1 mod b
    const uint64_t *qhat_floor;      // [bsk][L] qhat_j prod(q)^-1 mod b
    const uint64_t *t_inv_q_bsk;     // [bsk][2] t prod(q)^-1 mod b
    const uint64_t *t_qhat_inv;      // [L][2]   (t * qhat_j^-1) mod q_j
    const uint64_t *bhat_inv;        // [bsz][2] Bhat_j^-1 mod B_j
    const uint64_t *bhat_mod_q;      // [L][bsz] Bhat_j mod q_i
    const uint64_t *bhat_mod_msk;    // [bsz]    Bhat_j mod m_sk
    const uint64_t *prod_B_mod_q;    // [L]      prod(B) mod q_i
    // decryptScaleAndRound (src/utils/rns.cpp:1039-1095): base q -> {t, gamma}
    uint64_t gamma;
    uint64_t neg_inv_q_t, neg_inv_q_t_shoup, neg_inv_q_g, neg_inv_q_g_shoup;
    uint64_t inv_gamma_t, inv_gamma_t_shoup;
    uint64_t t_ratio_hi, t_ratio_lo, g_ratio_hi, g_ratio_lo;  // floor(2^128/t), floor(2^128/gamma)
    const uint64_t *tg_qhat_inv;     // [L][2]   (t gamma qhat_j^-1) mod q_j
    const uint64_t *qhat_mod_t;      // [L]
    const uint64_t *qhat_mod_g;      // [L]
    // multiplyAddPlainWithScalingVariant (src/utils/scalingvariant.cpp:48-86)
    uint64_t q_mod_t;                // prod(q) mod t
    const uint64_t *delta;           // [L][2]   floor(prod(q)/t) mod q_j
};

struct BehzLevel {
    BehzDev dev{};
    std::vector<uint64_t> bsk_moduli;  // host copy of Bsk (B..., m_sk)
    uint64_t gamma = 0;
    DevBuf tables;                     // backing store of the BehzDev pointers
    DevBuf bsk_primes;                 // DevPrime[bsk]
};

// Builds and uploads the tables of level L (q = the first L key-level primes).
void build_behz_level(Context &c, uint32_t L, BehzLevel &out);

size_t bfv_multiply_workspace(const Context &c, uint32_t count, uint32_t L);
// a, b: [count][2][L][N] coefficient form; out: [count][3][L][N].
void bfv_multiply(const Context &c, const uint64_t *a, const uint64_t *b, uint64_t *out, uint32_t count, uint32_t L,
                  void *ws, size_t ws_bytes, hipStream_t st);

}  // namespace shc
